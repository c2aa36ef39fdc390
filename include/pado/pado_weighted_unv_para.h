// pado/pado_weighted_unv_para.h -- drop-in for the reference's
// WeightedParaVertexCentricPLL<BATCH_SIZE> (includes/pado_weighted_unv_para.h:
// 50-292) and WeightedVertexCentricPLL<BATCH_SIZE> (includes/
// pado_weighted_unv_unp.h), backed by the MI355X library's weighted build.
//
// Call sites such as (pado.cpp:270-291)
//     WeightedGraph G(file); auto rank = G.make_rank(); G.id_transfer(rank);
//     WeightedParaVertexCentricPLL<512> *p = new WeightedParaVertexCentricPLL<512>(G);
//     p->store_index_to_file(out, rank);
// compile unchanged.  The label sets are the reference's (the canonical pruned
// landmark labels); distances are exact.  Errors print to stderr and
// exit(EXIT_FAILURE), as the reference does.
#ifndef PADO_GPU_DROPIN_WEIGHTED_PARA_H_
#define PADO_GPU_DROPIN_WEIGHTED_PARA_H_

#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "pado/globals.h"
#include "pado/graph.h"
#include "pado/pado_unw_unv_para.h"
#include "pado_gpu.h"

namespace PADO {

template <inti BATCH_SIZE = 512>
class WeightedParaVertexCentricPLL {
    static_assert(BATCH_SIZE >= 1 && BATCH_SIZE <= 512, "BATCH_SIZE must be in [1, 512]");

    pado_gpu_index *idx_ = nullptr;
    idi num_v_ = 0;

    void release()
    {
        if (idx_) pado_gpu_index_free(idx_);
        idx_ = nullptr;
    }

public:
    WeightedParaVertexCentricPLL() = default;
    // construct (WPARA:1571-1706) on a rank-space WeightedGraph.
    WeightedParaVertexCentricPLL(const WeightedGraph &G) { construct(G); }
    WeightedParaVertexCentricPLL(const WeightedParaVertexCentricPLL &) = delete;
    WeightedParaVertexCentricPLL &operator=(const WeightedParaVertexCentricPLL &) = delete;
    ~WeightedParaVertexCentricPLL() { release(); }

    void construct(const WeightedGraph &G)
    {
        release();
        pado_gpu_ctx *ctx = detail::shared_ctx();
        std::vector<uint64_t> off;
        detail::to_u64(G.vertices, G.num_v, G.num_e, off);
        pado_gpu_graph *g = nullptr;
        double t = -WallTimer::get_time_mark();
        detail::die_on(pado_gpu_graph_upload(ctx, G.num_v, G.num_e, off.data(), G.out_edges, G.out_weights, &g),
                       "pado_gpu_graph_upload");
        pado_gpu_build_cfg cfg = {};
        cfg.batch_size = BATCH_SIZE;
        pado_gpu_status s = pado_gpu_build(ctx, g, &cfg, &idx_);
        pado_gpu_graph_free(g);
        detail::die_on(s, "pado_gpu_build");
        t += WallTimer::get_time_mark();
        num_v_ = G.num_v;
        printf("Threads: %u Batch_size: %u\n", NUM_THREADS, BATCH_SIZE);
        printf("Total_labeling_time: %.2f seconds\n", t);
    }

    // order_labels (WPARA:2058-2122): queries then take original ids.
    void order_labels(const std::vector<idi> &rank2id, const std::vector<idi> &rank)
    {
        (void)rank2id;
        detail::die_on(pado_gpu_index_order(detail::shared_ctx(), idx_, rank.data(), (uint32_t)rank.size()),
                       "pado_gpu_index_order");
        pado_gpu_index_info info;
        pado_gpu_index_get_info(idx_, &info);
        printf("Label_size: %lu mean: %f\n", (unsigned long)info.num_labels,
               num_v_ ? (double)info.num_labels / num_v_ : 0.0);
    }

    // query_distance (WPARA:2125-2243): the reference's int16 answer
    // (>= 32765 and unreachable -> WEIGHTILARGE_MAX).
    weightiLarge query_distance(idi a, idi b)
    {
        uint32_t out = 0;
        detail::die_on(pado_gpu_query_pair(detail::shared_ctx(), idx_, a, b, PADO_GPU_QUERY_REF_COMPAT, &out),
                       "pado_gpu_query_pair");
        return (weightiLarge)out;
    }
    // Exact u32 distance (PADO_GPU_WINF if unreachable).
    uint32_t query_distance_exact(idi a, idi b)
    {
        uint32_t out = 0;
        detail::die_on(pado_gpu_query_pair(detail::shared_ctx(), idx_, a, b, 0, &out), "pado_gpu_query_pair");
        return out;
    }
    // Batched exact form (one GPU launch for all pairs).
    void query_distance(const std::vector<idi> &a, const std::vector<idi> &b, std::vector<uint32_t> &out)
    {
        if (a.size() != b.size()) {
            fprintf(stderr, "Error: query_distance: pair arrays differ in length\n");
            exit(EXIT_FAILURE);
        }
        out.resize(a.size());
        detail::die_on(pado_gpu_wquery(detail::shared_ctx(), idx_, a.data(), b.data(), a.size(), out.data(), 0,
                                       nullptr),
                       "pado_gpu_wquery");
    }

    // store_index_to_file (WPARA:1982-2018).
    void store_index_to_file(const char *filename, const std::vector<idi> &rank)
    {
        pado_gpu_status s =
            pado_gpu_windex_store(detail::shared_ctx(), idx_, filename, rank.data(), (uint32_t)rank.size());
        if (s == PADO_GPU_ERR_IO) {
            fprintf(stderr, "Error: cannot open file %s\n", filename);
            exit(EXIT_FAILURE);
        }
        detail::die_on(s, "pado_gpu_windex_store");
        pado_gpu_index_info info;
        pado_gpu_index_get_info(idx_, &info);
        printf("Label_size: %lu mean: %f\n", (unsigned long)info.num_labels,
               num_v_ ? (double)info.num_labels / num_v_ : 0.0);
    }

    // load_index_from_file (WPARA:2021-2055).
    void load_index_from_file(const char *filename)
    {
        release();
        pado_gpu_status s = pado_gpu_windex_load(detail::shared_ctx(), filename, &idx_);
        if (s == PADO_GPU_ERR_IO) {
            fprintf(stderr, "Error: cannot open file %s\n", filename);
            exit(EXIT_FAILURE);
        }
        detail::die_on(s, "pado_gpu_windex_load");
        pado_gpu_index_info info;
        pado_gpu_index_get_info(idx_, &info);
        num_v_ = info.num_v;
        printf("Label_size_loaded: %lu mean: %f\n", (unsigned long)info.num_labels,
               num_v_ ? (double)info.num_labels / num_v_ : 0.0);
    }

    pado_gpu_index *gpu_index() { return idx_; }
};

// The sequential weighted class builds the same label sets; here both run on
// the GPU.
template <inti BATCH_SIZE = 512>
class WeightedVertexCentricPLL : public WeightedParaVertexCentricPLL<BATCH_SIZE> {
public:
    WeightedVertexCentricPLL() = default;
    WeightedVertexCentricPLL(const WeightedGraph &G) : WeightedParaVertexCentricPLL<BATCH_SIZE>(G) {}
};

}  // namespace PADO

#endif  // PADO_GPU_DROPIN_WEIGHTED_PARA_H_
