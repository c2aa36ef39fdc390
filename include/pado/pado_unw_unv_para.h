// pado/pado_unw_unv_para.h -- drop-in for the reference's
// ParaVertexCentricPLL<BATCH_SIZE> (includes/pado_unw_unv_para*.h:42-304) and
// VertexCentricPLL<BATCH_SIZE> (includes/pado_unw_unv_unp.h:42-257), backed
// by the MI355X library through its C ABI (include/pado_gpu.h).
//
// Call sites such as
//     Graph G(file); auto rank = G.make_rank(); auto rank2id = G.id_transfer(rank);
//     ParaVertexCentricPLL<1024> *p = new ParaVertexCentricPLL<1024>(G);
//     p->order_labels(rank2id, rank); p->query_distance(a, b);
//     p->store_index_to_file(out, rank);
// compile unchanged.  Errors follow the reference: I/O failures print to
// stderr and exit(EXIT_FAILURE); device failures do the same (the reference
// has no other error channel).  The device is HIP device 0, or
// $PADO_GPU_DEVICE.
#ifndef PADO_GPU_DROPIN_PARA_H_
#define PADO_GPU_DROPIN_PARA_H_

#include <stdio.h>
#include <stdlib.h>

#include <utility>
#include <vector>

#include "pado/globals.h"
#include "pado/graph.h"
#include "pado_gpu.h"

namespace PADO {

namespace detail {

inline void die_on(pado_gpu_status s, const char *what)
{
    if (s != PADO_GPU_OK) {
        fprintf(stderr, "Error: %s: %s\n", what, pado_gpu_last_error());
        exit(EXIT_FAILURE);
    }
}

// One context per process (one device), created on first use.
inline pado_gpu_ctx *shared_ctx()
{
    static pado_gpu_ctx *ctx = [] {
        const char *env = getenv("PADO_GPU_DEVICE");
        pado_gpu_ctx *c = nullptr;
        die_on(pado_gpu_ctx_create(env ? atoi(env) : 0, &c), "pado_gpu_ctx_create");
        return c;
    }();
    return ctx;
}

}  // namespace detail

template <inti BATCH_SIZE = 1024>
class ParaVertexCentricPLL {
    static_assert(BATCH_SIZE >= 1 && BATCH_SIZE <= 1024, "BATCH_SIZE must be in [1, 1024]");

    pado_gpu_index *idx_ = nullptr;
    idi num_v_ = 0;

    void release()
    {
        if (idx_) pado_gpu_index_free(idx_);
        idx_ = nullptr;
    }

public:
    static const inti BITPARALLEL_SIZE = 50;

    ParaVertexCentricPLL() = default;
    // Builds the index of a rank-space graph (construct, PARA:309-313, 1740-1895).
    ParaVertexCentricPLL(const Graph &G) { construct(G); }
    ParaVertexCentricPLL(const ParaVertexCentricPLL &) = delete;
    ParaVertexCentricPLL &operator=(const ParaVertexCentricPLL &) = delete;
    ~ParaVertexCentricPLL() { release(); }

    void construct(const Graph &G)
    {
        release();
        pado_gpu_ctx *ctx = detail::shared_ctx();
        std::vector<uint64_t> off;
        detail::to_u64(G.vertices, G.num_v, G.num_e, off);
        pado_gpu_graph *g = nullptr;
        double t = -WallTimer::get_time_mark();
        detail::die_on(pado_gpu_graph_upload(ctx, G.num_v, G.num_e, off.data(), G.out_edges, nullptr, &g),
                       "pado_gpu_graph_upload");
        pado_gpu_build_cfg cfg = {};
        cfg.batch_size = BATCH_SIZE;
        pado_gpu_status s = pado_gpu_build(ctx, g, &cfg, &idx_);
        pado_gpu_graph_free(g);
        detail::die_on(s, "pado_gpu_build");
        t += WallTimer::get_time_mark();
        num_v_ = G.num_v;
        pado_gpu_index_info info;
        pado_gpu_index_get_info(idx_, &info);
        // The reference's construct() report (PARA:1845-1889).
        printf("Threads: %u Batch_size: %u\n", NUM_THREADS, BATCH_SIZE);
        printf("BP_Roots_Size: %u\n", BITPARALLEL_SIZE);
        printf("Total_labeling_time: %.2f seconds bp_hit_count: %lu\n", t, (unsigned long)info.bp_pruned);
    }

    // order_labels (PARA:2198-2281): queries then take original ids.
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

    // query_distance (PARA:2283-2416): exact hop distance, 255 if unreachable
    // or >= 253; ids >= num_v give (a == b ? 0 : 255).  The reference's
    // unsigned bit-parallel gate over-estimates some distances to its
    // bit-parallel roots; this returns the exact value (see
    // query_distance_reference for the reference's answer).
    // One launch per call, no copies (pado_gpu_query_pair); a loop over many
    // pairs is faster through the batched form below.
    weighti query_distance(idi a, idi b)
    {
        uint32_t out = 0;
        detail::die_on(pado_gpu_query_pair(detail::shared_ctx(), idx_, a, b, 0, &out), "pado_gpu_query_pair");
        return (weighti)out;
    }
    weighti query_distance_reference(idi a, idi b)
    {
        uint32_t out = 0;
        detail::die_on(pado_gpu_query_pair(detail::shared_ctx(), idx_, a, b, PADO_GPU_QUERY_REF_COMPAT, &out),
                       "pado_gpu_query_pair");
        return (weighti)out;
    }
    // Batched form (one GPU launch for all pairs).
    void query_distance(const std::vector<idi> &a, const std::vector<idi> &b, std::vector<weighti> &out)
    {
        out.resize(a.size());
        if (a.size() != b.size()) {
            fprintf(stderr, "Error: query_distance: pair arrays differ in length\n");
            exit(EXIT_FAILURE);
        }
        detail::die_on(pado_gpu_query(detail::shared_ctx(), idx_, a.data(), b.data(), a.size(), out.data(), 0, nullptr),
                       "pado_gpu_query");
    }

    // Writes the reference's index file (UNP:1121-1180; the OpenMP class's
    // writer only prints Label_size, PARA:2082-2151).
    void store_index_to_file(const char *filename, const std::vector<idi> &rank)
    {
        pado_gpu_status s = pado_gpu_index_store(detail::shared_ctx(), idx_, filename, rank.data(), (uint32_t)rank.size());
        if (s == PADO_GPU_ERR_IO) {
            fprintf(stderr, "Error: cannot open file %s\n", filename);
            exit(EXIT_FAILURE);
        }
        detail::die_on(s, "pado_gpu_index_store");
        pado_gpu_index_info info;
        pado_gpu_index_get_info(idx_, &info);
        printf("Label_size: %lu mean: %f\n", (unsigned long)info.num_labels,
               num_v_ ? (double)info.num_labels / num_v_ : 0.0);
    }

    // load_index_from_file (PARA:2154-2195 / UNP:1183-1224).
    void load_index_from_file(const char *filename)
    {
        release();
        pado_gpu_status s = pado_gpu_index_load(detail::shared_ctx(), filename, &idx_);
        if (s == PADO_GPU_ERR_IO) {
            fprintf(stderr, "Error: cannot open file %s\n", filename);
            exit(EXIT_FAILURE);
        }
        detail::die_on(s, "pado_gpu_index_load");
        pado_gpu_index_info info;
        pado_gpu_index_get_info(idx_, &info);
        num_v_ = info.num_v;
        printf("Label_size_loaded: %lu mean: %f\n", (unsigned long)info.num_labels,
               num_v_ ? (double)info.num_labels / num_v_ : 0.0);
    }

    // switch_labels_to_old_id (PARA:2419-2466) only counts labels.
    void switch_labels_to_old_id(const std::vector<idi> &rank2id, const std::vector<idi> &rank)
    {
        (void)rank2id;
        pado_gpu_index_info info;
        pado_gpu_index_get_info(idx_, &info);
        printf("Label sum: %lu %lu mean: %f\n", (unsigned long)info.num_labels, (unsigned long)info.num_labels,
               rank.empty() ? 0.0 : (double)info.num_labels / rank.size());
    }

    // Canonical export (rank space) for callers that want the raw index.
    pado_gpu_index *gpu_index() { return idx_; }
};

template <inti BATCH_SIZE>
const inti ParaVertexCentricPLL<BATCH_SIZE>::BITPARALLEL_SIZE;

// The sequential class of the reference builds the same label sets
// (UNP:42-257); here both run on the GPU.
template <inti BATCH_SIZE = 1024>
class VertexCentricPLL : public ParaVertexCentricPLL<BATCH_SIZE> {
public:
    VertexCentricPLL() = default;
    VertexCentricPLL(const Graph &G) : ParaVertexCentricPLL<BATCH_SIZE>(G) {}
    std::pair<uint64_t, uint64_t> length_larger_than_16 = std::make_pair(0, 0);
};

}  // namespace PADO

#endif  // PADO_GPU_DROPIN_PARA_H_
