// pado/graph.h -- the reference's Graph / WeightedGraph classes
// (includes/graph.h:31-424) for the GPU drop-in.  Same public fields, same
// constructors, same make_rank / id_transfer semantics; the work is done by
// the host functions of libpado_gpu.so (include/pado_gpu.h).
#ifndef PADO_GPU_DROPIN_GRAPH_H_
#define PADO_GPU_DROPIN_GRAPH_H_

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "pado/globals.h"
#include "pado_gpu.h"

namespace PADO {

namespace detail {
// CSR of the reference's shape (idi offsets) <-> the ABI's u64 offsets.
inline void to_u64(const idi *vertices, idi num_v, idi num_e, std::vector<uint64_t> &off)
{
    off.resize((size_t)num_v + 1);
    for (idi v = 0; v < num_v; ++v) off[v] = vertices[v];
    off[num_v] = num_e;
}
}  // namespace detail

class Graph final {
public:
    idi num_v = 0;
    idi num_e = 0;  // directed edge count, 2 x input lines (includes/graph.h:127)
    idi *vertices = nullptr;
    idi *out_edges = nullptr;
    idi *out_degrees = nullptr;

    Graph() = default;
    // The text edge list, parsed on all host cores with the reference's line
    // rules (pado_read_edge_text).
    explicit Graph(const char *filename)
    {
        uint32_t *flat = nullptr;
        uint64_t m = 0;
        if (pado_read_edge_text(filename, 2, &flat, &m) != 0) {
            fprintf(stderr, "Error: cannot open file %s\n", filename);
            exit(EXIT_FAILURE);
        }
        construct_flat(flat, m);
        pado_free(flat);
    }
    explicit Graph(std::vector<std::pair<idi, idi>> &edge_list) { construct(edge_list); }
    Graph(const Graph &) = delete;
    Graph &operator=(const Graph &) = delete;
    ~Graph()
    {
        free(vertices);
        free(out_edges);
        free(out_degrees);
    }
    idi get_num_v() const { return num_v; }
    idi get_num_e() const { return num_e; }
    idi ith_get_edge(idi i, idi e) const { return out_edges[vertices[i] + e]; }
    idi ith_get_out_degree(idi i) const { return out_degrees[i]; }

    void construct(const std::vector<std::pair<idi, idi>> &edge_list)
    {
        std::vector<uint32_t> flat(2 * edge_list.size());
        for (size_t i = 0; i < edge_list.size(); ++i) {
            flat[2 * i] = edge_list[i].first;
            flat[2 * i + 1] = edge_list[i].second;
        }
        construct_flat(flat.data(), edge_list.size());
    }

    // Graph::construct (includes/graph.h:125-153) on m (head, tail) pairs.
    void construct_flat(const uint32_t *flat, uint64_t m)
    {
        num_e = (idi)(2 * m);
        num_v = m == 0 ? 0 : pado_graph_num_v(flat, m);
        std::vector<uint64_t> off((size_t)num_v + 1, 0);
        vertices = (idi *)malloc(sizeof(idi) * (num_v ? num_v : 1));
        out_edges = (idi *)malloc(sizeof(idi) * (num_e ? num_e : 1));
        out_degrees = (idi *)malloc(sizeof(idi) * (num_v ? num_v : 1));
        if (num_v) pado_graph_csr(flat, m, num_v, off.data(), out_edges);
        for (idi v = 0; v < num_v; ++v) {
            vertices[v] = (idi)off[v];
            out_degrees[v] = (idi)(off[v + 1] - off[v]);
        }
    }

    // Degree order with rand() jitter, consuming the process-wide rand() as
    // the reference does (includes/graph.h:156-174).
    std::vector<idi> make_rank() const
    {
        std::vector<uint64_t> off;
        detail::to_u64(vertices, num_v, num_e, off);
        std::vector<idi> rank(num_v);
        pado_make_rank(off.data(), num_v, 1, rank.data());
        return rank;
    }

    // Renumber into rank space, adjacency sorted descending; returns rank2id
    // (includes/graph.h:176-206).
    std::vector<idi> id_transfer(const std::vector<idi> &rank)
    {
        std::vector<uint64_t> off, off2((size_t)num_v + 1);
        detail::to_u64(vertices, num_v, num_e, off);
        std::vector<idi> adj2(num_e ? num_e : 1), rank2id(num_v);
        pado_id_transfer(num_v, off.data(), out_edges, rank.data(), off2.data(), adj2.data(), rank2id.data());
        for (idi v = 0; v < num_v; ++v) {
            vertices[v] = (idi)off2[v];
            out_degrees[v] = (idi)(off2[v + 1] - off2[v]);
        }
        if (num_e) memcpy(out_edges, adj2.data(), sizeof(idi) * num_e);
        return rank2id;
    }

    void print()
    {
        printf("num_v: %u, num_e: %u\n", num_v, num_e / 2);
        for (idi h = 0; h < num_v; ++h)
            for (idi e = vertices[h]; e < vertices[h] + out_degrees[h]; ++e) printf("%u %u\n", h, out_edges[e]);
    }
};

// WeightedGraph (includes/graph.h:223-408): Graph's CSR plus out_weights.
class WeightedGraph final {
public:
    idi num_v = 0;
    idi num_e = 0;
    idi *vertices = nullptr;
    idi *out_edges = nullptr;
    idi *out_degrees = nullptr;
    weighti *out_weights = nullptr;

    WeightedGraph() = default;
    // "head tail weight" lines (includes/graph.h:277-304), parsed on all host cores with the reference's line rules
    // (pado_read_edge_text); weights cast to weighti as the reference does.
    explicit WeightedGraph(const char *filename)
    {
        uint32_t *rows = nullptr;
        uint64_t m = 0;
        if (pado_read_edge_text(filename, 3, &rows, &m) != 0) {
            fprintf(stderr, "Error: cannot open file %s\n", filename);
            exit(EXIT_FAILURE);
        }
        std::vector<std::pair<idi, idi>> edge_list(m);
        std::vector<weighti> weight_list(m);
        for (uint64_t i = 0; i < m; ++i) {
            edge_list[i] = std::make_pair(rows[3 * i], rows[3 * i + 1]);
            weight_list[i] = static_cast<weighti>(rows[3 * i + 2]);
        }
        pado_free(rows);
        construct(edge_list, weight_list);
    }
    WeightedGraph(const std::vector<std::pair<idi, idi>> &edge_list, const std::vector<weighti> &weight_list)
    {
        construct(edge_list, weight_list);
    }
    WeightedGraph(const WeightedGraph &) = delete;
    WeightedGraph &operator=(const WeightedGraph &) = delete;
    ~WeightedGraph()
    {
        free(vertices);
        free(out_edges);
        free(out_degrees);
        free(out_weights);
    }
    idi get_num_v() const { return num_v; }
    idi get_num_e() const { return num_e; }
    idi ith_get_edge(idi i, idi e) const { return out_edges[vertices[i] + e]; }
    idi ith_get_out_degree(idi i) const { return out_degrees[i]; }

    void construct(const std::vector<std::pair<idi, idi>> &edge_list, const std::vector<weighti> &weight_list)
    {
        const size_t m = edge_list.size();
        std::vector<uint32_t> flat(2 * m), w32(m);
        for (size_t i = 0; i < m; ++i) {
            flat[2 * i] = edge_list[i].first;
            flat[2 * i + 1] = edge_list[i].second;
            w32[i] = i < weight_list.size() ? weight_list[i] : 0;
        }
        num_e = (idi)(2 * m);
        num_v = m ? pado_graph_num_v(flat.data(), m) : 0;
        std::vector<uint64_t> off((size_t)num_v + 1, 0);
        vertices = (idi *)malloc(sizeof(idi) * (num_v ? num_v : 1));
        out_edges = (idi *)malloc(sizeof(idi) * (num_e ? num_e : 1));
        out_degrees = (idi *)malloc(sizeof(idi) * (num_v ? num_v : 1));
        out_weights = (weighti *)malloc(sizeof(weighti) * (num_e ? num_e : 1));
        if (num_v) pado_wgraph_csr(flat.data(), w32.data(), m, num_v, off.data(), out_edges, out_weights);
        for (idi v = 0; v < num_v; ++v) {
            vertices[v] = (idi)off[v];
            out_degrees[v] = (idi)(off[v + 1] - off[v]);
        }
    }

    // Same degree order as Graph::make_rank (includes/graph.h:353-372).
    std::vector<idi> make_rank()
    {
        std::vector<uint64_t> off;
        detail::to_u64(vertices, num_v, num_e, off);
        std::vector<idi> rank(num_v);
        pado_make_rank(off.data(), num_v, 1, rank.data());
        return rank;
    }

    // (id, weight) lists sorted descending (includes/graph.h:374-408).
    std::vector<idi> id_transfer(const std::vector<idi> &rank)
    {
        std::vector<uint64_t> off, off2((size_t)num_v + 1);
        detail::to_u64(vertices, num_v, num_e, off);
        std::vector<idi> adj2(num_e ? num_e : 1), rank2id(num_v);
        std::vector<weighti> w2(num_e ? num_e : 1);
        pado_wid_transfer(num_v, off.data(), out_edges, out_weights, rank.data(), off2.data(), adj2.data(),
                          w2.data(), rank2id.data());
        for (idi v = 0; v < num_v; ++v) {
            vertices[v] = (idi)off2[v];
            out_degrees[v] = (idi)(off2[v + 1] - off2[v]);
        }
        if (num_e) {
            memcpy(out_edges, adj2.data(), sizeof(idi) * num_e);
            memcpy(out_weights, w2.data(), sizeof(weighti) * num_e);
        }
        return rank2id;
    }

    void print()
    {
        printf("num_v: %u, num_e: %u\n", num_v, num_e / 2);
        for (idi h = 0; h < num_v; ++h)
            for (idi e = vertices[h]; e < vertices[h] + out_degrees[h]; ++e)
                printf("%u %u %u\n", h, out_edges[e], (unsigned)out_weights[e]);
    }
};

// The reference's binary edge list (distributed_version/tools/graph-convert/
// edgelist2binary.cpp:28-91: u32 num_v, u64 num_e, then the pairs) as the
// edge list Graph(std::vector<pair>) takes; exits on a missing or malformed
// file, as the reference's readers do.
inline std::vector<std::pair<idi, idi>> read_binary_edge_list(const char *filename)
{
    uint32_t *flat = nullptr;
    uint64_t m = 0;
    if (pado_read_edge_binary(filename, &flat, &m, nullptr) != 0) {
        fprintf(stderr, "Error: cannot read binary edge list %s\n", filename);
        exit(EXIT_FAILURE);
    }
    std::vector<std::pair<idi, idi>> edge_list(m);
    for (uint64_t i = 0; i < m; ++i) edge_list[i] = std::make_pair(flat[2 * i], flat[2 * i + 1]);
    pado_free(flat);
    return edge_list;
}

}  // namespace PADO

#endif  // PADO_GPU_DROPIN_GRAPH_H_
