// ingest.hip -- device-side graph ingest: the reference's Graph::construct,
// make_rank and id_transfer (includes/graph.h:120-206; WeightedGraph
// :313-408) from an edge list, producing the rank-space CSR in HBM.
//
// Per undirected input edge (a, b, w): both directions, no de-duplication,
// self-loops kept (GR:138-141).  The degree order is the reference's
// make_rank on the host (it consumes rand() in vertex order; only the n
// degrees cross PCIe).  Renumbering scatters every direction as a 64-bit key
// (rank[b] << 8 | w) into its tail's segment, and a segmented descending sort
// gives each list in id_transfer's order: neighbours descending, parallel
// edges by weight descending (the (id, weight) pair sort of GR:374-408).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "internal.hpp"

using namespace pado;

namespace {

__global__ void k_degrees(const uint32_t *e, uint64_t m, uint32_t *deg)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        atomicAdd(&deg[e[2 * i]], 1u);
        atomicAdd(&deg[e[2 * i + 1]], 1u);
    }
}

__global__ void k_permute(uint32_t n, const uint32_t *rank, const uint32_t *deg, uint32_t *deg_new)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) deg_new[rank[v]] = deg[v];
}

__global__ void k_scatter(const uint32_t *e, const uint8_t *w, uint64_t m, const uint32_t *rank, const uint64_t *off,
                          uint32_t *cursor, unsigned long long *keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const uint32_t a = rank[e[2 * i]], b = rank[e[2 * i + 1]];
        const unsigned long long wi = w ? w[i] : 0u;
        keys[off[a] + atomicAdd(&cursor[a], 1u)] = ((unsigned long long)b << 8) | wi;
        keys[off[b] + atomicAdd(&cursor[b], 1u)] = ((unsigned long long)a << 8) | wi;
    }
}

__global__ void k_split_keys(uint64_t total, const unsigned long long *keys, uint32_t *adj, uint8_t *w)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        adj[i] = (uint32_t)(keys[i] >> 8);
        if (w) w[i] = (uint8_t)(keys[i] & 255u);
    }
}

}  // namespace

extern "C" pado_gpu_status pado_gpu_graph_from_edges(pado_gpu_ctx *ctx, const uint32_t *edges, uint64_t m,
                                                     const uint32_t *weights, const uint32_t *rank_in,
                                                     int use_global_rand, pado_gpu_graph **out, uint32_t *num_v,
                                                     uint32_t *rank_out, uint32_t *rank2id_out)
{
    clear_error();
    if (!ctx || !out || (m && !edges)) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    PADO_HIP(hipSetDevice(ctx->device));
    if (m == 0) {
        set_error("empty edge list");
        return PADO_GPU_ERR_ARG;
    }
    if (2 * m >= (1ull << 31)) {
        set_error("device ingest takes fewer than 2^30 edges (2^31 directed)");
        return PADO_GPU_ERR_ARG;
    }
    const uint32_t n = pado_graph_num_v(edges, m);
    if (n >= kMaxVertices) {
        set_error("vertex count must be < 2^24");
        return PADO_GPU_ERR_ARG;
    }
    const uint64_t md = 2 * m;
    std::vector<uint8_t> w8;
    if (weights) {
        w8.resize(m);
        for (uint64_t i = 0; i < m; ++i) {
            w8[i] = (uint8_t)weights[i];  // static_cast<weighti>(weight), GR:293
            if (w8[i] == 0) {
                set_error("edge weights must be in [1, 255]");
                return PADO_GPU_ERR_ARG;
            }
        }
    }
    hipStream_t st = ctx->stream;
    DevBuf<uint32_t> d_e, deg, rank, deg_new, cursor;
    DevBuf<uint8_t> d_w;
    DevBuf<unsigned long long> keys, sorted;
    PADO_TRY(d_e.alloc(md));
    PADO_TRY(deg.alloc(n));
    PADO_HIP(hipMemcpyAsync(d_e.p, edges, 4 * md, hipMemcpyHostToDevice, st));
    if (weights) {
        PADO_TRY(d_w.alloc(m));
        PADO_HIP(hipMemcpyAsync(d_w.p, w8.data(), m, hipMemcpyHostToDevice, st));
    }
    PADO_HIP(hipMemsetAsync(deg.p, 0, 4ull * n, st));
    hipLaunchKernelGGL(k_degrees, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, d_e.p, m, deg.p);
    // make_rank on the host from the degrees (GR:156-174).
    std::vector<uint32_t> h_rank(n);
    if (rank_in) {
        std::vector<uint8_t> seen(n, 0);
        for (uint32_t v = 0; v < n; ++v) {
            if (rank_in[v] >= n || seen[rank_in[v]]) {
                set_error("rank is not a permutation of [0, num_v)");
                return PADO_GPU_ERR_ARG;
            }
            seen[rank_in[v]] = 1;
            h_rank[v] = rank_in[v];
        }
    } else {
        std::vector<uint32_t> h_deg(n);
        PADO_HIP(hipMemcpyAsync(h_deg.data(), deg.p, 4ull * n, hipMemcpyDeviceToHost, st));
        PADO_HIP(hipStreamSynchronize(st));
        std::vector<uint64_t> off0((uint64_t)n + 1, 0);
        for (uint32_t v = 0; v < n; ++v) off0[v + 1] = off0[v] + h_deg[v];
        if (pado_make_rank(off0.data(), n, use_global_rand, h_rank.data()) != 0) {
            set_error("make_rank failed");
            return PADO_GPU_ERR_HIP;
        }
    }
    // id_transfer (GR:176-206): renumber, then each list sorted descending.
    auto *g = new pado_gpu_graph();
    struct Guard {
        pado_gpu_graph *&g;
        ~Guard() { delete g; }
    } guard{g};
    g->ctx = ctx;
    g->n = n;
    g->m = md;
    g->weighted = weights != nullptr;
    PADO_TRY(rank.alloc(n));
    PADO_TRY(deg_new.alloc(n));
    PADO_TRY(cursor.alloc(n));
    PADO_TRY(g->off.alloc((uint64_t)n + 1));
    PADO_TRY(g->adj.alloc(md));
    if (weights) PADO_TRY(g->w.alloc(md));
    PADO_TRY(keys.alloc(md));
    PADO_TRY(sorted.alloc(md));
    PADO_HIP(hipMemcpyAsync(rank.p, h_rank.data(), 4ull * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_permute, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, rank.p, deg.p, deg_new.p);
    PADO_TRY(scan_u32_to_u64(deg_new.p, g->off.p, n, st));
    PADO_HIP(hipMemsetAsync(cursor.p, 0, 4ull * n, st));
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(m, 256, 8192)), dim3(256), 0, st, d_e.p,
                       weights ? d_w.p : (const uint8_t *)nullptr, m, rank.p, g->off.p, cursor.p, keys.p);
    size_t bytes = 0;
    PADO_HIP(hipcub::DeviceSegmentedSort::SortKeysDescending(nullptr, bytes, keys.p, sorted.p, (int)md, (int)n,
                                                             g->off.p, g->off.p + 1, st));
    DevBuf<uint8_t> ws;
    PADO_TRY(ws.alloc(bytes));
    PADO_HIP(hipcub::DeviceSegmentedSort::SortKeysDescending(ws.p, bytes, keys.p, sorted.p, (int)md, (int)n,
                                                             g->off.p, g->off.p + 1, st));
    hipLaunchKernelGGL(k_split_keys, dim3(grid_for(md, 256, 8192)), dim3(256), 0, st, md, sorted.p, g->adj.p,
                       weights ? g->w.p : (uint8_t *)nullptr);
    // Host copies (the bit-parallel root selection is a serial host pass).
    g->h_off.resize((uint64_t)n + 1);
    g->h_adj.resize(md);
    PADO_HIP(hipMemcpyAsync(g->h_off.data(), g->off.p, 8ull * (n + 1), hipMemcpyDeviceToHost, st));
    PADO_HIP(hipMemcpyAsync(g->h_adj.data(), g->adj.p, 4ull * md, hipMemcpyDeviceToHost, st));
    if (weights) {
        g->h_w.resize(md);
        PADO_HIP(hipMemcpyAsync(g->h_w.data(), g->w.p, md, hipMemcpyDeviceToHost, st));
    }
    PADO_HIP(hipStreamSynchronize(st));
    if (num_v) *num_v = n;
    if (rank_out) std::copy(h_rank.begin(), h_rank.end(), rank_out);
    if (rank2id_out)
        for (uint32_t v = 0; v < n; ++v) rank2id_out[h_rank[v]] = v;
    *out = g;
    g = nullptr;  // ownership passes to the caller
    return PADO_GPU_OK;
}

extern "C" pado_gpu_status pado_gpu_graph_export(pado_gpu_ctx *ctx, const pado_gpu_graph *g, uint64_t *off,
                                                 uint32_t *adj, uint8_t *w)
{
    clear_error();
    if (!ctx || !g) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    PADO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (off) PADO_HIP(hipMemcpyAsync(off, g->off.p, 8ull * (g->n + 1), hipMemcpyDeviceToHost, st));
    if (adj && g->m) PADO_HIP(hipMemcpyAsync(adj, g->adj.p, 4ull * g->m, hipMemcpyDeviceToHost, st));
    if (w && g->weighted && g->m) PADO_HIP(hipMemcpyAsync(w, g->w.p, g->m, hipMemcpyDeviceToHost, st));
    PADO_HIP(hipStreamSynchronize(st));
    return PADO_GPU_OK;
}

extern "C" pado_gpu_status pado_gpu_graph_get_info(const pado_gpu_graph *g, uint32_t *num_v, uint64_t *m_dir,
                                                   int *weighted)
{
    clear_error();
    if (!g) {
        set_error("null graph");
        return PADO_GPU_ERR_ARG;
    }
    if (num_v) *num_v = g->n;
    if (m_dir) *m_dir = g->m;
    if (weighted) *weighted = g->weighted ? 1 : 0;
    return PADO_GPU_OK;
}
