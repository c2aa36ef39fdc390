// bp.hip -- bit-parallel labeling on gfx950.
//
// Reference: bit_parallel_labeling, includes/pado_unw_unv_para.201912311607.
// parallel_bp_labeling.h:544-840.  For each of the 50 slots the reference picks
// the first unused vertex r (rank order) and up to 64 of its unused neighbours
// (scanned from the END of r's descending adjacency), then runs a BFS from r in
// which every vertex v carries two 64-bit sets over the selected neighbours:
//   S^-1[v] = selected u with d(u, v) = d(r, v) - 1
//   S^0[v]  = selected u with d(u, v) = d(r, v)
// Level d first ORs S^-1 of same-level neighbours into S^0 (sibling edges,
// PARA:783-792), then pushes both sets along parent->child edges
// (PARA:804-813).  The stored record is (d, S^-1, S^0 \ S^-1) (PARA:829-835).
//
// GPU mapping: root/neighbour selection is a serial host pass over the CSR
// (it depends only on the adjacency and the used flags, in slot order).  Each
// BFS level is two edge-parallel kernels over the frontier's adjacency -- a
// sibling pass and a child pass, in the reference's order -- with 64-bit
// atomicOr for the set unions (order-free, so the result is deterministic).
// Work is split into 16-edge chunks, one per quarter-wave, so a hub with 10^5
// neighbours spreads over thousands of lanes.
#include "internal.hpp"

namespace pado {
namespace {

__device__ inline void emit_chunks(uint32_t v, uint64_t deg, unsigned long long *chunks, uint32_t *cnt)
{
    uint32_t nch = (uint32_t)((deg + kChunk - 1) / kChunk);
    if (nch == 0) return;
    uint32_t c0 = atomicAdd(cnt, nch);
    for (uint32_t q = 0; q < nch; ++q)
        chunks[c0 + q] = ((unsigned long long)(q * kChunk) << 32) | v;
}

__global__ void k_bp_reset(uint32_t n, uint8_t *lvl, unsigned long long *s1, unsigned long long *s0,
                           uint32_t *vis)
{
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        lvl[v] = kInf8;
        s1[v] = 0;
        s0[v] = 0;
        if ((v & 31) == 0) vis[v >> 5] = 0;
    }
}

// Root at level 0 and its selected neighbours at level 1 (PARA:680-709).
// cur receives the root's chunks, next the selected neighbours' chunks.
__global__ void k_bp_seed(uint32_t r, const uint32_t *sel, uint32_t ns, const uint64_t *off, uint8_t *lvl,
                          unsigned long long *s1, uint32_t *vis, unsigned long long *cur, uint32_t *cnt_cur,
                          unsigned long long *next, uint32_t *cnt_next)
{
    if (threadIdx.x == 0) {
        *cnt_cur = 0;
        *cnt_next = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        lvl[r] = 0;
        atomicOr(&vis[r >> 5], 1u << (r & 31));
        emit_chunks(r, off[r + 1] - off[r], cur, cnt_cur);
    }
    for (uint32_t j = threadIdx.x; j < ns; j += blockDim.x) {
        uint32_t v = sel[j];
        lvl[v] = 1;
        s1[v] = 1ull << j;
        atomicOr(&vis[v >> 5], 1u << (v & 31));
        emit_chunks(v, off[v + 1] - off[v], next, cnt_next);
    }
}

// Sibling pass of level d: S^0[v] |= S^-1[u] for every neighbour u of a level-d
// vertex v that is itself at level d (u != v).  Each endpoint updates its own
// set, which covers both directions of the reference's (v < u) pair.
__global__ __launch_bounds__(256) void k_bp_sib(uint32_t d, const uint64_t *off, const uint32_t *adj,
                                                const uint8_t *lvl, const unsigned long long *s1,
                                                unsigned long long *s0, const unsigned long long *cur,
                                                const uint32_t *cnt_cur)
{
    const uint32_t lane = threadIdx.x & (kChunk - 1);
    const uint32_t ngroups = (gridDim.x * blockDim.x) / kChunk;
    const uint32_t nc = *cnt_cur;
    for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) / kChunk; c < nc; c += ngroups) {
        unsigned long long desc = cur[c];
        uint32_t v = (uint32_t)desc;
        uint64_t e = off[v] + (desc >> 32) + lane;
        if (e >= off[v + 1]) continue;
        uint32_t u = adj[e];
        if (u != v && lvl[u] == d) {
            unsigned long long x = s1[u];
            if (x) atomicOr(&s0[v], x);
        }
    }
}

// Child pass of level d (PARA:731-771, 804-813): unvisited neighbours join
// level d+1 (the visited-bitmap atomicOr decides who appends them), and every
// edge into level d+1 carries S^-1 and S^0 of the parent.
__global__ __launch_bounds__(256) void k_bp_child(uint32_t d, const uint64_t *off, const uint32_t *adj,
                                                  uint8_t *lvl, unsigned long long *s1, unsigned long long *s0,
                                                  uint32_t *vis, const unsigned long long *cur,
                                                  const uint32_t *cnt_cur, unsigned long long *next,
                                                  uint32_t *cnt_next)
{
    const uint32_t lane = threadIdx.x & (kChunk - 1);
    const uint32_t ngroups = (gridDim.x * blockDim.x) / kChunk;
    const uint32_t nc = *cnt_cur;
    for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) / kChunk; c < nc; c += ngroups) {
        unsigned long long desc = cur[c];
        uint32_t v = (uint32_t)desc;
        uint64_t e = off[v] + (desc >> 32) + lane;
        if (e >= off[v + 1]) continue;
        uint32_t u = adj[e];
        uint32_t l = lvl[u];
        if (l <= d) continue;
        if (l == kInf8) {
            uint32_t bit = 1u << (u & 31);
            uint32_t old = atomicOr(&vis[u >> 5], bit);
            if (!(old & bit)) {
                lvl[u] = (uint8_t)(d + 1);
                emit_chunks(u, off[u + 1] - off[u], next, cnt_next);
            }
        }
        unsigned long long x1 = s1[v], x0 = s0[v];
        if (x1) atomicOr(&s1[u], x1);
        if (x0) atomicOr(&s0[u], x0);
    }
}

// Slot record (PARA:829-835): bp_dist = level, sets = (S^-1, S^0 \ S^-1).
__global__ void k_bp_store(uint32_t n, uint32_t slot, const uint8_t *lvl, const unsigned long long *s1,
                           const unsigned long long *s0, uint8_t *bpd, unsigned long long *bps)
{
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        bpd[(uint64_t)v * kBpdStride + slot] = lvl[v];
        unsigned long long a = s1[v], b = s0[v];
        bps[((uint64_t)v * kBP + slot) * 2 + 0] = a;
        bps[((uint64_t)v * kBP + slot) * 2 + 1] = b & ~a;
    }
}

__global__ void k_fill_u8(uint8_t *p, uint64_t count, uint8_t val)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = val;
}

__global__ void k_fill_u64(unsigned long long *p, uint64_t count, unsigned long long val)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = val;
}

}  // namespace

pado_gpu_status build_bp(pado_gpu_graph *g, pado_gpu_index *idx, DevBuf<uint8_t> &d_used,
                         std::vector<uint8_t> &h_used)
{
    hipStream_t st = g->ctx->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    PADO_TRY(idx->bpd.alloc((uint64_t)n * kBpdStride));
    PADO_TRY(idx->bps.alloc((uint64_t)n * kBP * 2));
    PADO_TRY(d_used.alloc(n));
    // Unused slots keep bp_dist = 255 and empty sets (PARA:674-679).
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for((uint64_t)n * kBpdStride, 256, 8192)), dim3(256), 0, st,
                       idx->bpd.p, (uint64_t)n * kBpdStride, kInf8);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for((uint64_t)n * kBP * 2, 256, 8192)), dim3(256), 0, st,
                       (unsigned long long *)idx->bps.p, (uint64_t)n * kBP * 2, 0ull);

    // Serial root / neighbour selection (PARA:670-709).
    h_used.assign(n, 0);
    std::vector<uint32_t> sel_all(kBP * 64, 0);
    std::vector<uint32_t> nsel(kBP, 0);
    uint32_t r = 0;
    for (uint32_t i = 0; i < kBP; ++i) {
        while (r < n && h_used[r]) ++r;
        if (r == n) {
            idx->bp_roots[i] = kEmpty;
            continue;
        }
        idx->bp_roots[i] = r;
        h_used[r] = 1;
        uint64_t deg = g->h_off[r + 1] - g->h_off[r];
        for (uint64_t di = 0; di < deg; ++di) {
            uint32_t v = g->h_adj[g->h_off[r] + deg - 1 - di];
            if (!h_used[v]) {
                h_used[v] = 1;
                sel_all[i * 64 + nsel[i]] = v;
                if (++nsel[i] == 64) break;
            }
        }
    }
    PADO_HIP(hipMemcpyAsync(d_used.p, h_used.data(), n, hipMemcpyHostToDevice, st));

    DevBuf<uint8_t> lvl;
    DevBuf<unsigned long long> s1, s0, ca, cb;
    DevBuf<uint32_t> vis, sel, cnt;
    PADO_TRY(lvl.alloc(n));
    PADO_TRY(s1.alloc(n));
    PADO_TRY(s0.alloc(n));
    PADO_TRY(vis.alloc((n + 31) / 32 + 1));
    const uint64_t max_chunks = m / kChunk + n + 64;
    PADO_TRY(ca.alloc(max_chunks));
    PADO_TRY(cb.alloc(max_chunks));
    PADO_TRY(sel.alloc(kBP * 64));
    PADO_TRY(cnt.alloc(2));
    PADO_HIP(hipMemcpyAsync(sel.p, sel_all.data(), sizeof(uint32_t) * kBP * 64, hipMemcpyHostToDevice, st));

    uint32_t *h_cnt = nullptr;
    PADO_HIP(hipHostMalloc((void **)&h_cnt, 2 * sizeof(uint32_t), hipHostMallocDefault));
    struct PinnedFree {
        uint32_t *p;
        ~PinnedFree()
        {
            std::lock_guard<std::recursive_mutex> g(capture_mutex());
            (void)hipHostFree(p);
        }
    } pin{h_cnt};

    const uint32_t edge_grid = 2048;
    uint32_t max_level = 0;
    for (uint32_t i = 0; i < kBP; ++i) {
        if (idx->bp_roots[i] == kEmpty) continue;
        hipLaunchKernelGGL(k_bp_reset, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, lvl.p, s1.p, s0.p,
                           vis.p);
        unsigned long long *cur = ca.p, *next = cb.p;
        uint32_t *cnt_cur = cnt.p, *cnt_next = cnt.p + 1;
        hipLaunchKernelGGL(k_bp_seed, dim3(1), dim3(64), 0, st, idx->bp_roots[i], sel.p + i * 64, nsel[i],
                           g->off.p, lvl.p, s1.p, vis.p, cur, cnt_cur, next, cnt_next);
        for (uint32_t d = 0;; ++d) {
            hipLaunchKernelGGL(k_bp_sib, dim3(edge_grid), dim3(256), 0, st, d, g->off.p, g->adj.p, lvl.p, s1.p,
                               s0.p, cur, cnt_cur);
            hipLaunchKernelGGL(k_bp_child, dim3(edge_grid), dim3(256), 0, st, d, g->off.p, g->adj.p, lvl.p, s1.p,
                               s0.p, vis.p, cur, cnt_cur, next, cnt_next);
            PADO_HIP(hipMemcpyAsync(h_cnt, cnt_next, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            PADO_HIP(hipStreamSynchronize(st));
            if (h_cnt[0] == 0) break;
            if (d + 1 > max_level) max_level = d + 1;
            if (d + 1 >= kInf8) {
                set_error("bit-parallel BFS reached level 255: graph eccentricity is outside the uint8 "
                          "distance domain of the reference (its level counter wraps at 255)");
                return PADO_GPU_ERR_DOMAIN;
            }
            std::swap(cur, next);
            std::swap(cnt_cur, cnt_next);
            // Reset the (now) next counter for level d+2.
            PADO_HIP(hipMemsetAsync(cnt_next, 0, sizeof(uint32_t), st));
        }
        hipLaunchKernelGGL(k_bp_store, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, i, lvl.p, s1.p, s0.p,
                           idx->bpd.p, (unsigned long long *)idx->bps.p);
    }
    PADO_HIP(hipStreamSynchronize(st));
    idx->bp_max_level = max_level;
    return PADO_GPU_OK;
}

}  // namespace pado
