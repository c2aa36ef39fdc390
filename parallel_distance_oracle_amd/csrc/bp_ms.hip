// bp_ms.hip -- bit-parallel labeling with all 50 slots in one level-synchronous
// sweep (multi-source BFS), gfx950.
//
// Reference: bit_parallel_labeling, includes/pado_unw_unv_para.201912311607.
// parallel_bp_labeling.h:544-840 (semantics restated in bp.hip's header).  The
// 50 BFSes are independent once the roots and their selected neighbours are
// chosen (a serial host pass, PARA:670-709), so they advance together: a
// vertex carries a 64-bit mask of the slots in which it sits on the current
// level, and one pass over its adjacency serves every such slot.  A vertex is
// expanded once per distinct level it has across the slots (typically 2-4 on
// R-MAT graphs) instead of once per slot.
//
// Per level d (ping-pong masks fm = level d, nm = level d+1):
//   k_ms_sib   : S^0[v][i] |= S^-1[u][i] for edges (v,u), u != v, both at level d in slot i
//   k_ms_child : slots where v is at d and u is unvisited -> u joins level d+1
//                there (nm), bp_dist = d+1, S^-1/S^0 of v pushed into u's slot
//   k_ms_update: vis |= nm for the new frontier; fm cleared for the old one
// Finally S^0 is masked with ~S^-1 (PARA:829-835).  The set unions are 64-bit
// atomicOr, order-free, so the records are deterministic.
#include "internal.hpp"

namespace pado {
namespace {

using u64 = unsigned long long;

__device__ inline void ms_emit(uint32_t v, uint64_t deg, u64 *chunks, uint32_t *cnt_chunks, uint32_t *list,
                               uint32_t *cnt_list)
{
    list[atomicAdd(cnt_list, 1u)] = v;
    uint32_t nch = (uint32_t)((deg + kChunk - 1) / kChunk);
    if (nch == 0) return;
    uint32_t c0 = atomicAdd(cnt_chunks, nch);
    for (uint32_t q = 0; q < nch; ++q) chunks[c0 + q] = ((u64)(q * kChunk) << 32) | v;
}

// Seeds: roots at level 0 (fm), selected neighbours at level 1 (nm, vis, S^-1 bit).
// seeds[j] = (vertex, slot, neighbour index or 255 for the root).
__global__ void k_ms_seed(uint32_t nseeds, const uint32_t *seed_v, const uint32_t *seed_slot,
                          const uint32_t *seed_j, const uint64_t *off, uint8_t *bpd, ulonglong2 *bps, u64 *fm,
                          u64 *nm, u64 *vis, u64 *cur_chunks, uint32_t *cnt_cur_chunks, uint32_t *cur_list,
                          uint32_t *cnt_cur_list, u64 *nxt_chunks, uint32_t *cnt_nxt_chunks, uint32_t *nxt_list,
                          uint32_t *cnt_nxt_list)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nseeds; s += gridDim.x * blockDim.x) {
        const uint32_t v = seed_v[s], i = seed_slot[s], j = seed_j[s];
        const u64 bit = 1ull << i;
        atomicOr(&vis[v], bit);
        if (j == 255u) {
            bpd[(uint64_t)v * kBpdStride + i] = 0;
            if (atomicOr(&fm[v], bit) == 0ull) ms_emit(v, off[v + 1] - off[v], cur_chunks, cnt_cur_chunks, cur_list,
                                                       cnt_cur_list);
        } else {
            bpd[(uint64_t)v * kBpdStride + i] = 1;
            bps[(uint64_t)v * kBP + i].x = 1ull << j;
            if (atomicOr(&nm[v], bit) == 0ull) ms_emit(v, off[v + 1] - off[v], nxt_chunks, cnt_nxt_chunks, nxt_list,
                                                       cnt_nxt_list);
        }
    }
}

// The 16 lanes of a chunk share v: per slot their S^-1 sets are OR-reduced
// across the 16 lanes and one lane applies the union, so v's record sees one
// atomic per slot and chunk instead of one per edge.
__global__ __launch_bounds__(256) void k_ms_sib(const uint64_t *off, const uint32_t *adj, const u64 *fm,
                                                ulonglong2 *bps, const u64 *chunks, const uint32_t *cnt_chunks)
{
    const uint32_t lane = threadIdx.x & (kChunk - 1);
    const uint32_t ngroups = (gridDim.x * blockDim.x) / kChunk;
    const uint32_t nc = *cnt_chunks;
    // c is uniform over each 16-lane group, so the group runs the loop body
    // together (the shuffles below stay inside the group).
    for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) / kChunk; c < nc; c += ngroups) {
        const u64 desc = chunks[c];
        const uint32_t v = (uint32_t)desc;
        const uint64_t e = off[v] + (desc >> 32) + lane;
        u64 m = 0ull;
        uint32_t u = v;
        if (e < off[v + 1]) {
            u = adj[e];
            if (u != v) m = fm[v] & fm[u];
        }
        u64 un = m;
        for (int o = 1; o < (int)kChunk; o <<= 1) un |= __shfl_xor(un, o);
        while (un) {
            const uint32_t i = __ffsll(un) - 1;
            un &= un - 1;
            u64 x = ((m >> i) & 1ull) ? bps[(uint64_t)u * kBP + i].x : 0ull;
            for (int o = 1; o < (int)kChunk; o <<= 1) x |= __shfl_xor(x, o);
            if (lane == 0 && x) atomicOr(&bps[(uint64_t)v * kBP + i].y, x);
        }
    }
}

__global__ __launch_bounds__(256) void k_ms_child(uint32_t d, const uint64_t *off, const uint32_t *adj, const u64 *fm,
                                                  u64 *nm, const u64 *vis, uint8_t *bpd, ulonglong2 *bps,
                                                  const u64 *chunks, const uint32_t *cnt_chunks, u64 *nxt_chunks,
                                                  uint32_t *cnt_nxt_chunks, uint32_t *nxt_list, uint32_t *cnt_nxt_list)
{
    const uint32_t lane = threadIdx.x & (kChunk - 1);
    const uint32_t ngroups = (gridDim.x * blockDim.x) / kChunk;
    const uint32_t nc = *cnt_chunks;
    for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) / kChunk; c < nc; c += ngroups) {
        const u64 desc = chunks[c];
        const uint32_t v = (uint32_t)desc;
        const uint64_t e = off[v] + (desc >> 32) + lane;
        if (e >= off[v + 1]) continue;
        const uint32_t u = adj[e];
        const u64 m0 = fm[v] & ~vis[u];
        if (!m0) continue;
        const u64 prev = atomicOr(&nm[u], m0);
        if (prev == 0ull) ms_emit(u, off[u + 1] - off[u], nxt_chunks, cnt_nxt_chunks, nxt_list, cnt_nxt_list);
        // The lane that sets u's level bit of a slot writes its distance
        // (every parent would write the same d + 1).
        const u64 fresh = m0 & ~prev;
        u64 m = m0;
        while (m) {
            const uint32_t i = __ffsll(m) - 1;
            m &= m - 1;
            if ((fresh >> i) & 1ull) bpd[(uint64_t)u * kBpdStride + i] = (uint8_t)(d + 1);
            const ulonglong2 sv = bps[(uint64_t)v * kBP + i];
            if (sv.x) atomicOr(&bps[(uint64_t)u * kBP + i].x, sv.x);
            if (sv.y) atomicOr(&bps[(uint64_t)u * kBP + i].y, sv.y);
        }
    }
}

__global__ __launch_bounds__(256) void k_ms_update(const uint32_t *old_list, const uint32_t *cnt_old, u64 *fm_old,
                                                   const uint32_t *new_list, const uint32_t *cnt_new, const u64 *nm,
                                                   u64 *vis)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t no = *cnt_old, nn = *cnt_new;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += stride) {
        const uint32_t u = new_list[i];
        vis[u] |= nm[u];
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < no; i += stride) fm_old[old_list[i]] = 0ull;
}

__global__ void k_ms_finish(uint32_t n, ulonglong2 *bps)
{
    const uint64_t total = (uint64_t)n * kBP;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        ulonglong2 s = bps[i];
        s.y &= ~s.x;
        bps[i] = s;
    }
}

__global__ void k_ms_fill(uint8_t *bpd, uint64_t nbpd, u64 *z0, uint64_t nz0)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbpd; i += stride) bpd[i] = kInf8;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nz0; i += stride) z0[i] = 0ull;
}

__global__ void k_ms_zero_counts(uint32_t *c) { c[threadIdx.x] = 0; }

}  // namespace

void select_bp_roots(uint32_t n, const uint64_t *off, const uint32_t *adj, std::vector<uint8_t> &used,
                     uint32_t *roots, BpSeeds &seeds, const uint32_t *map)
{
    used.assign(n, 0);
    uint32_t r = 0;
    for (uint32_t i = 0; i < kBP; ++i) {
        while (r < n && used[r]) ++r;
        if (r == n) {
            if (roots) roots[i] = kEmpty;
            continue;
        }
        if (roots) roots[i] = r;
        used[r] = 1;
        seeds.v.push_back(map ? map[r] : r);
        seeds.slot.push_back(i);
        seeds.j.push_back(255u);
        const uint64_t deg = off[r + 1] - off[r];
        uint32_t ns = 0;
        for (uint64_t di = 0; di < deg && ns < 64; ++di) {
            const uint32_t v = adj[off[r] + deg - 1 - di];
            if (!used[v]) {
                used[v] = 1;
                seeds.v.push_back(map ? map[v] : v);
                seeds.slot.push_back(i);
                seeds.j.push_back(ns);
                ++ns;
            }
        }
    }
}

pado_gpu_status build_bp_multi(pado_gpu_graph *g, pado_gpu_index *idx, DevBuf<uint8_t> &d_used,
                               std::vector<uint8_t> &h_used)
{
    // Serial root / neighbour selection (PARA:670-709).
    BpSeeds seeds;
    select_bp_roots(g->n, g->h_off.data(), g->h_adj.data(), h_used, idx->bp_roots, seeds, nullptr);
    return build_bp_seeded(g, idx, d_used, h_used, seeds);
}

pado_gpu_status build_bp_seeded(pado_gpu_graph *g, pado_gpu_index *idx, DevBuf<uint8_t> &d_used,
                                const std::vector<uint8_t> &h_used, const BpSeeds &seeds)
{
    hipStream_t st = g->ctx->stream;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    PADO_TRY(idx->bpd.alloc((uint64_t)n * kBpdStride));
    PADO_TRY(idx->bps.alloc((uint64_t)n * kBP * 2));
    PADO_TRY(d_used.alloc(n));
    const std::vector<uint32_t> &sv = seeds.v, &ss = seeds.slot, &sj = seeds.j;
    PADO_HIP(hipMemcpyAsync(d_used.p, h_used.data(), n, hipMemcpyHostToDevice, st));

    DevBuf<u64> fma, fmb, vis, ca, cb;
    DevBuf<uint32_t> la, lb, cnt, dseeds;
    PADO_TRY(fma.alloc(n));
    PADO_TRY(fmb.alloc(n));
    PADO_TRY(vis.alloc(n));
    const uint64_t max_chunks = m / kChunk + n + 64;
    PADO_TRY(ca.alloc(max_chunks));
    PADO_TRY(cb.alloc(max_chunks));
    PADO_TRY(la.alloc(n));
    PADO_TRY(lb.alloc(n));
    PADO_TRY(cnt.alloc(4));  // [chunks a, list a, chunks b, list b]
    const uint32_t ns = (uint32_t)sv.size();
    PADO_TRY(dseeds.alloc(3ull * (ns ? ns : 1)));
    std::vector<uint32_t> packed(sv);
    packed.insert(packed.end(), ss.begin(), ss.end());
    packed.insert(packed.end(), sj.begin(), sj.end());
    if (ns) PADO_HIP(hipMemcpyAsync(dseeds.p, packed.data(), 4ull * packed.size(), hipMemcpyHostToDevice, st));

    hipLaunchKernelGGL(k_ms_fill, dim3(8192), dim3(256), 0, st, idx->bpd.p, (uint64_t)n * kBpdStride,
                       (u64 *)idx->bps.p, (uint64_t)n * kBP * 2);
    hipLaunchKernelGGL(k_ms_fill, dim3(4096), dim3(256), 0, st, (uint8_t *)nullptr, 0ull, fma.p, (uint64_t)n);
    hipLaunchKernelGGL(k_ms_fill, dim3(4096), dim3(256), 0, st, (uint8_t *)nullptr, 0ull, fmb.p, (uint64_t)n);
    hipLaunchKernelGGL(k_ms_fill, dim3(4096), dim3(256), 0, st, (uint8_t *)nullptr, 0ull, vis.p, (uint64_t)n);
    hipLaunchKernelGGL(k_ms_zero_counts, dim3(1), dim3(4), 0, st, cnt.p);

    uint32_t *h_cnt = nullptr;
    PADO_HIP(hipHostMalloc((void **)&h_cnt, 4 * sizeof(uint32_t), hipHostMallocDefault));
    struct PinnedFree {
        uint32_t *p;
        ~PinnedFree()
        {
            std::lock_guard<std::recursive_mutex> g(capture_mutex());
            (void)hipHostFree(p);
        }
    } pin{h_cnt};

    ulonglong2 *bps = (ulonglong2 *)idx->bps.p;
    u64 *fm = fma.p, *nm = fmb.p, *cur_c = ca.p, *nxt_c = cb.p;
    uint32_t *cur_l = la.p, *nxt_l = lb.p;
    uint32_t *cnt_cur_c = cnt.p, *cnt_cur_l = cnt.p + 1, *cnt_nxt_c = cnt.p + 2, *cnt_nxt_l = cnt.p + 3;
    if (ns)
        hipLaunchKernelGGL(k_ms_seed, dim3(grid_for(ns, 256, 4096)), dim3(256), 0, st, ns, dseeds.p, dseeds.p + ns, dseeds.p + 2 * ns, g->off.p,
                           idx->bpd.p, bps, fm, nm, vis.p, cur_c, cnt_cur_c, cur_l, cnt_cur_l, nxt_c, cnt_nxt_c,
                           nxt_l, cnt_nxt_l);
    // Grids follow the frontier: a level's edge kernels get enough threads
    // for its descriptors in one pass (16 lanes each), up to kMsGrid blocks;
    // level 0's size is not read back, so it gets the most.
    constexpr uint32_t kMsGrid = 4096;
    uint32_t cur_chunks = ~0u;
    uint32_t max_level = 0;
    hipEvent_t e0, e1;
    PADO_HIP(hipEventCreate(&e0));
    PADO_HIP(hipEventCreate(&e1));
    for (uint32_t d = 0; ns; ++d) {
        const uint32_t egrid = cur_chunks == ~0u
                                   ? kMsGrid
                                   : (uint32_t)std::min<uint64_t>(kMsGrid, std::max<uint64_t>(1, ((uint64_t)cur_chunks * kChunk + 255) / 256));
        PADO_HIP(hipEventRecord(e0, st));
        hipLaunchKernelGGL(k_ms_sib, dim3(egrid), dim3(256), 0, st, g->off.p, g->adj.p, fm, bps, cur_c, cnt_cur_c);
        hipLaunchKernelGGL(k_ms_child, dim3(egrid), dim3(256), 0, st, d, g->off.p, g->adj.p, fm, nm, vis.p,
                           idx->bpd.p, bps, cur_c, cnt_cur_c, nxt_c, cnt_nxt_c, nxt_l, cnt_nxt_l);
        // (the new frontier has at most one vertex per edge of this one)
        hipLaunchKernelGGL(k_ms_update, dim3(egrid), dim3(256), 0, st, cur_l, cnt_cur_l, fm, nxt_l, cnt_nxt_l, nm,
                           vis.p);
        PADO_HIP(hipEventRecord(e1, st));
        PADO_HIP(hipMemcpyAsync(h_cnt, cnt.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        PADO_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        PADO_HIP(hipEventElapsedTime(&ms, e0, e1));
        idx->ktimes.ms_bp += ms;
        idx->ktimes.n_bp += 1;
        // the counter slots swap roles every level: read them through the pointers
        const uint32_t next_count = h_cnt[cnt_nxt_l - cnt.p];
        cur_chunks = h_cnt[cnt_nxt_c - cnt.p];
        if (next_count == 0) break;
        if (d + 1 > max_level) max_level = d + 1;
        if (d + 1 >= kInf8) {
            // Wide hop domain: the sweep stops after level 254 (whose sibling
            // sets are complete); vertices farther away keep INF (255) in
            // that slot -- level 255 was written as 255 -- and the labeling
            // and query kernels ignore such slots.
            if (idx->wide) break;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            set_error("bit-parallel BFS reached level 255: graph eccentricity is outside the uint8 "
                      "distance domain of the reference (its level counter wraps at 255)");
            return PADO_GPU_ERR_DOMAIN;
        }
        // The next level becomes current; the old current buffers (mask cleared
        // by k_ms_update, counters reset here) receive level d+2.
        std::swap(fm, nm);
        std::swap(cur_c, nxt_c);
        std::swap(cur_l, nxt_l);
        std::swap(cnt_cur_c, cnt_nxt_c);
        std::swap(cnt_cur_l, cnt_nxt_l);
        PADO_HIP(hipMemsetAsync(cnt_nxt_c, 0, sizeof(uint32_t), st));
        PADO_HIP(hipMemsetAsync(cnt_nxt_l, 0, sizeof(uint32_t), st));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hipLaunchKernelGGL(k_ms_finish, dim3(8192), dim3(256), 0, st, n, bps);
    PADO_HIP(hipStreamSynchronize(st));
    idx->bp_max_level = max_level;
    return PADO_GPU_OK;
}

}  // namespace pado
