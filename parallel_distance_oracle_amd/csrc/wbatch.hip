// wbatch.hip -- weighted pruned landmark labeling on gfx950.
//
// Reference: WeightedParaVertexCentricPLL<512>, includes/pado_weighted_unv_para.h
// (initialize_tables :305-383, send_messages :435-522, distance_query
// :527-914, filter_out_labels :1100-1228, reset_tables :918-956,
// update_index :961-990, vertex_centric_labeling_in_batches :1231-1566).
//
// The reference's labels are the canonical pruned landmark labels: (r, d) is
// a label of v iff r ranks at or above v, d = d(v, r), and no vertex ranked
// above r lies on a shortest v-r path.  Its schedule (queue orders, first
// witness returned by the distance query, in-place filtering) does not change
// that set, and this build reaches the same set with a schedule-free scheme,
// per batch of roots:
//   1. Bellman-Ford rounds from all batch roots at once.  A root's distance
//      travels only to lower-ranked tails (send_messages, WPARA:457-470); a
//      (tail, root) value that improves is checked against the labels of
//      earlier batches (distance_query's first loop, WPARA:569-580): if a
//      2-hop path through an earlier hub is as short, the pair is marked
//      pruned and does not propagate.  Rounds end when nothing improves.
//   2. Filter (filter_out_labels): a surviving pair (v, r) is dropped if some
//      batch root h < r has surviving pairs (v, h), (r, h) with
//      S(v,h) + S(r,h) <= S(v,r).
// Why this is exact: a canonical pair's shortest paths avoid every vertex
// ranked above r, so the restricted search finds d(v,r) and no earlier hub can
// prune it; a non-canonical pair has some higher vertex on a shortest path,
// and the highest such vertex h* either belongs to an earlier batch (then the
// earlier labels prune the pair at any value >= d) or to this batch (then
// (v,h*) and (r,h*) are canonical, exact, and witness the filter).
//
// Layout per batch (B = 512 roots): T[n][B] u32 tentative distances, P and I
// bitmaps (pruned at the current value / improved this round) [n][B/64] u64,
// the earlier-batch root-distance table (hmin[n], dcw[hub][B] u32), and the
// frontier as (vertex, root bitmap) entries with 16-edge push chunks.  Labels
// are packed u64 (hub << 32 | dist) in per-vertex growable arrays.
#include <cstring>

#include "internal.hpp"

namespace pado {
namespace {

constexpr uint32_t kWB = 512;         // roots per batch (WPARA BATCH_SIZE)
constexpr uint32_t kWW = kWB / 64;    // 64-bit words per root bitmap
constexpr uint32_t kWInf = 0xFFFFFFFFu;
using u64 = unsigned long long;

// Frontier / improved-tail counts of round i live in slot i & 1: each round's
// kernels read their slot and empty the other one for the next kernel.
struct WCounters {
    uint32_t n_act[2], n_chunk[2], n_imp[2], n_touched, overflow;
    uint32_t a_top[2];  // active-list slots handed out (>= n_act: per-wave reservations)
    u64 pool_top, relax, improved, pruned, labels, filtered, rounds, need;
    // Pool words the batch's commit can reserve at most: per touched vertex
    // the reservation its list would take if every batch root were kept.
    u64 need_bound;
};

struct WStore {
    u64 *pool;
    u64 *base;
    uint32_t *len, *cap;
    u64 pool_size;
};

__device__ inline uint32_t lane_id() { return threadIdx.x & 63; }

__device__ inline u64 wave_sum(u64 x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Tails > lo form a prefix of a descending adjacency list (WPARA:457 breaks
// at the first tail <= roots_start).
__device__ inline uint32_t prefix_above(const uint64_t *off, const uint32_t *adj, uint32_t v, uint32_t lo)
{
    const uint32_t lane = lane_id();
    const uint64_t base = off[v];
    uint32_t a = 0, b = (uint32_t)(off[v + 1] - base);
    while (b - a > 64) {
        uint32_t step = (b - a + 63) / 64;
        uint32_t pos = a + lane * step;
        bool p = pos < b && adj[base + pos] > lo;
        uint32_t c = (uint32_t)__popcll(__ballot(p));
        if (c == 0) {
            b = a;
        } else {
            uint32_t na = a + (c - 1) * step + 1, nb = a + c * step;
            a = na;
            if (nb < b) b = nb;
        }
    }
    uint32_t pos = a + lane;
    bool p = pos < b && adj[base + pos] > lo;
    return a + (uint32_t)__popcll(__ballot(p));
}

__device__ inline u64 store_reserve(WStore S, WCounters *C, uint32_t v, uint32_t len, uint32_t add)
{
    const uint32_t lane = lane_id();
    u64 base = S.base[v];
    uint32_t cap = S.cap[v];
    if (len + add <= cap) return base;
    uint32_t ncap = cap * 2;
    if (ncap < len + add) ncap = len + add;
    if (ncap < 8) ncap = 8;
    u64 nb = 0;
    if (lane == 0) nb = atomicAdd(&C->pool_top, (u64)ncap);
    nb = __shfl(nb, 0);
    if (nb + ncap > S.pool_size) {
        if (lane == 0) atomicOr(&C->overflow, 1u);
        return ~0ull;
    }
    for (uint32_t j = lane; j < len; j += 64) S.pool[nb + j] = S.pool[base + j];
    if (lane == 0) {
        S.base[v] = nb;
        S.cap[v] = ncap;
    }
    return nb;
}

// Pool words a commit of up to `add` new labels may reserve for a list of
// length len and capacity cap (store_reserve's doubling rule).
__device__ inline u64 reserve_words(uint32_t len, uint32_t cap, uint32_t add)
{
    if (len + add <= cap) return 0;
    uint32_t ncap = cap * 2;
    if (ncap < len + add) ncap = len + add;
    if (ncap < 8) ncap = 8;
    return ncap;
}

// Round 1's frontier is the batch's roots in order: root offset ro holds
// active-list slot ro and touched-list entry ro (k_w_init).
__global__ void k_w_reset_counters(WCounters *C, uint32_t rsz)
{
    C->n_act[0] = C->n_chunk[0] = C->n_chunk[1] = C->n_imp[0] = C->n_imp[1] = 0;
    C->a_top[0] = 0;
    C->n_act[1] = C->a_top[1] = rsz;
    C->n_touched = rsz;
    C->need = C->need_bound = 0;
}

// Earlier-batch labels of the batch roots -> root-distance table
// (roots_labels_buffer, WPARA:322-348, transposed): dcw[h][ro] = d(root
// rs + ro, hub h), indexed by hub id, and hmin[h] = the least of them.  Both
// are INF outside the batch roots' hubs (k_w_commit resets what was written).
__device__ inline void w_hub_fill(uint32_t ro, uint32_t r, WStore S, uint32_t *hmin, uint32_t *dcw)
{
    const uint32_t lane = lane_id();
    const uint32_t len = S.len[r];
    const u64 base = S.base[r];
    // Lists are hub-sorted and every root holds the top-ranked hubs: walking
    // them from the end staggers the roots' arrival at those hubs, so the
    // plain read of hmin skips most atomics.
    for (uint32_t j0 = (len + 63) & ~63u; j0 > 0; j0 -= 64) {
        const uint32_t j = j0 - 64 + lane;
        if (j >= len) continue;
        const u64 lab = S.pool[base + j];
        const uint32_t h = (uint32_t)(lab >> 32), d = (uint32_t)lab;
        dcw[(uint64_t)h * kWB + ro] = d;
        if (d < hmin[h]) atomicMin(&hmin[h], d);
    }
}

// initialize_tables (WPARA:305-383): every root at distance 0 from itself,
// active with its own bit, and its earlier labels in the root-distance table.
// One wave per root, one chunk reservation per block.
__global__ __launch_bounds__(256) void k_w_init(uint32_t rs, uint32_t rsz, uint32_t gen_batch, const uint64_t *off,
                                                const uint32_t *adj, uint32_t *T, uint32_t *tmark, uint32_t *tlist,
                                                uint32_t *alist, u64 *ag, uint32_t *amask, uint32_t *aend, u64 *chunks,
                                                uint32_t *hmin, uint32_t *dcw, WStore S, WCounters *C)
{
    __shared__ uint32_t s_nch[4], s_c0;
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t ro = blockIdx.x * 4 + wv;
    uint32_t nch = 0;
    if (ro < rsz) {
        const uint32_t r = rs + ro;
        if (lane == 0) {
            T[(uint64_t)r * kWB + ro] = 0;
            tmark[r] = gen_batch;
            tlist[ro] = r;
            alist[ro] = r;
            amask[ro] = 1u << (ro >> 6);
            const u64 nb = reserve_words(S.len[r], S.cap[r], rsz);
            if (nb) atomicAdd(&C->need_bound, nb);
        }
        if (lane < kWW) ag[(uint64_t)ro * kWW + lane] = (lane == (ro >> 6)) ? (1ull << (ro & 63)) : 0ull;
        const uint32_t pdeg = prefix_above(off, adj, r, rs);
        if (lane == 0) aend[ro] = pdeg;
        nch = (pdeg + kChunk - 1) / kChunk;
    }
    if (lane == 0) s_nch[wv] = nch;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = s_nch[0] + s_nch[1] + s_nch[2] + s_nch[3];
        s_c0 = tot ? atomicAdd(&C->n_chunk[1], tot) : 0u;
    }
    __syncthreads();
    uint32_t c0 = s_c0;
    for (uint32_t w = 0; w < wv; ++w) c0 += s_nch[w];
    for (uint32_t q = lane; q < nch; q += 64) chunks[c0 + q] = ((u64)(q * kChunk) << 32) | ro;
    if (ro < rsz) w_hub_fill(ro, rs + ro, S, hmin, dcw);
}


// Position of the rank-th set bit of m (rank < popcount(m)).
__device__ inline uint32_t w_nth_set_bit(u64 m, uint32_t rank)
{
    uint32_t pos = 0;
    uint32_t c = __popc((uint32_t)m);
    if (rank >= c) {
        rank -= c;
        m >>= 32;
        pos = 32;
    }
    uint32_t x = (uint32_t)m;
#pragma unroll
    for (uint32_t width = 16; width; width >>= 1) {
        c = __popc(x & ((1u << width) - 1u));
        if (rank >= c) {
            rank -= c;
            x >>= width;
            pos += width;
        }
    }
    return pos;
}

// Orders LDS writes of some lanes of a wave before reads by other lanes.
__device__ inline void w_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

constexpr uint32_t kWPushWaves = 4;  // waves per k_w_push block

// Touched vertices up to which k_w_filter reads the roots' entries in place.
constexpr uint32_t kWDirectFilter = 4096;

// Improved pairs of a tail up to which k_w_check reads single table entries.
constexpr uint32_t kWGather = 8;

// k_w_check reserves next-frontier slots per wave: kWAlistGrab active-list
// entries and up to kWChunkGrab push descriptors per atomic on the shared
// counters; a descriptor slot reserved but not used holds kWChunkPad (edge
// offset past every list), which k_w_push skips.
constexpr uint32_t kWAlistGrab = 64, kWChunkGrab = 64;
constexpr u64 kWChunkPad = 0xFFFFFFF0ull << 32;

// send_messages (WPARA:435-522): relax the frontier's edges toward tails
// ranked below each root.  A wave takes four 16-edge chunks (one edge per
// lane) and walks the root words of its edges' groups in step; per word the
// (edge, root) relaxations of all 64 lanes are spread over the wave, 64 per
// step, so an edge with many roots does not hold up the others.
__global__ __launch_bounds__(256) void k_w_push(uint32_t round, uint32_t rs, uint32_t rsz, uint32_t gen_batch,
                                                const uint64_t *off, const uint32_t *adj, const uint8_t *wgt,
                                                const uint32_t *alist, const u64 *ag, const uint32_t *amask,
                                                const uint32_t *aend, const u64 *chunks, uint32_t *T, u64 *I,
                                                uint32_t *tmark, uint32_t *tlist, u64 *imark, uint32_t *ilist,
                                                WStore S, WCounters *C)
{
    __shared__ u64 s_g[kWPushWaves][64], s_imp[kWPushWaves][64];
    __shared__ uint32_t s_ex[kWPushWaves][64], s_u[kWPushWaves][64], s_t[kWPushWaves][64], s_w[kWPushWaves][64];
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t wave = blockIdx.x * kWPushWaves + wv;
    const uint32_t nwaves = gridDim.x * kWPushWaves;
    // Round `round` of the batch (slot round & 1; empty past the batch's
    // end).  Block 0 empties the other slot, which k_w_check fills.
    const uint32_t p = round & 1u;
    const uint32_t nc = C->n_chunk[p];
    const u64 gen_round = ((u64)gen_batch << 24) | round;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (C->n_act[p]) C->rounds += 1;
        C->n_act[p ^ 1u] = 0;
        C->a_top[p ^ 1u] = 0;
        C->n_chunk[p ^ 1u] = 0;
    }
    uint32_t n_relax = 0, n_imp = 0;
    u64 n_bound = 0;
    for (uint32_t c0 = wave * 4; c0 < nc; c0 += nwaves * 4) {
        const uint32_t c = c0 + (lane >> 4);
        uint32_t k = 0, u = 0, t = 0, w = 0, lim = 0, words = 0;
        const u64 desc = c < nc ? chunks[c] : kWChunkPad;
        if (desc != kWChunkPad) {
            k = (uint32_t)desc;
            const uint32_t e = (uint32_t)(desc >> 32) + (lane & (kChunk - 1));
            if (e < aend[k]) {
                u = alist[k];
                const uint64_t ei = off[u] + e;
                t = adj[ei];
                w = wgt[ei];
                lim = t - rs;  // roots with offset < lim rank above t
                words = amask[k] & (lim >= kWB ? ((1u << kWW) - 1u) : ((1u << ((lim + 63) / 64)) - 1u));
            }
        }
        s_u[wv][lane] = u;
        s_t[wv][lane] = t;
        s_w[wv][lane] = w;
        uint32_t wall = words;
        for (int o = 32; o > 0; o >>= 1) wall |= __shfl_xor(wall, o);
        bool got = false;
        while (wall) {
            const uint32_t wd = __ffs(wall) - 1;
            wall &= wall - 1;
            u64 g = 0;
            if ((words >> wd) & 1u) {
                g = ag[(uint64_t)k * kWW + wd];
                if (lim < (wd + 1) * 64) g &= (1ull << (lim - wd * 64)) - 1ull;
            }
            const uint32_t cnt = (uint32_t)__popcll(g);
            uint32_t incl = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o);
                if (lane >= (uint32_t)o) incl += y;
            }
            const uint32_t total = __shfl(incl, 63);
            if (total == 0) continue;
            s_g[wv][lane] = g;
            s_ex[wv][lane] = incl - cnt;
            s_imp[wv][lane] = 0ull;
            w_lds_sync();
            for (uint32_t x0 = 0; x0 < total; x0 += 64) {
                const uint32_t x = x0 + lane;
                if (x >= total) continue;
                uint32_t o = 0;
#pragma unroll
                for (uint32_t sft = 32; sft; sft >>= 1)
                    if (s_ex[wv][o + sft] <= x) o += sft;
                const uint32_t b = w_nth_set_bit(s_g[wv][o], x - s_ex[wv][o]);
                const uint32_t ro = wd * 64 + b;
                const uint32_t du = T[(uint64_t)s_u[wv][o] * kWB + ro];
                if (du == kWInf) continue;
                const uint32_t cand = du + s_w[wv][o];
                ++n_relax;
                // A plain read first: values only fall during the kernel, so
                // one read as <= cand (even a stale one) makes the atomic a no-op.
                uint32_t *tt = &T[(uint64_t)s_t[wv][o] * kWB + ro];
                if (cand < *tt) {
                    const uint32_t old = atomicMin(tt, cand);
                    if (cand < old) atomicOr(&s_imp[wv][o], 1ull << b);
                }
            }
            w_lds_sync();
            const u64 imp = s_imp[wv][lane];
            if (imp) {
                atomicOr(&I[(uint64_t)t * kWW + wd], imp);
                got = true;
            }
            // The next word's writes must not overtake this word's reads.
            w_lds_sync();
        }
        const bool new_t = got && tmark[t] != gen_batch && atomicExch(&tmark[t], gen_batch) != gen_batch;
        const bool new_i = got && imark[t] != gen_round && atomicExch(&imark[t], gen_round) != gen_round;
        const uint32_t ti = wave_append(&C->n_touched, new_t);
        if (new_t) {
            tlist[ti] = t;
            n_bound += reserve_words(S.len[t], S.cap[t], rsz);
        }
        const uint32_t ii = wave_append(&C->n_imp[p], new_i);
        if (new_i) {
            ilist[ii] = t;
            ++n_imp;
        }
        // s_u / s_t / s_w are rewritten by the next step.
        w_lds_sync();
    }
    const u64 t_relax = wave_sum(n_relax), t_imp = wave_sum(n_imp), t_bound = wave_sum(n_bound);
    if (lane == 0) {
        if (t_relax) atomicAdd(&C->relax, t_relax);
        if (t_imp) atomicAdd(&C->improved, t_imp);
        if (t_bound) atomicAdd(&C->need_bound, t_bound);
    }
}

// distance_query's first loop (WPARA:569-580) for every improved pair of an
// improved tail: one wave per tail, lane l owns roots 8l..8l+7.  Pruned pairs
// stop propagating; the others become the tail's next frontier bitmap.
__global__ __launch_bounds__(256) void k_w_check(uint32_t round, uint32_t rs, const uint32_t *ilist, const uint64_t *off,
                                                 const uint32_t *adj, uint32_t *T, u64 *I, u64 *P, WStore S,
                                                 const uint32_t *hmin, const uint32_t *dcw,
                                                 uint32_t *alist, u64 *ag, uint32_t *amask, uint32_t *aend,
                                                 u64 *chunks, WCounters *C)
{
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t p = round & 1u;
    const uint32_t nimp = C->n_imp[p];
    if (blockIdx.x == 0 && threadIdx.x == 0) C->n_imp[p ^ 1u] = 0;
    u64 n_pruned = 0;
    uint32_t a_next = 0, a_left = 0, c_next = 0, c_left = 0, my_act = 0;
    for (uint32_t i = wave; i < nimp; i += nwaves) {
        const uint32_t t = ilist[i];
        uint8_t *Ib = reinterpret_cast<uint8_t *>(I + (uint64_t)t * kWW);
        uint8_t *Pb = reinterpret_cast<uint8_t *>(P + (uint64_t)t * kWW);
        const uint32_t imp = Ib[lane];
        // Clear after every lane has read (the store depends on the ballot).
        if (__ballot(imp != 0) != 0ull) Ib[lane] = 0;
        uint32_t dv[8];
        uint32_t maxd = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            dv[q] = (imp >> q) & 1u ? T[(uint64_t)t * kWB + lane * 8 + q] : 0u;
            if (((imp >> q) & 1u) && dv[q] > maxd) maxd = dv[q];
        }
        uint32_t wmax = maxd;
        for (int o = 32; o > 0; o >>= 1) {
            uint32_t y = __shfl_xor(wmax, o);
            wmax = y > wmax ? y : wmax;
        }
        uint32_t prune = 0;
        const uint32_t len = uniform32(S.len[t]);
        const u64 base = uniform64(S.base[t]);
        // Row phase: while more than kWGather improved pairs are unpruned,
        // lanes hold roots (8 each) and read the passing hubs' row slices.
        // The first hubs usually prune most pairs; once a 4-hub group leaves
        // at most kWGather unpruned, the labels not yet read go to the gather
        // phase.
        uint32_t j_gather = 0;
        if (wave_sum((u64)__popc(imp)) > kWGather) {
            j_gather = len;
            for (uint32_t j0 = 0; j0 < len; j0 += 64) {
                const uint32_t j = j0 + lane;
                uint32_t h = 0, d = kWInf;
                if (j < len) {
                    const u64 lab = S.pool[base + j];
                    h = (uint32_t)(lab >> 32);
                    d = (uint32_t)lab;
                }
                // hmin is INF for hubs no batch root holds
                const bool pass = j < len && (uint64_t)d + hmin[h] <= wmax;
                u64 m = __ballot(pass);
                // Up to four passing hubs per step: their row loads go out
                // together (32 B per lane each), and only lanes that still hold an
                // unpruned improved pair load.
                bool few = false;
                while (m) {
                    int src[4];
                    bool ok[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        ok[u] = m != 0ull;
                        src[u] = ok[u] ? __ffsll(m) - 1 : 0;
                        if (ok[u]) m &= m - 1;
                    }
                    const bool need = (imp & ~prune) != 0u;
                    uint32_t dd[4];
                    uint4 r0[4], r1[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        dd[u] = __shfl(d, src[u]);
                        const uint32_t sl = __shfl(h, src[u]);
                        r0[u] = r1[u] = make_uint4(kWInf, kWInf, kWInf, kWInf);
                        if (ok[u] && need) {
                            const uint4 *row = reinterpret_cast<const uint4 *>(dcw + (uint64_t)sl * kWB + lane * 8);
                            r0[u] = row[0];
                            r1[u] = row[1];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t x[8] = {r0[u].x, r0[u].y, r0[u].z, r0[u].w, r1[u].x, r1[u].y, r1[u].z, r1[u].w};
#pragma unroll
                        for (int q = 0; q < 8; ++q)
                            if (((imp >> q) & 1u) && x[q] != kWInf && (uint64_t)dd[u] + x[q] <= dv[q]) prune |= 1u << q;
                    }
                    if (wave_sum((u64)__popc(imp & ~prune)) <= kWGather) {
                        few = true;
                        break;
                    }
                }
                if (few) {
                    // the step's passing hubs not yet read start the gather phase
                    const uint32_t nxt = m ? j0 + (uint32_t)__ffsll(m) - 1 : j0 + 64;
                    j_gather = nxt < len ? nxt : len;
                    break;
                }
            }
        }
        // Gather phase: the (at most kWGather) improved pairs still unpruned
        // are checked hub-parallel over labels j_gather.. : lanes hold labels
        // and read one table entry per pair, 128 labels per step with the
        // next step's labels loaded while this one runs.  croot / cdv: the
        // pairs' root offsets and tentative distances (wave-uniform); cpr:
        // which are pruned.
        uint32_t croot[kWGather], cdv[kWGather];
        uint32_t ncand_t = 0, cpr = 0;
        {
            uint32_t rest = imp & ~prune;
#pragma unroll
            for (uint32_t k = 0; k < kWGather; ++k) {
                const unsigned long long m = __ballot(rest != 0u);
                croot[k] = 0u;
                cdv[k] = 0u;
                if (m) {
                    const uint32_t l0 = (uint32_t)__ffsll(m) - 1;
                    const uint32_t bits = uniform32(__shfl(rest, (int)l0));
                    const uint32_t q = (uint32_t)__ffs(bits) - 1;
                    // dv[q] of lane l0 (unrolled select: no runtime index into dv)
                    uint32_t myd = 0;
#pragma unroll
                    for (uint32_t qq = 0; qq < 8; ++qq)
                        if (qq == q) myd = dv[qq];
                    croot[k] = 8 * l0 + q;
                    cdv[k] = uniform32(__shfl(myd, (int)l0));
                    if (lane == l0) rest &= rest - 1;
                    ++ncand_t;
                }
            }
        }
        if (ncand_t && j_gather < len) {
            auto gather_pass = [&](uint32_t d, uint32_t h, uint32_t hm) {
                uint32_t x[kWGather];
#pragma unroll
                for (uint32_t k = 0; k < kWGather; ++k) {
                    x[k] = kWInf;
                    if (k < ncand_t && !((cpr >> k) & 1u) && (uint64_t)d + hm <= cdv[k])
                        x[k] = dcw[(uint64_t)h * kWB + croot[k]];
                }
#pragma unroll
                for (uint32_t k = 0; k < kWGather; ++k)
                    if (__ballot(x[k] != kWInf && (uint64_t)d + x[k] <= cdv[k]) != 0ull) cpr |= 1u << k;
            };
            const u64 gb = base + j_gather;
            const uint32_t glen = len - j_gather;
            u64 la = lane < glen ? S.pool[gb + lane] : 0ull, lb = lane + 64 < glen ? S.pool[gb + lane + 64] : 0ull;
            for (uint32_t j0 = 0; j0 < glen; j0 += 128) {
                const uint32_t ja = j0 + lane, jb = j0 + 64 + lane;
                const u64 ca = la, cbl = lb;
                la = ja + 128 < glen ? S.pool[gb + ja + 128] : 0ull;
                lb = jb + 128 < glen ? S.pool[gb + jb + 128] : 0ull;
                // (past the list: d = hmin = INF, so nothing is read)
                uint32_t da = kWInf, db = kWInf, sa = 0, sb = 0, ma = kWInf, mb = kWInf;
                if (ja < glen) {
                    sa = (uint32_t)(ca >> 32);
                    da = (uint32_t)ca;
                    ma = hmin[sa];
                }
                if (jb < glen) {
                    sb = (uint32_t)(cbl >> 32);
                    db = (uint32_t)cbl;
                    mb = hmin[sb];
                }
                gather_pass(da, sa, ma);
                gather_pass(db, sb, mb);
                if (cpr == (1u << ncand_t) - 1u) break;
            }
#pragma unroll
            for (uint32_t k = 0; k < kWGather; ++k)
                if (((cpr >> k) & 1u) && lane == (croot[k] >> 3)) prune |= 1u << (croot[k] & 7u);
        }
        const uint32_t live = imp & ~prune;
        n_pruned += __popc(imp & prune);
        // pruned-at-current-value bits: set for pruned, cleared for live
        if (imp) Pb[lane] = (uint8_t)((Pb[lane] & ~imp) | (imp & prune));
        const u64 any = __ballot(live != 0);
        if (!any) continue;
        // Next-frontier slots come from per-wave reservations, capped by the
        // tails this wave has left in the launch (grid-stride).
        const uint32_t rem = (nimp - 1 - i) / nwaves + 1;
        if (a_left == 0) {
            const uint32_t grab = rem < kWAlistGrab ? rem : kWAlistGrab;
            uint32_t b0 = 0;
            if (lane == 0) b0 = atomicAdd(&C->a_top[p ^ 1u], grab);
            a_next = __shfl(b0, 0);
            a_left = grab;
        }
        const uint32_t k = a_next++;
        --a_left;
        ++my_act;
        reinterpret_cast<uint8_t *>(ag + (uint64_t)k * kWW)[lane] = (uint8_t)live;
        uint32_t wmask = 0;
#pragma unroll
        for (int w = 0; w < (int)kWW; ++w)
            if ((any >> (8 * w)) & 0xFFull) wmask |= 1u << w;
        const uint32_t pdeg = prefix_above(off, adj, t, rs);
        if (lane == 0) {
            alist[k] = t;
            amask[k] = wmask;
            aend[k] = pdeg;
        }
        const uint32_t nch = (pdeg + kChunk - 1) / kChunk;
        if (nch > c_left) {
            for (uint32_t q = lane; q < c_left; q += 64) chunks[c_next + q] = kWChunkPad;
            const uint64_t guess = (uint64_t)nch * rem;  // descriptors left if t is typical
            const uint32_t cap = guess < kWChunkGrab ? (uint32_t)guess : kWChunkGrab;
            const uint32_t want = nch > cap ? nch : cap;
            uint32_t b0 = 0;
            if (lane == 0) b0 = atomicAdd(&C->n_chunk[p ^ 1u], want);
            c_next = __shfl(b0, 0);
            c_left = want;
        }
        for (uint32_t q = lane; q < nch; q += 64) chunks[c_next + q] = ((u64)(q * kChunk) << 32) | k;
        c_next += nch;
        c_left -= nch;
    }
    for (uint32_t q = lane; q < c_left; q += 64) chunks[c_next + q] = kWChunkPad;
    if (lane == 0 && my_act) atomicAdd(&C->n_act[p ^ 1u], my_act);
    n_pruned = wave_sum(n_pruned);
    if (lane == 0 && n_pruned) atomicAdd(&C->pruned, n_pruned);
}

// The batch roots' mutual entries, transposed for the filter: RT[h][r] =
// T[rs + r][h] and bit r of PT[h] = pruned bit h of root rs + r, so a witness
// h's column reads as one contiguous row.  One block per 64 x 64 tile (four
// waves, sixteen rows each); roots past the batch read as INF / unpruned.
__global__ __launch_bounds__(256) void k_w_roots_t(uint32_t rs, uint32_t rsz, const uint32_t *T, const u64 *P,
                                                   uint32_t *RT, u64 *PT)
{
    __shared__ uint32_t tile[64][65];
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t r0 = (blockIdx.x % (kWB / 64)) * 64, h0 = (blockIdx.x / (kWB / 64)) * 64;
    for (uint32_t i = wv; i < 64; i += 4) {
        const uint32_t r = r0 + i;
        tile[i][lane] = r < rsz ? T[(uint64_t)(rs + r) * kWB + h0 + lane] : kWInf;
    }
    __syncthreads();
    for (uint32_t j = wv; j < 64; j += 4) RT[(uint64_t)(h0 + j) * kWB + r0 + lane] = tile[lane][j];
    // Pruned bits: lane = root r0 + lane, ballot over the roots per witness.
    if (wv == 0) {
        const uint32_t r = r0 + lane;
        const u64 pw = r < rsz ? P[(uint64_t)(rs + r) * kWW + h0 / 64] : 0ull;
        u64 mine = 0;
        for (uint32_t h = 0; h < 64; ++h) {
            const u64 bits = __ballot((pw >> h) & 1ull);
            if (lane == h) mine = bits;
        }
        PT[(uint64_t)(h0 + lane) * kWW + r0 / 64] = mine;
    }
}

// filter_out_labels (WPARA:1100-1228): one wave per touched vertex; surviving
// pairs (finite, not pruned) are dropped if an earlier batch root h witnesses
// S(v,h) + S(r,h) <= S(v,r).  Reads T/P only, so every wave sees the same
// pre-filter state; the kept bitmap goes to I and the pool words the commit
// may reserve are summed into C->need.  kDirect: the roots' entries come from
// T/P rows directly (8 scattered reads per witness and lane, no transpose
// kernel; for batches that touch few vertices), else from RT/PT.
template <bool kDirect>
__global__ __launch_bounds__(256) void k_w_filter(uint32_t rs, uint32_t rsz, const uint32_t *tlist,
                                                  const uint32_t *T, const u64 *P, const uint32_t *RT, const u64 *PT,
                                                  u64 *I, WStore S, WCounters *C)
{
    __shared__ uint8_t sv[4][64];
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x >> 6;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t nt = C->n_touched;
    u64 n_lab = 0, n_filt = 0, n_need = 0;
    for (uint32_t i = wave; i < nt; i += nwaves) {
        const uint32_t v = tlist[i];
        const uint8_t pb = reinterpret_cast<const uint8_t *>(P + (uint64_t)v * kWW)[lane];
        uint32_t dv[8];
        uint32_t surv = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            dv[q] = T[(uint64_t)v * kWB + lane * 8 + q];
            if (dv[q] != kWInf && !((pb >> q) & 1u)) surv |= 1u << q;
        }
        sv[wib][lane] = (uint8_t)surv;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        uint32_t drop = 0;
        for (uint32_t w = 0; w < kWW; ++w) {
            u64 word = 0;
#pragma unroll
            for (int bb = 0; bb < 8; ++bb) word |= (u64)sv[wib][w * 8 + bb] << (8 * bb);
            while (word) {
                const uint32_t h = w * 64 + __ffsll(word) - 1;  // witness root offset
                word &= word - 1;
                if (h >= lane * 8 + 7) break;  // only roots r > h of this lane can use h
                const uint32_t dvh = T[(uint64_t)v * kWB + h];
                // Column h of the roots' entries, 8 consecutive roots per lane.
                uint32_t drh[8];
                uint32_t prh = 0;
                if constexpr (kDirect) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const uint32_t ro = lane * 8 + q;
                        drh[q] = kWInf;
                        if (((surv >> q) & 1u) && ro > h && ro < rsz) {
                            drh[q] = T[(uint64_t)(rs + ro) * kWB + h];
                            prh |= (uint32_t)((P[(uint64_t)(rs + ro) * kWW + h / 64] >> (h % 64)) & 1ull) << q;
                        }
                    }
                } else {
                    const uint4 *rt = reinterpret_cast<const uint4 *>(RT + (uint64_t)h * kWB + lane * 8);
                    const uint4 ra = rt[0], rb = rt[1];
                    drh[0] = ra.x, drh[1] = ra.y, drh[2] = ra.z, drh[3] = ra.w;
                    drh[4] = rb.x, drh[5] = rb.y, drh[6] = rb.z, drh[7] = rb.w;
                    prh = reinterpret_cast<const uint8_t *>(PT + (uint64_t)h * kWW)[lane];
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const uint32_t ro = lane * 8 + q;
                    if (!((surv >> q) & 1u) || ro <= h) continue;
                    if (drh[q] == kWInf || ((prh >> q) & 1u)) continue;
                    if ((uint64_t)dvh + drh[q] <= dv[q]) drop |= 1u << q;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        const uint32_t keep = surv & ~drop;
        n_filt += __popc(drop);
        // The improved bitmap is all zero once the rounds end; it carries the
        // kept pairs to the commit kernel.
        reinterpret_cast<uint8_t *>(I + (uint64_t)v * kWW)[lane] = (uint8_t)keep;
        const uint32_t total = (uint32_t)wave_sum((u64)__popc(keep));
        if (!total) continue;
        n_lab += total;
        if (lane == 0) n_need += reserve_words(S.len[v], S.cap[v], total);
    }
    n_filt = wave_sum(n_filt);
    if (lane == 0) {
        if (n_lab) atomicAdd(&C->labels, n_lab);
        if (n_filt) atomicAdd(&C->filtered, n_filt);
        if (n_need) atomicAdd(&C->need, n_need);
    }
}

// update_index (WPARA:961-990): append the kept pairs of each touched vertex,
// hubs ascending (batches are committed in root order and each lane's roots
// follow the lower lanes'), then clear the carried bitmap.
// update_index (WPARA:961-990) and reset_tables (WPARA:918-956) in one pass
// over the touched vertices: each vertex's wave appends its kept pairs and
// then empties its T / P rows; the batch roots (touched-list entries
// 0..rsz-1) first clear the root-distance entries their labels wrote.
__global__ __launch_bounds__(256) void k_w_commit(uint32_t rs, uint32_t rsz, const uint32_t *tlist, uint32_t *T, u64 *P,
                                                  u64 *I, uint32_t *hmin, uint32_t *dcw, WStore S, WCounters *C)
{
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t nt = C->n_touched;
    for (uint32_t i = wave; i < nt; i += nwaves) {
        const uint32_t v = tlist[i];
        if (i < rsz) {
            // root offset i: its labels are those k_w_init entered
            const uint32_t len = S.len[v];
            const u64 base = S.base[v];
            for (uint32_t j = lane; j < len; j += 64) {
                const uint32_t h = (uint32_t)(S.pool[base + j] >> 32);
                dcw[(uint64_t)h * kWB + i] = kWInf;
                hmin[h] = kWInf;
            }
        }
        uint8_t *Kb = reinterpret_cast<uint8_t *>(I + (uint64_t)v * kWW);
        const uint32_t keep = Kb[lane];
        const uint32_t cnt = __popc(keep);
        uint32_t incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(incl, o);
            if (lane >= (uint32_t)o) incl += y;
        }
        const uint32_t total = __shfl(incl, 63);
        if (total) {
            if (keep) Kb[lane] = 0;
            const uint32_t len = S.len[v];
            const u64 nb = store_reserve(S, C, v, len, total);
            if (nb != ~0ull) {
                uint32_t pos = len + incl - cnt;
                for (uint32_t kb = keep; kb; kb &= kb - 1) {
                    const uint32_t q = __ffs(kb) - 1;
                    S.pool[nb + pos++] = ((u64)(rs + lane * 8 + q) << 32) | T[(uint64_t)v * kWB + lane * 8 + q];
                }
                if (lane == 0) S.len[v] = len + total;
            }
        }
        // This lane's 8 entries were read above; the row goes back to INF.
        uint4 *row = reinterpret_cast<uint4 *>(T + (uint64_t)v * kWB + lane * 8);
        row[0] = row[1] = make_uint4(kWInf, kWInf, kWInf, kWInf);
        if (lane < kWW) P[(uint64_t)v * kWW + lane] = 0ull;
    }
}

__global__ void k_w_fill(uint32_t *p, uint64_t count, uint32_t val)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = val;
}

__global__ void k_w_zero64(u64 *p, uint64_t count)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = 0ull;
}

// Copies list v (len[v] words at pool[base[v]]) to dst[dst_off[v]].
__global__ __launch_bounds__(256) void k_w_move(uint32_t n, const u64 *pool, const u64 *base, const uint32_t *len,
                                                const uint64_t *dst_off, u64 *dst)
{
    const uint32_t lane = lane_id();
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < n; v += nw) {
        const uint64_t o = dst_off[v], b = base[v];
        const uint32_t L = len[v];
        for (uint32_t j = lane; j < L; j += 64) dst[o + j] = pool[b + j];
    }
}

__global__ void k_w_split(uint64_t total, const u64 *lab, uint32_t *hub, uint32_t *dist)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        hub[i] = (uint32_t)(lab[i] >> 32);
        dist[i] = (uint32_t)lab[i];
    }
}

}  // namespace

pado_gpu_status build_weighted(pado_gpu_graph *g, pado_gpu_index *idx, uint32_t batch_req)
{
    hipStream_t st = g->ctx->stream;
    const uint32_t n = g->n;
    const uint32_t batch = (batch_req == 0 || batch_req > kWB) ? kWB : batch_req;
    DevBuf<uint32_t> T, tmark, tlist, ilist, alist, amask, aend, hmin, dcw, len, cap;
    DevBuf<u64> P, I, ag, chunks, pool, base, imark;
    DevBuf<uint32_t> RT;  // batch roots' entries, transposed for the filter
    DevBuf<u64> PT;
    DevBuf<WCounters> ctr;
    PADO_TRY(T.alloc((uint64_t)n * kWB));
    PADO_TRY(RT.alloc((uint64_t)kWB * kWB));
    PADO_TRY(PT.alloc((uint64_t)kWB * kWW));
    PADO_TRY(P.alloc((uint64_t)n * kWW));
    PADO_TRY(I.alloc((uint64_t)n * kWW));
    // Per-wave reservations in k_w_check leave holes: up to kWAlistGrab
    // active-list slots per wave beyond the real ones, and padded
    // descriptors fewer than the real ones plus kWChunkGrab per wave.
    const uint64_t check_waves = 4ull * kAddGrid;
    const uint64_t alist_cap = (uint64_t)n + kWAlistGrab * check_waves;
    PADO_TRY(ag.alloc(alist_cap * kWW));
    PADO_TRY(chunks.alloc(2 * (g->m / kChunk + (uint64_t)n) + kWChunkGrab * check_waves + 64));
    PADO_TRY(tmark.alloc(n));
    PADO_TRY(tlist.alloc(n));
    PADO_TRY(imark.alloc(n));
    PADO_TRY(ilist.alloc(n));
    PADO_TRY(alist.alloc(alist_cap));
    PADO_TRY(amask.alloc(alist_cap));
    PADO_TRY(aend.alloc(alist_cap));
    PADO_TRY(hmin.alloc(n));
    PADO_TRY(len.alloc(n));
    PADO_TRY(cap.alloc(n));
    PADO_TRY(base.alloc(n));
    PADO_TRY(ctr.alloc(1));
    u64 pool_size = std::max<u64>(1ull << 22, 32ull * n);
    PADO_TRY(pool.alloc(pool_size));
    uint32_t dc_slots = std::min<uint32_t>(n, 1u << 14);
    PADO_TRY(dcw.alloc((uint64_t)dc_slots * kWB));
    const uint32_t fg = 4096;
    hipLaunchKernelGGL(k_w_fill, dim3(fg), dim3(256), 0, st, dcw.p, (uint64_t)dc_slots * kWB, kWInf);
    hipLaunchKernelGGL(k_w_fill, dim3(fg), dim3(256), 0, st, T.p, (uint64_t)n * kWB, kWInf);
    hipLaunchKernelGGL(k_w_zero64, dim3(fg), dim3(256), 0, st, P.p, (uint64_t)n * kWW);
    hipLaunchKernelGGL(k_w_zero64, dim3(fg), dim3(256), 0, st, I.p, (uint64_t)n * kWW);
    hipLaunchKernelGGL(k_w_fill, dim3(fg), dim3(256), 0, st, tmark.p, (uint64_t)n, 0u);
    hipLaunchKernelGGL(k_w_zero64, dim3(fg), dim3(256), 0, st, imark.p, (uint64_t)n);
    hipLaunchKernelGGL(k_w_fill, dim3(fg), dim3(256), 0, st, hmin.p, (uint64_t)n, kWInf);
    hipLaunchKernelGGL(k_w_fill, dim3(fg), dim3(256), 0, st, len.p, (uint64_t)n, 0u);
    hipLaunchKernelGGL(k_w_fill, dim3(fg), dim3(256), 0, st, cap.p, (uint64_t)n, 0u);
    hipLaunchKernelGGL(k_w_zero64, dim3(fg), dim3(256), 0, st, base.p, (uint64_t)n);
    PADO_HIP(hipMemsetAsync(ctr.p, 0, sizeof(WCounters), st));
    WCounters *h = nullptr;
    PADO_HIP(hipHostMalloc((void **)&h, sizeof(WCounters), hipHostMallocDefault));
    std::memset(h, 0, sizeof(WCounters));
    struct PinnedFree {
        WCounters *p;
        ~PinnedFree()
        {
            std::lock_guard<std::recursive_mutex> g(capture_mutex());
            (void)hipHostFree(p);
        }
    } pin{h};

    WStore S{pool.p, base.p, len.p, cap.p, pool_size};
    // The first batch takes a quarter of the roots: nothing prunes its
    // searches, and its labels prune the next batch's (measured on weighted
    // R-MAT 20 / 22: 1.04 -> 0.95 s / 6.5 -> 6.3 s).  Any split of the rank
    // order into batches yields the same label set.
    const uint32_t first = std::max<uint32_t>(1u, batch / 4);
    uint32_t prev_rounds = 1;
    uint32_t rs_next = 0;
    for (uint32_t b = 0; rs_next < n; ++b) {
        const uint32_t rs = rs_next, rsz = std::min(b ? batch : first, n - rs), gen_batch = b + 1;
        rs_next = rs + rsz;
        const uint32_t rb = (rsz * 64 + 255) / 256;
        hipLaunchKernelGGL(k_w_reset_counters, dim3(1), dim3(1), 0, st, ctr.p, rsz);
        // The hubs of the batch roots' labels are ids below the batch end, so
        // that many table rows always suffice (no device round trip).
        if (rs + rsz > dc_slots) {
            dc_slots = std::min<uint32_t>(n, std::max<uint32_t>(rs + rsz, dc_slots * 2));
            PADO_TRY(dcw.alloc((uint64_t)dc_slots * kWB));
            hipLaunchKernelGGL(k_w_fill, dim3(fg), dim3(256), 0, st, dcw.p, (uint64_t)dc_slots * kWB, kWInf);
        }
        hipLaunchKernelGGL(k_w_init, dim3(rb), dim3(256), 0, st, rs, rsz, gen_batch, g->off.p, g->adj.p, T.p, tmark.p,
                           tlist.p, alist.p, ag.p, amask.p, aend.p, chunks.p, hmin.p, dcw.p, S, ctr.p);
        // Rounds per host check: the previous batch's round count + 1 first
        // (enough when this batch takes no more), then blocks of 2^j that
        // double while a batch runs long.
        // Round 1's frontier is the batch's roots (never empty); the counters
        // seen last (after the previous batch's filter) hold every round so far.
        uint32_t live = rsz;
        const uint64_t rounds_before = h->rounds;
        uint32_t launched = 0;
        int j = 0;
        while ((1u << j) < prev_rounds + 1 && j < 6) ++j;
        while (live) {
            const uint32_t blk = launched ? 1u << j : std::min<uint32_t>(prev_rounds + 1, 64u);
            if (launched + blk >= (1u << 24)) {
                set_error("a weighted batch needs more than 2^24 Bellman-Ford rounds");
                return PADO_GPU_ERR_DOMAIN;
            }
            for (uint32_t u = 0; u < blk; ++u) {
                const uint32_t round = launched + 1 + u;
                hipLaunchKernelGGL(k_w_push, dim3(kPushGrid), dim3(256), 0, st, round, rs, rsz, gen_batch, g->off.p,
                                   g->adj.p, g->w.p, alist.p, ag.p, amask.p, aend.p, chunks.p, T.p, I.p, tmark.p,
                                   tlist.p, imark.p, ilist.p, S, ctr.p);
                hipLaunchKernelGGL(k_w_check, dim3(kAddGrid), dim3(256), 0, st, round, rs, ilist.p, g->off.p, g->adj.p,
                                   T.p, I.p, P.p, S, hmin.p, dcw.p, alist.p, ag.p, amask.p, aend.p,
                                   chunks.p, ctr.p);
            }
            launched += blk;
            PADO_HIP(hipMemcpyAsync(h, ctr.p, sizeof(WCounters), hipMemcpyDeviceToHost, st));
            PADO_HIP(hipStreamSynchronize(st));
            live = h->n_act[(launched + 1) & 1u];  // frontier of the next round
            if (j < 6) ++j;
        }
        prev_rounds = (uint32_t)(h->rounds - rounds_before);
        const u64 labels_before = h->labels;
        // Few touched vertices: the filter reads the roots' entries in place
        // instead of transposing them first.
        if (h->n_touched <= kWDirectFilter) {
            hipLaunchKernelGGL(k_w_filter<true>, dim3(2048), dim3(256), 0, st, rs, rsz, tlist.p, T.p, P.p, RT.p, PT.p,
                               I.p, S, ctr.p);
        } else {
            hipLaunchKernelGGL(k_w_roots_t, dim3((kWB / 64) * (kWB / 64)), dim3(256), 0, st, rs, rsz, T.p, P.p, RT.p,
                               PT.p);
            hipLaunchKernelGGL(k_w_filter<false>, dim3(2048), dim3(256), 0, st, rs, rsz, tlist.p, T.p, P.p, RT.p,
                               PT.p, I.p, S, ctr.p);
        }
        // The touched vertices are known once the rounds end: when even the
        // reservation bound fits the pool, the commit needs no host check.
        bool grow = h->pool_top + h->need_bound > pool_size;
        if (grow) {
            PADO_HIP(hipMemcpyAsync(h, ctr.p, sizeof(WCounters), hipMemcpyDeviceToHost, st));
            PADO_HIP(hipStreamSynchronize(st));
            grow = h->pool_top + h->need > pool_size;
        }
        if (grow) {
            // Compact every list into a larger pool.  After compaction a list of
            // length L has capacity L + L/4 + 4, so the commit can reserve at
            // most 2.5 * live + new + 8 * touched words.
            const u64 added = h->labels - labels_before;
            std::vector<uint32_t> hl(n);
            PADO_HIP(copy_sync(hl.data(), len.p, 4ull * n, hipMemcpyDeviceToHost, st));
            std::vector<uint64_t> no(n + 1, 0);
            std::vector<uint32_t> nc(n);
            for (uint32_t v = 0; v < n; ++v) {
                nc[v] = hl[v] + (hl[v] >> 2) + 4;
                no[v + 1] = no[v] + nc[v];
            }
            const u64 nsize = std::max<u64>(pool_size * 2, 4 * no[n] + 2 * added + 8ull * h->n_touched + 1024);
            DevBuf<u64> np;
            DevBuf<uint64_t> dno;
            PADO_TRY(np.alloc(nsize));
            PADO_TRY(dno.alloc((uint64_t)n + 1));
            PADO_HIP(copy_sync(dno.p, no.data(), 8ull * (n + 1), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_w_move, dim3(grid_for((uint64_t)n * 64, 256, 8192)), dim3(256), 0, st, n, pool.p,
                               base.p, len.p, dno.p, np.p);
            PADO_HIP(hipStreamSynchronize(st));
            std::vector<u64> nb64(no.begin(), no.end() - 1);
            PADO_HIP(copy_sync(base.p, nb64.data(), 8ull * n, hipMemcpyHostToDevice, st));
            PADO_HIP(copy_sync(cap.p, nc.data(), 4ull * n, hipMemcpyHostToDevice, st));
            std::swap(pool.p, np.p);
            std::swap(pool.n, np.n);
            pool_size = nsize;
            S.pool = pool.p;
            S.pool_size = pool_size;
            const u64 top = no[n];
            PADO_HIP(copy_sync(&ctr.p->pool_top, &top, 8, hipMemcpyHostToDevice, st));
        }
        hipLaunchKernelGGL(k_w_commit, dim3(2048), dim3(256), 0, st, rs, rsz, tlist.p, T.p, P.p, I.p, hmin.p, dcw.p,
                           S, ctr.p);
    }
    PADO_HIP(hipMemcpyAsync(h, ctr.p, sizeof(WCounters), hipMemcpyDeviceToHost, st));
    PADO_HIP(hipStreamSynchronize(st));
    if (h->overflow) {
        set_error("weighted label pool exhausted");
        return PADO_GPU_ERR_NOMEM;
    }
    idx->iterations = h->rounds;
    idx->counters.labels = h->labels;
    idx->counters.pushes = h->relax;
    idx->counters.cands = h->improved;
    idx->counters.dq_pruned = h->pruned;
    idx->counters.bp_pruned = h->filtered;

    // Canonical CSR.  Every list is already sorted by hub: batches commit in
    // root order and k_w_commit writes a batch's roots in ascending order, as
    // the reference's update_index appends them (WPARA:961-990).
    PADO_TRY(idx->loff.alloc((uint64_t)n + 1));
    PADO_TRY(scan_u32_to_u64(len.p, idx->loff.p, n, st));
    uint64_t total = 0;
    PADO_HIP(copy_sync(&total, idx->loff.p + n, 8, hipMemcpyDeviceToHost, st));
    idx->num_labels = total;
    DevBuf<u64> packed;
    PADO_TRY(packed.alloc(total));
    PADO_TRY(idx->whub.alloc(total));
    PADO_TRY(idx->wdist.alloc(total));
    hipLaunchKernelGGL(k_w_move, dim3(grid_for((uint64_t)n * 64, 256, 8192)), dim3(256), 0, st, n, pool.p, base.p,
                       len.p, idx->loff.p, packed.p);
    hipLaunchKernelGGL(k_w_split, dim3(grid_for(total, 256, 8192)), dim3(256), 0, st, total, packed.p, idx->whub.p,
                       idx->wdist.p);
    PADO_HIP(hipStreamSynchronize(st));
    idx->weighted = 1;
    return PADO_GPU_OK;
}

}  // namespace pado
