// batch.hip -- batched pruned labeling (the reference's batch_process) on gfx950.
//
// Reference (includes/pado_unw_unv_para.201912311607.parallel_bp_labeling.h):
//   initialize            :848-1080   roots -> active, self labels, dist_matrix
//   push_labels           :1084-1211  indicator test-and-set, BP check, candidates
//   distance_query        :1218-1270  2-hop check against labels of earlier iterations
//   insert_label_only     :1276-1291  append, root rows of dist_matrix
//   update_label_indices  :1295-1322
//   reset_at_end          :1328-1410
//   batch_process         :1413-1736
//   construct             :1740-1836  batches of BATCH_SIZE roots in rank order
//
// Semantics kept exactly: a (tail, root) pair is offered once per batch (the
// indicator is set BEFORE the BP check, PARA:1147-1153); the BP check uses the
// reference's unsigned `td - 2 <= iter` (slots with td < 2 never prune,
// PARA:1169-1170); the distance query sees v's labels with dist < iter and
// the batch roots' labels, hubs h < root only (PARA:1245-1260); pushes come
// from the last distance group only and stop at tails < roots_start.  Every
// decision is a set operation, so the label sets are independent of the
// schedule (the reference's own OpenMP order is arbitrary too).
//
// GPU layout (HBM):
//   ind[n][16] u64   indicator bitmap, 1024 roots per vertex row (rows of the
//                    vertices touched in a batch are cleared at its end)
//   cand[n][16] u64  candidate bitmap of the current iteration
//   ag[A][16] u64    last distance group of each active vertex, as a bitmap
//   chunks[]         16-edge push work descriptors (active index, edge start)
//   pool[] u32       label store: per-vertex growable arrays of packed
//                    (hub << 8 | dist) labels, grown fourfold (at least 32)
//                    on overflow by a bump allocator (one wave owns one
//                    vertex's array); rec[n] = (base, len, cap), 16 B each
//   lmask[n][16] u64 bit-parallel level masks (bit i of word k < 8: slot i at
//                    distance k), nonempty-set masks (words 8, 9: slots with
//                    S^-1 / S^0 nonempty) and vmark (word 10: ((batch << 16 |
//                    iteration) << 1 | candidate) of the last push into the
//                    vertex, raised by atomicMax); rmask / rset: the batch
//                    roots' mask words and set pairs, word- / slot-major
//   hinfo[n] 32 B    per hub: lane masks of the batch roots at distance 1, 2,
//                    3 (bit l: roots [16l, 16l+16)) and 0xFFFFFF00 | min
//                    distance from any batch root
//   dc[h][1024] u8   root-distance table: dc[h][r - rs] = dist(root r, hub h),
//                    the transposed form of dist_matrix, rows by hub id
// Kernels per iteration: k_push (edge-parallel, one edge per lane, 16-edge
// chunks) and k_add (wave per candidate vertex: distance query + label
// append + next active entry).  The host enqueues blocks of iterations and
// checks the frontier count once per block, running one batch ahead behind a
// device-side gate.
#include <algorithm>
#include <functional>
#include <map>

#include "internal.hpp"

namespace pado {
namespace {

__device__ inline uint32_t lane_id() { return threadIdx.x & 63; }

__device__ inline uint32_t wave_sum_u32(uint32_t x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Push descriptor written into reserved chunk slots that carry no work: its
// edge offset is past every adjacency list, so k_push skips it.
constexpr unsigned long long kChunkPad = 0xFFFFFFF0ull << 32;

// k_add reserves next-frontier push descriptors per wave, up to kChunkGrab
// per atomic on the shared counter (the unused rest of a reservation is
// padding).
constexpr uint32_t kChunkGrab = 64;

// Number of neighbours of v with id >= rs: the adjacency is sorted
// descending, so they form a prefix (PARA:1121-1123 returns at the first tail
// below roots_start).  Wave-cooperative 64-ary search.
__device__ inline uint32_t push_degree_range(const uint32_t *adj, uint64_t base, uint64_t end, uint32_t rs)
{
    const uint32_t lane = lane_id();
    uint32_t lo = 0, hi = (uint32_t)(end - base);
    while (hi - lo > 64) {
        uint32_t step = (hi - lo + 63) / 64;
        uint32_t pos = lo + lane * step;
        bool p = pos < hi && adj[base + pos] >= rs;
        uint32_t c = (uint32_t)__popcll(__ballot(p));
        if (c == 0) {
            hi = lo;
        } else {
            uint32_t nlo = lo + (c - 1) * step + 1;
            uint32_t nhi = lo + c * step;
            lo = nlo;
            if (nhi < hi) hi = nhi;
        }
    }
    uint32_t pos = lo + lane;
    bool p = pos < hi && adj[base + pos] >= rs;
    return lo + (uint32_t)__popcll(__ballot(p));
}

__device__ inline uint32_t push_degree(const uint64_t *off, const uint32_t *adj, uint32_t v, uint32_t rs)
{
    return push_degree_range(adj, off[v], off[v + 1], rs);
}

// Bit-parallel check of PARA:1164-1186 with the reference's unsigned
// arithmetic.  Per slot the reference tests `td - 2 <= iter` (unsigned, so
// td < 2 never passes), subtracts 2 or 1 by the S^-1 / S^0 intersections and
// prunes if the result is <= iter.  As the sets can only lower td, that is:
// prune if 2 <= td <= iter; prune if td == iter + 1 and S^-1_r meets S^-1_t
// or S^0_t (or S^0_r meets S^-1_t); prune if td == iter + 2 and S^-1_r meets
// S^-1_t.  The slots are an OR, so the check runs in two phases: this scan
// of the distance lines (no memory traffic) settles the first case and
// returns the slots at iter + 1 / iter + 2 (`cand`, with `one` marking
// iter + 1) whose set pairs the second phase loads.
// Bytes of x that are not 0xFF: bit 7 of each such byte set.
__device__ inline uint32_t bytes_not_ff(uint32_t x)
{
    const uint32_t y = ~x;  // a byte of y is nonzero iff x's byte is not 0xFF
    return (((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y) & 0x80808080u;
}

template <bool W>
__device__ inline bool bp_dist_scan(const uint32_t (&dr)[13], const uint32_t (&dt)[13], uint32_t iter,
                                    unsigned long long &cand, unsigned long long &one)
{
    // Four slots per 32-bit word, as 16-bit lanes: lo = slots 4w and 4w+2,
    // hi = slots 4w+1 and 4w+3 (td <= 510).  (0x8000 | c) - td keeps bit 15
    // of a lane set iff td <= c (c < 0x8000), and no borrow crosses lanes.
    // Bytes 50..63 of the distance line are padding: slots 50 and 51 are
    // masked off.  In the wide domain iter can pass 255, so a slot with INF
    // (255) on either side is masked off explicitly; in the reference's domain
    // such a slot takes part as the reference's arithmetic has it.
    const uint32_t c_it = 0x80008000u | (iter * 0x00010001u);
    const uint32_t c_it2 = 0x80008000u | ((iter + 2u) * 0x00010001u);
    const uint32_t c_1 = 0x80018001u;
    uint32_t hit = 0;
    cand = 0;
    one = 0;
#pragma unroll
    for (int w = 0; w < 13; ++w) {
        uint32_t valid = w == 12 ? 0x00008000u : 0x80008000u;
        uint32_t valid_hi = valid;
        if (W) {
            // Bit 7 / 23 (bytes 0 / 2) move to lane bits 15 / 31 for lo; bytes
            // 1 / 3 already sit at bits 15 / 31 for hi.
            const uint32_t ok = bytes_not_ff(dr[w]) & bytes_not_ff(dt[w]);
            valid_hi = valid & ok;
            valid &= (ok & 0x00800080u) << 8;
        }
        const uint32_t lo = (dr[w] & 0x00FF00FFu) + (dt[w] & 0x00FF00FFu);
        const uint32_t hi = ((dr[w] >> 8) & 0x00FF00FFu) + ((dt[w] >> 8) & 0x00FF00FFu);
        const uint32_t lo_it = c_it - lo, hi_it = c_it - hi;
        hit |= ((lo_it & ~(c_1 - lo)) & valid) | ((hi_it & ~(c_1 - hi)) & valid_hi);  // 2 <= td <= iter
        const uint32_t lo_n = (c_it2 - lo) & ~lo_it & valid, hi_n = (c_it2 - hi) & ~hi_it & valid_hi;
        const uint32_t lo_1 = (c_it + 0x00010001u - lo) & ~lo_it, hi_1 = (c_it + 0x00010001u - hi) & ~hi_it;
        const uint32_t nib = ((lo_n >> 15) & 1u) | ((hi_n >> 14) & 2u) | ((lo_n >> 29) & 4u) | ((hi_n >> 28) & 8u);
        const uint32_t o = ((lo_1 >> 15) & 1u) | ((hi_1 >> 14) & 2u) | ((lo_1 >> 29) & 4u) | ((hi_1 >> 28) & 8u);
        cand |= (unsigned long long)nib << (4 * w);
        one |= (unsigned long long)(o & nib) << (4 * w);
    }
    return hit != 0;
}

// Level masks of a vertex's bit-parallel distances: bit i of word k is set
// iff slot i is at distance k (k < 8; larger distances and INF are in no
// word).  For iter <= kMaskIters every td the check can act on is at most
// iter + 2 <= 7, so the scan above becomes a few 64-bit ANDs: td == a + b for
// the slots in L_r[a] & L_t[b].
constexpr uint32_t kMaskIters = 5;

// Words of a vertex's mask row: 8 level masks, then the slots whose S^-1 /
// S^0 set is nonempty, the touched mark, then padding to one 128-byte line.
constexpr uint32_t kLmWords = 16, kLmNz1 = 8, kLmNz0 = 9;
// Word 10 of the row is the vertex's touched mark (vmark, see push_body):
// kept in the row whose other words the bit-parallel check of a fresh pair
// reads anyway, so marking a tail touches no further cache line.
constexpr uint32_t kLmMark = 10;

// The candidate slots of bp_dist_scan / bp_mask_scan whose set test can
// prune at all (bp_set_hit): at td == iter + 2 both S^-1 sets must be
// nonempty, at td == iter + 1 the S^-1 set of one side and the S^-1 or S^0
// set of the other.  The rest are dropped before any set pair is loaded.
__device__ inline unsigned long long bp_nz_filter(unsigned long long cs, unsigned long long one,
                                                  unsigned long long nz1r, unsigned long long nz0r,
                                                  unsigned long long nz1t, unsigned long long nz0t)
{
    const unsigned long long two = cs & ~one, o1 = cs & one;
    return (two & nz1r & nz1t) | (o1 & ((nz1r & (nz1t | nz0t)) | (nz0r & nz1t)));
}

template <int IT>
__device__ inline bool bp_mask_scan_it(const unsigned long long *rm, uint32_t ro, const unsigned long long *lt,
                                       unsigned long long &cand, unsigned long long &one)
{
    // Levels above IT + 2 never matter: only those below are loaded.
    constexpr int L = IT + 3 < 8 ? IT + 3 : 8;
    unsigned long long lr[L], tl[L];
#pragma unroll
    for (int a = 0; a < L; ++a) {
        lr[a] = rm[a * kBatch + ro];
        tl[a] = lt[a];
    }
    unsigned long long hit = 0, e1 = 0, e2 = 0;
#pragma unroll
    for (int a = 0; a < L; ++a) {
#pragma unroll
        for (int b = 0; b < L; ++b) {
            const int td = a + b;
            if (td >= 2 && td <= IT)
                hit |= lr[a] & tl[b];
            else if (td == IT + 1)
                e1 |= lr[a] & tl[b];
            else if (td == IT + 2)
                e2 |= lr[a] & tl[b];
        }
    }
    cand = e1 | e2;
    one = e1;
    return hit != 0ull;
}

// Same contract as bp_dist_scan for 1 <= iter <= kMaskIters (wave-uniform):
// root offset ro's masks come from the batch's word-major root table rm
// (rm[k * kBatch + ro]: lanes holding consecutive roots read consecutive
// words), the tail's from its mask row lt.
__device__ inline bool bp_mask_scan(uint32_t iter, const unsigned long long *rm, uint32_t ro,
                                    const unsigned long long *lt, unsigned long long &cand,
                                    unsigned long long &one)
{
    switch (iter) {
    case 1: return bp_mask_scan_it<1>(rm, ro, lt, cand, one);
    case 2: return bp_mask_scan_it<2>(rm, ro, lt, cand, one);
    case 3: return bp_mask_scan_it<3>(rm, ro, lt, cand, one);
    case 4: return bp_mask_scan_it<4>(rm, ro, lt, cand, one);
    default: return bp_mask_scan_it<5>(rm, ro, lt, cand, one);
    }
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ inline uint32_t wave_incl_scan(uint32_t x)
{
    const uint32_t lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

// Orders LDS writes of some lanes of a wave before reads by other lanes of
// the same wave.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Position of the rank-th set bit of m (rank < popcount(m)).
__device__ inline uint32_t nth_set_bit(unsigned long long m, uint32_t rank)
{
    uint32_t pos = 0;
    uint32_t c = __popc((uint32_t)m);
    if (rank >= c) {
        rank -= c;
        m >>= 32;
        pos = 32;
    }
    uint32_t x = (uint32_t)m;
    c = __popc(x & 0xFFFFu);
    if (rank >= c) {
        rank -= c;
        x >>= 16;
        pos += 16;
    }
    c = __popc(x & 0xFFu);
    if (rank >= c) {
        rank -= c;
        x >>= 8;
        pos += 8;
    }
    c = __popc(x & 0xFu);
    if (rank >= c) {
        rank -= c;
        x >>= 4;
        pos += 4;
    }
    c = __popc(x & 0x3u);
    if (rank >= c) {
        rank -= c;
        x >>= 2;
        pos += 2;
    }
    if (rank >= (x & 1u)) pos += 1;
    return pos;
}

// Last lane o whose exclusive prefix ex[o] <= x (x below the total): the
// lane that owns flattened item x (it has at least one item).
__device__ inline uint32_t item_owner(const uint32_t *ex, uint32_t x)
{
    uint32_t o = 0;
#pragma unroll
    for (uint32_t sft = 32; sft; sft >>= 1)
        if (ex[o + sft] <= x) o += sft;
    return o;
}

// One set-pair test: slot i, td == iter + 2 (S^-1_r meets S^-1_t) or, with
// `one`, td == iter + 1 (also S^-1 against S^0 either way).
__device__ inline bool bp_set_hit(ulonglong2 r, ulonglong2 t, bool one)
{
    return (r.x & t.x) != 0ull || (one && ((r.x & t.y) | (r.y & t.x)) != 0ull);
}

// LDS scratch of one wave for the set phase.
struct SetScratch {
    unsigned long long *rem, *one;
    uint32_t *r, *t, *ex, *hit;
};

// Second phase of the bit-parallel check, across the 64 lanes of a wave
// (every lane calls it; rem = 0 for a lane with nothing to test).  Each lane
// first tests its two lowest candidate slots, which settles most pruned
// pairs; the (pair, slot) tests left over, mostly those of pairs that pass,
// are then spread over the wave 64 at a time, so a pair with dozens of
// candidate slots costs a few wave steps instead of dozens.  Root set pairs
// always come from the batch's slot-major table (800 KB, which stays in the
// L2 cache), tail set pairs from the vertex-major table (in the spread steps
// lanes hold slots of the same few tails, whose records they share).
// Returns whether the lane's pair is pruned.
__device__ inline bool bp_set_phase(unsigned long long rem, unsigned long long one, const ulonglong2 *rset,
                                    uint32_t ro, const ulonglong2 *bps, uint32_t t, const SetScratch &sc)
{
    const uint32_t lane = lane_id();
    bool pr = false;
    if (rem) {
        const uint32_t i0 = (uint32_t)__ffsll(rem) - 1;
        rem &= rem - 1;
        const bool two = rem != 0ull;
        const uint32_t i1 = two ? (uint32_t)__ffsll(rem) - 1 : i0;
        if (two) rem &= rem - 1;
        const ulonglong2 r0 = rset[i0 * kBatch + ro], t0 = bps[(uint64_t)t * kBP + i0];
        ulonglong2 r1 = make_ulonglong2(0ull, 0ull), t1 = r1;
        if (two) {
            r1 = rset[i1 * kBatch + ro];
            t1 = bps[(uint64_t)t * kBP + i1];
        }
        if (bp_set_hit(r0, t0, (one >> i0) & 1ull) || (two && bp_set_hit(r1, t1, (one >> i1) & 1ull))) {
            pr = true;
            rem = 0ull;
        }
    }
    const uint32_t cnt = (uint32_t)__popcll(rem);
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = __shfl(incl, 63);
    if (total == 0) return pr;
    sc.rem[lane] = rem;
    sc.one[lane] = one;
    sc.r[lane] = ro;
    sc.t[lane] = t;
    sc.ex[lane] = incl - cnt;
    sc.hit[lane] = 0u;
    wave_lds_sync();
    for (uint32_t x0 = 0; x0 < total; x0 += 64) {
        const uint32_t x = x0 + lane;
        if (x >= total) continue;
        const uint32_t o = item_owner(sc.ex, x);
        const unsigned long long ro_rem = sc.rem[o];
        const uint32_t i = nth_set_bit(ro_rem, x - sc.ex[o]);
        const ulonglong2 rr = rset[i * kBatch + sc.r[o]], tt = bps[(uint64_t)sc.t[o] * kBP + i];
        if (bp_set_hit(rr, tt, (sc.one[o] >> i) & 1ull)) atomicOr(&sc.hit[o], 1u);
    }
    wave_lds_sync();
    pr = pr || sc.hit[lane] != 0u;
    // The scratch is rewritten by the next call.
    wave_lds_sync();
    return pr;
}

__device__ inline void load_dline(const uint8_t *bpd, uint32_t v, uint32_t (&d)[13])
{
    const uint4 *p = reinterpret_cast<const uint4 *>(bpd + (uint64_t)v * kBpdStride);
    uint4 a = p[0], b = p[1], c = p[2], e = p[3];
    d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
    d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
    d[8] = c.x; d[9] = c.y; d[10] = c.z; d[11] = c.w;
    d[12] = e.x;
}

// A vertex's label array in the pool: one 16-byte record, so the three
// fields come in with one access.
struct alignas(16) LRec {
    unsigned long long base;  // first pool slot
    uint32_t len;             // labels held
    uint32_t cap;             // slots owned
};

struct LabelStore {
    uint32_t *pool;
    LRec *rec;
    unsigned long long pool_size;
};

// Each vertex's labels at distance 1, kept apart from the label arrays for
// the distance query of iteration 2 (the only one that needs them alone).  A
// distance-1 label comes from a root pushing along an edge in iteration 1, so
// its hub is a neighbour and a vertex has at most deg(v) of them: vertex v's
// list lives in the adjacency slots [off[v], off[v + 1]) of lab, packed like
// the pool, with len[v] entries.
struct HubList1 {
    uint32_t *lab;
    uint32_t *len;
};

// Summary of one row of the root-distance table (hub h): near[k] has bit l
// set iff some batch root in [16l, 16l+16) -- the roots lane l of k_add owns --
// is at distance k + 1 from h; info is 0xFFFFFF00 | the minimum distance of
// any batch root (0xFFFFFFFF when h is no hub of any of them).  A hub can
// prune a candidate root r of v only if d_r(h) <= iter - d_v(h), so for
// thresholds up to kNear the masks name the lanes whose row slices can
// matter, and the minimum rules out the rest.
constexpr uint32_t kNear = 3;
struct alignas(32) HubInfo {
    unsigned long long near[kNear];
    uint32_t info, pad;
};

__device__ inline void hub_info_clear(HubInfo &x)
{
    for (uint32_t k = 0; k < kNear; ++k) x.near[k] = 0ull;
    x.info = kEmpty;
}

__global__ void k_hub_info_init(HubInfo *p, uint64_t count)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        hub_info_clear(p[i]);
        p[i].pad = 0u;
    }
}

// Reserve room for `add` more labels of v (wave-uniform call).  Returns the
// (possibly new) base; on pool exhaustion sets the overflow flag and returns
// ~0ull.
__device__ inline unsigned long long store_reserve_at(LabelStore S, BuildCounters *C, uint32_t v, uint32_t len,
                                                      unsigned long long base, uint32_t cap, uint32_t add)
{
    const uint32_t lane = lane_id();
    if (len + add <= cap) return base;
    uint32_t ncap = cap * 4;  // few reallocations: each is an atomic on one shared counter
    if (ncap < len + add) ncap = len + add;
    if (ncap < 32) ncap = 32;
    unsigned long long nb = 0;
    if (lane == 0) nb = atomicAdd(&C->pool_top, (unsigned long long)ncap);
    nb = __shfl(nb, 0);
    if (nb + ncap > S.pool_size) {
        if (lane == 0) atomicOr(&C->overflow, 1u);
        return ~0ull;
    }
    for (uint32_t j = lane; j < len; j += 64) S.pool[nb + j] = S.pool[base + j];
    if (lane == 0) {
        S.rec[v].base = nb;
        S.rec[v].cap = ncap;
    }
    return nb;
}

__device__ inline unsigned long long store_reserve(LabelStore S, BuildCounters *C, uint32_t v, uint32_t len,
                                                   uint32_t add)
{
    const LRec r = S.rec[v];
    return store_reserve_at(S, C, v, len, r.base, r.cap, add);
}

__global__ void k_fill_u32(uint32_t *p, uint64_t count, uint32_t val)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = val;
}

__global__ void k_zero_u64(unsigned long long *p, uint64_t count)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = 0ull;
}

// Level masks (see bp_mask_scan) and nonempty-set masks (bp_nz_filter) of
// every vertex: one 128-byte row per vertex.
__global__ __launch_bounds__(256) void k_bp_masks(uint32_t n, const uint8_t *bpd, const ulonglong2 *bps,
                                                  unsigned long long *lmask)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        uint32_t d[13];
        load_dline(bpd, (uint32_t)v, d);
        unsigned long long m[kLmWords] = {};
#pragma unroll
        for (int i = 0; i < (int)kBP; ++i) {
            const uint32_t x = (d[i >> 2] >> (8 * (i & 3))) & 255u;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (x == (uint32_t)k) m[k] |= 1ull << i;
            const ulonglong2 st = bps[v * kBP + i];
            if (st.x) m[8] |= 1ull << i;
            if (st.y) m[9] |= 1ull << i;
        }
        uint4 *dst = reinterpret_cast<uint4 *>(lmask + v * kLmWords);
#pragma unroll
        for (int q = 0; q < (int)kLmWords / 2; ++q)
            dst[q] = make_uint4((uint32_t)m[2 * q], (uint32_t)(m[2 * q] >> 32), (uint32_t)m[2 * q + 1],
                                (uint32_t)(m[2 * q + 1] >> 32));
    }
}

// Root-distance table (dist_matrix, PARA:986-1013, transposed): row h holds
// dist(root r, hub h) at column r - rs for the hubs of the batch roots'
// labels, 255 elsewhere; hinfo[h] summarises the row (HubInfo).  Root
// offset ro's labels pool[base, base + len) are entered (fill) or their
// entries reset, by one wave; each lane handles four labels per step so the
// dependent loads of a step overlap.
template <bool W>
__device__ __forceinline__ void hub_rows(uint32_t fill, uint32_t ro, unsigned long long base, uint32_t len, LabelStore S,
                                         HubInfo *hinfo, typename Hop<W>::Dist *dc)
{
    constexpr uint32_t B = Hop<W>::kBits, INF = Hop<W>::kInf;
    const uint32_t lane = lane_id();
    for (uint32_t j0 = 0; j0 < len; j0 += 256) {
        uint32_t lab[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t j = j0 + 64 * u + lane;
            lab[u] = j < len ? S.pool[base + j] : kEmpty;
        }
        if (!fill) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (lab[u] == kEmpty) continue;
                const uint32_t h = lab[u] >> B;
                dc[(uint64_t)h * kBatch + ro] = (typename Hop<W>::Dist)~0u;
                hub_info_clear(hinfo[h]);
            }
            continue;
        }
        HubInfo hi[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (lab[u] != kEmpty) hi[u] = hinfo[lab[u] >> B];
        const unsigned long long bit = 1ull << (ro >> 4);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (lab[u] == kEmpty) continue;
            const uint32_t h = lab[u] >> B, d = lab[u] & INF;
            dc[(uint64_t)h * kBatch + ro] = (typename Hop<W>::Dist)d;
            // The minimum only falls and mask bits are only set during the
            // kernel, so a value read as already there (even a stale one)
            // makes the atomic a no-op.
            if (d < (hi[u].info & INF)) atomicMin(&hinfo[h].info, ~INF | d);
            if (d >= 1u && d <= kNear) {
                // (no runtime index into hi: it would live in scratch memory)
                const unsigned long long had = d == 1u ? hi[u].near[0] : (d == 2u ? hi[u].near[1] : hi[u].near[2]);
                if (!(had & bit)) atomicOr(&hinfo[h].near[d - 1], bit);
            }
        }
    }
}

// initialize (PARA:848-1080): non-BP roots become active with their own bit as
// last group, get the self label (r, 0), and set their own indicator bit;
// then their labels go into the root-distance table (PARA:986-1013).  One
// wave per root, four roots per block.  Root offset ro owns touched-list
// slot ro (k_batch_begin sets the count and the frontier size), so the shared
// counters left are the push-descriptor count, updated once per block, and
// the frontier's push-edge count.
template <bool W, uint32_t CH>
__device__ __forceinline__ void batch_init_body(uint32_t bid, uint32_t nblk, const uint8_t *used, const uint64_t *off, const uint32_t *adj,
                                                     LabelStore S, HubInfo *hinfo, typename Hop<W>::Dist *dc,
                                                     unsigned long long *ind, unsigned long long *vmark,
                                                     uint32_t *tlist, unsigned long long *abits, unsigned long long *ag,
                                                     uint32_t *amask, uint32_t *aend, unsigned long long *chunks,
                                                     const ulonglong2 *bps, ulonglong2 *rset,
                                                     const unsigned long long *lmask, unsigned long long *rmask,
                                                     BuildCounters *C)
{
    __shared__ uint32_t s_nch[16], s_pdeg[16];
    __shared__ uint32_t s_base;
    if (C->hold) return;
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t ro = (bid * blockDim.x + threadIdx.x) >> 6;
    const uint32_t rs = C->rs, rsz = C->rsz, gen_batch = C->gen_batch;
    const uint32_t r = rs + ro;
    // Every root of the batch (bit-parallel roots included) is on the
    // touched list; a row that stays empty is cleared at no harm.
    if (ro < rsz && lane == 0) {
        vmark[(uint64_t)r * kLmWords] = (unsigned long long)gen_batch << 17;
        tlist[ro] = r;
    }
    // Slot-major copy of the batch roots' bit-parallel sets and word-major
    // copy of their mask rows (push's bit-parallel check).
    if (ro < rsz && lane < kBP) rset[lane * kBatch + ro] = bps[(uint64_t)r * kBP + lane];
    if (ro < rsz && lane < kLmNz0 + 1) rmask[lane * kBatch + ro] = lmask[(uint64_t)r * kLmWords + lane];
    uint32_t pdeg = 0;
    // the root's labels before its self label (their array, even if it moves)
    uint32_t len0 = 0;
    unsigned long long base0 = 0;
    bool table = false;
    if (ro < rsz && !used[r]) {
        const LRec rec0 = S.rec[r];
        len0 = rec0.len;
        base0 = rec0.base;
        const uint32_t len = len0;
        const unsigned long long base = store_reserve_at(S, C, r, len, rec0.base, rec0.cap, 1);
        if (base != ~0ull) {
            table = true;
            if (lane == 0) {
                S.pool[base + len] = r << Hop<W>::kBits;  // (hub r, dist 0), PARA:932-949
                S.rec[r].len = len + 1;
                ind[(uint64_t)r * kWords + (ro >> 6)] |= 1ull << (ro & 63);
                amask[r] = 1u << (ro >> 6);
                // frontier of iteration 1: slot 1 of the bitmap
                atomicOr(&abits[C->abits_words + (r >> 6)], 1ull << (r & 63));
            }
            if (lane < kWords) ag[(uint64_t)r * kWords + lane] = (lane == (ro >> 6)) ? (1ull << (ro & 63)) : 0ull;
            pdeg = push_degree(off, adj, r, rs);
            if (lane == 0) aend[r] = pdeg;
        }
    }
    const uint32_t nch = (pdeg + CH - 1) / CH;
    if (lane == 0) {
        s_nch[wv] = nch;
        s_pdeg[wv] = pdeg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        unsigned long long tot_e = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
            tot += s_nch[w];
            tot_e += s_pdeg[w];
        }
        // The frontier of iteration i is counted in slot i & 1: iteration 1 -> 1.
        // One update per block of each shared counter (every root would
        // otherwise queue on the same two addresses).
        s_base = tot ? atomicAdd(&C->n_chunk[1], tot) : 0u;
        if (tot_e) atomicAdd(&C->n_pedge[1], tot_e);
    }
    __syncthreads();
    if (nch) {
        uint32_t c0 = s_base;
        for (uint32_t w = 0; w < wv; ++w) c0 += s_nch[w];
        for (uint32_t q = lane; q < nch; q += 64) chunks[c0 + q] = ((unsigned long long)(q * CH) << 32) | r;
    }
    if (table) {
        // root-distance entries: the earlier labels, then the self label
        hub_rows<W>(1u, ro, base0, len0, S, hinfo, dc);
        if (lane == 0) {
            dc[(uint64_t)r * kBatch + ro] = 0;
            atomicMin(&hinfo[r].info, ~Hop<W>::kInf);
        }
    }
}

template <bool W, uint32_t CH>
__global__ __launch_bounds__(256) void k_batch_init(const uint8_t *used, const uint64_t *off, const uint32_t *adj,
                                                     LabelStore S, HubInfo *hinfo, typename Hop<W>::Dist *dc,
                                                     unsigned long long *ind, unsigned long long *vmark,
                                                     uint32_t *tlist, unsigned long long *abits, unsigned long long *ag,
                                                     uint32_t *amask, uint32_t *aend, unsigned long long *chunks,
                                                     const ulonglong2 *bps, ulonglong2 *rset,
                                                     const unsigned long long *lmask, unsigned long long *rmask,
                                                     BuildCounters *C)
{
    batch_init_body<W, CH>(blockIdx.x, gridDim.x, used, off, adj, S, hinfo, dc, ind, vmark, tlist, abits, ag, amask, aend,
                           chunks, bps, rset, lmask, rmask, C);
}

// One root's entries of the root-distance table from the batch roots'
// offsets (one wave per root): with `fill` its labels are entered, without,
// the entries of the previous batch's roots (the labels they held at its end,
// which include every entry written during it) are reset.
template <bool W>
__device__ __forceinline__ void hub_table_body(uint32_t bid, uint32_t nblk, uint32_t fill, const uint8_t *used, LabelStore S, HubInfo *hinfo,
                                                   typename Hop<W>::Dist *dc, const BuildCounters *C)
{
    if (C->hold) return;
    const uint32_t ro = (bid * blockDim.x + threadIdx.x) >> 6;
    const uint32_t rs = C->rs, rsz = C->rsz;
    if (ro >= rsz) return;
    const uint32_t r = rs + ro;
    if (used[r]) return;
    const LRec rec = S.rec[r];
    hub_rows<W>(fill, ro, rec.base, rec.len, S, hinfo, dc);
}


constexpr uint32_t kPushWaves = 4;  // default waves per k_push block

// Default of pado_gpu_build_cfg::pull_ratio: an iteration runs in pull form
// when its frontier's push edges exceed this many thousandths of the
// adjacency entries of the vertices >= roots_start (what the pull form scans).
constexpr uint32_t kPullRatio = ~0u;

// The pull form of one push step (no reference counterpart: the reference
// always pushes, PARA:1486-1559).  Heavy iterations offer most tails many
// roots over many edges; pushing, every (edge, root word) pair is a random
// test-and-set of the tail's indicator row.  Pulling, each tail t >= rs
// gathers its own offers: the lanes that own t scan t's adjacency prefix
// (ids >= rs: exactly the push edges into t, adjacency being symmetric), OR
// the last-group rows of the neighbours on the frontier bitmap into a 1024-bit
// offer, and test-and-set t's indicator row once.  A (tail, root) pair passes
// iff the push form's edges would pass it, and the bit-parallel check, the
// candidate bits, the touched mark and the lists follow the same rules, so
// the iteration's outcome (a set) is the same.  Tails go four per wave, 16
// lanes each (lane 16 g + w: tail g, root word w); the fresh pairs of the four
// tails are then checked 64 at a time as in the push form, lane o of the
// wave owning word o & 15 of tail o >> 4.
template <bool W>
__device__ __forceinline__ void pull_step(uint32_t wave, uint32_t nwaves, uint32_t iter, uint32_t p, uint32_t rs,
                                          uint32_t n, unsigned long long gen_iter, uint32_t gen_batch,
                                          const uint64_t *off, const uint32_t *adj, const unsigned long long *bits,
                                          const unsigned long long *ag, unsigned long long *ind,
                                          unsigned long long *cand, unsigned long long *vmark, uint32_t *tlist,
                                          uint32_t *clist, const uint8_t *bpd, const unsigned long long *lmask,
                                          const unsigned long long *rmask, const ulonglong2 *bps,
                                          const ulonglong2 *rset, BuildCounters *C, uint32_t bp_prune,
                                          unsigned long long *s_nb, unsigned long long *s_keep, uint32_t *s_ex,
                                          uint32_t *s_t, const SetScratch &sc, uint32_t &n_push, uint32_t &n_pass,
                                          uint32_t &n_bp, uint32_t &n_cand, uint32_t &n_edge, uint32_t &n_tail)
{
    const uint32_t lane = lane_id(), g = lane >> 4, w = lane & 15;
    const bool use_masks = iter <= kMaskIters;
    for (uint32_t t0 = rs + 4 * wave; t0 < n; t0 += 4 * nwaves) {
        const uint32_t t = t0 + g;
        const bool tv = t < n;
        uint64_t e = tv ? off[t] : 0, e1 = tv ? off[t + 1] : 0;
        // roots below t (PARA:1134): offsets < lim; wm = those of word w
        const uint32_t lim = t - rs;
        const unsigned long long wm =
            lim >= (w + 1) * 64 ? ~0ull : (lim <= w * 64 ? 0ull : (1ull << (lim - w * 64)) - 1ull);
        unsigned long long acc = 0ull;
        bool more = tv && e < e1;
        while (__ballot(more) != 0ull) {
            // sixteen entries of the group's prefix per step
            const uint64_t j = e + w;
            uint32_t k = 0;
            bool inp = false, act = false;
            if (more && j < e1) {
                k = adj[j];
                inp = k >= rs;
            }
            if (inp) act = (bits[k >> 6] >> (k & 63)) & 1ull;
            n_edge += act ? 1u : 0u;
            uint32_t gm = (uint32_t)(__ballot(act) >> (16 * g)) & 0xFFFFu;
            // the last-group rows of the group's active neighbours, up to
            // four loads in flight per lane
            while (__ballot(gm != 0u) != 0ull) {
                unsigned long long x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t sl = gm ? (uint32_t)__ffs(gm) - 1u : 0u;
                    const uint32_t kk = __shfl(k, (int)(16 * g + sl));
                    x[u] = 0ull;
                    if (gm) {
                        x[u] = ag[(uint64_t)kk * kWords + w] & wm;
                        gm &= gm - 1u;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    acc |= x[u];
                    n_push += (uint32_t)__popcll(x[u]);
                }
            }
            // the prefix goes on past this step iff the group's last lane was in it
            const unsigned long long inm = __ballot(inp);
            more = more && ((inm >> (16 * g + 15)) & 1ull) && e + 16 < e1;
            e += 16;
        }
        // t's indicator row: one test-and-set per word (its only writer in
        // this iteration)
        unsigned long long nb = 0ull;
        if (acc) {
            const unsigned long long seen = ind[(uint64_t)t * kWords + w];
            nb = acc & ~seen;
            if (nb) ind[(uint64_t)t * kWords + w] = seen | acc;
        }
        const uint32_t cnt = (uint32_t)__popcll(nb);
        n_pass += cnt;
        const uint32_t incl = wave_incl_scan(cnt);
        const uint32_t total = __shfl(incl, 63);
        unsigned long long keep = 0ull;
        if (total) {
            s_nb[lane] = nb;
            s_keep[lane] = 0ull;
            s_ex[lane] = incl - cnt;
            s_t[lane] = t;
            wave_lds_sync();
            for (uint32_t x0 = 0; x0 < total; x0 += 64) {
                const uint32_t x = x0 + lane;
                const bool xa = x < total;
                uint32_t o = 0, b = 0, ro = 0, to = 0;
                unsigned long long cs = 0ull, one = 0ull;
                bool dist_pr = true;
                if (xa) {
                    o = item_owner(s_ex, x);
                    b = nth_set_bit(s_nb[o], x - s_ex[o]);
                    ro = (o & 15u) * 64 + b;
                    to = s_t[o];
                    const unsigned long long *lt = lmask + (uint64_t)to * kLmWords;
                    if (use_masks) {
                        dist_pr = bp_mask_scan(iter, rmask, ro, lt, cs, one);
                    } else {
                        uint32_t dr[13], dt[13];
                        load_dline(bpd, rs + ro, dr);
                        load_dline(bpd, to, dt);
                        dist_pr = bp_dist_scan<W>(dr, dt, iter, cs, one);
                    }
                    if (!bp_prune) {
                        dist_pr = false;
                        cs = 0ull;
                    }
                    if (!dist_pr && cs)
                        cs = bp_nz_filter(cs, one, rmask[kLmNz1 * kBatch + ro], rmask[kLmNz0 * kBatch + ro], lt[kLmNz1],
                                          lt[kLmNz0]);
                }
                const bool set_pr = bp_set_phase(dist_pr ? 0ull : cs, one, rset, ro, bps, to, sc);
                if (xa) {
                    if (dist_pr || set_pr)
                        ++n_bp;
                    else
                        atomicOr(&s_keep[o], 1ull << b);
                }
            }
            wave_lds_sync();
            keep = s_keep[lane];
            // the scratch is rewritten by the next step
            wave_lds_sync();
        }
        if (keep) {
            cand[(uint64_t)t * kWords + w] = keep;  // the row was empty (k_add clears what it takes)
            n_cand += (uint32_t)__popcll(keep);
        }
        const bool touched = ((__ballot(nb != 0ull) >> (16 * g)) & 0xFFFFull) != 0ull;
        const bool got = ((__ballot(keep != 0ull) >> (16 * g)) & 0xFFFFull) != 0ull;
        bool new_t = false, new_c = false;
        if (w == 0 && touched) {
            const unsigned long long want = (gen_iter << 1) | (got ? 1ull : 0ull);
            const unsigned long long prev = vmark[(uint64_t)t * kLmWords];
            if (prev < want) {
                vmark[(uint64_t)t * kLmWords] = want;
                new_t = (prev >> 17) != gen_batch;
                if ((prev >> 1) != gen_iter) ++n_tail;
                new_c = got;
            }
        }
        const uint32_t ti = wave_append(&C->n_touched, new_t);
        if (new_t) tlist[ti] = t;
        const uint32_t ci = wave_append(&C->n_cand[p], new_c);
        if (new_c) clist[ci] = t;
    }
}

// push (PARA:1486-1559 / push_labels PARA:1084-1211).  A wave takes 64 / CH
// work descriptors of CH edges each (one edge per lane; CH = kChunk and
// PW = kPushWaves waves per block by default, other instantiations exist
// for the autotuner) and walks the root words of its edges'
// last distance groups in step, four words at a time: each lane
// test-and-sets its indicator words, and the fresh (tail, root) pairs of all
// 64 lanes are then checked against the bit-parallel labels 64 at a time,
// one pair per lane, so a tail with many fresh roots does not hold up the
// wave.  Descriptors name the active vertex; its last group is row v of ag.
// An iteration whose frontier has more push edges than pull_ratio / 1000 of
// the adjacency entries at or above roots_start runs in pull form instead
// (pull_step; ~0u: never; PL = false compiles the pull form out).
template <bool W, uint32_t CH, uint32_t PW, bool PL = true>
__device__ __forceinline__ void push_body(uint32_t bid, uint32_t nblk, uint32_t iter, const uint64_t *off,
                                              const uint32_t *adj, unsigned long long *abits,
                                              const unsigned long long *ag,
                                              const uint32_t *amask, const uint32_t *aend,
                                              const unsigned long long *chunks, unsigned long long *ind,
                                              unsigned long long *cand, unsigned long long *vmark, uint32_t *tlist,
                                              uint32_t *clist,
                                              const uint8_t *bpd, const unsigned long long *lmask,
                                              const unsigned long long *rmask, const ulonglong2 *bps,
                                              const ulonglong2 *rset, BuildCounters *C, uint32_t bp_prune,
                                              uint32_t pull_ratio)
{
    static_assert(CH >= 4 && CH <= 64 && (CH & (CH - 1)) == 0, "chunk: a power of two in [4, 64]");
    constexpr uint32_t DPS = 64 / CH;  // descriptors per wave step
    __shared__ unsigned long long s_nb[PW][4][64]; // fresh roots of each lane's edge, per word of the step
    __shared__ unsigned long long s_keep[PW][4][64];  // of those, the ones the BP check keeps
    __shared__ uint32_t s_ex[PW][64];              // exclusive prefix of the lanes' fresh counts
    __shared__ uint32_t s_cu[PW][64];              // a lane's running counts over its four words
    __shared__ uint32_t s_t[PW][64];               // each lane's tail
    // set-phase scratch (bp_set_phase)
    __shared__ unsigned long long s_srem[PW][64], s_sone[PW][64];
    __shared__ uint32_t s_sro[PW][64], s_st[PW][64], s_sex[PW][64], s_shit[PW][64];
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t wave = bid * PW + wv;
    const uint32_t nwaves = nblk * PW;
    const SetScratch sc{s_srem[wv], s_sone[wv], s_sro[wv], s_st[wv], s_sex[wv], s_shit[wv]};
    // The check reads level masks while they cover every td it acts on, the
    // distance bytes after that.
    const bool use_masks = iter <= kMaskIters;
    // Iteration `iter` of batch_process's loop (PARA:1478-1670) runs if the
    // previous one left active vertices (frontier slot iter & 1); past the
    // batch's end the launch is empty.  Iteration 255 would leave the
    // reference's u8 domain: it is flagged, not run.  Block 0 also empties the
    // other slot, which k_add fills with the next frontier.
    if (C->hold) return;
    const uint32_t p = iter & 1u;
    const uint32_t n_act0 = C->n_act[p];
    const bool live = n_act0 != 0 && iter < Hop<W>::kInf;
    const uint32_t nc = live ? C->n_chunk[p] : 0u;
    const uint32_t rs = C->rs, gen_batch = C->gen_batch;
    const unsigned long long gen_iter = ((unsigned long long)gen_batch << 16) | iter;
    const uint32_t n = C->n_vert, nw = C->abits_words;
    // Pull or push: the same decision in every wave (none of its inputs
    // changes during the launch).
    const bool pull = PL && live && pull_ratio != ~0u &&
                      C->n_pedge[p] * 1000ull > (unsigned long long)(off[n] - off[rs]) * pull_ratio;
    if (bid == 0 && threadIdx.x == 0) {
        C->iter_last = iter;
        C->n_act[p ^ 1u] = 0;
        C->n_chunk[p ^ 1u] = 0;
        C->n_pedge[p ^ 1u] = 0;
        if (n_act0 != 0) {
            if (iter >= Hop<W>::kInf) {
                C->domain = 1;
            } else {
                C->iterations += 1;
                C->actives += n_act0;
                if (pull) C->pulls += 1;
            }
        }
    }
    // The other bitmap slot is emptied for k_add to mark the next frontier.
    if (live) {
        unsigned long long *nbits = abits + (uint64_t)(p ^ 1u) * nw;
        for (uint32_t i = bid * blockDim.x + threadIdx.x; i < nw; i += nblk * blockDim.x) nbits[i] = 0ull;
    }
    // Per-lane statistics of one launch fit 32 bits; summed in 64 at the end.
    uint32_t n_push = 0, n_pass = 0, n_bp = 0, n_cand = 0, n_edge = 0, n_tail = 0;
    if constexpr (PL)
        if (pull)
            pull_step<W>(wave, nwaves, iter, p, rs, n, gen_iter, gen_batch, off, adj, abits + (uint64_t)p * nw, ag, ind,
                     cand, vmark, tlist, clist, bpd, lmask, rmask, bps, rset, C, bp_prune, s_nb[wv][0], s_keep[wv][0],
                     s_ex[wv], s_t[wv], sc, n_push, n_pass, n_bp, n_cand, n_edge, n_tail);
    for (uint32_t c0 = wave * DPS; !pull && c0 < nc; c0 += nwaves * DPS) {
        const uint32_t c = c0 + lane / CH;
        const unsigned long long desc = c < nc ? chunks[c] : kChunkPad;
        uint32_t k = 0, t = 0, lim = 0, words = 0;
        if (desc != kChunkPad) {
            k = (uint32_t)desc;
            const uint32_t e = (uint32_t)(desc >> 32) + (lane & (CH - 1));
            if (e < aend[k]) {
                ++n_edge;
                t = adj[off[k] + e];
                lim = t - rs;  // roots with offset < lim satisfy root < tail (PARA:1134)
                // words w with w * 64 < lim
                words = amask[k] & (lim >= kBatch ? 0xFFFFu : ((1u << ((lim + 63) / 64)) - 1u));
            }
        }
        s_t[wv][lane] = t;
        uint32_t wall = words;
        for (int o = 32; o > 0; o >>= 1) wall |= __shfl_xor(wall, o);
        wall = uniform32(wall);  // the same in every lane: the step's word choice lives in scalar registers
        bool touched = false, got = false;
        while (wall) {
            // Up to four root words per step (wave-uniform): their indicator
            // test-and-sets are issued together, and their fresh pairs share
            // the bit-parallel passes.
            uint32_t wsel[4] = {0u, 0u, 0u, 0u}, gsel = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (wall) {
                    wsel[u] = __ffs(wall) - 1;
                    wall &= wall - 1;
                    gsel |= 1u << u;
                }
            }
            unsigned long long g[4], seen[4], nb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                g[u] = 0ull;
                seen[u] = 0ull;
                const uint32_t w = wsel[u];
                if (((gsel >> u) & 1u) && ((words >> w) & 1u)) {
                    g[u] = ag[(uint64_t)k * kWords + w];
                    if (lim < (w + 1) * 64) g[u] &= (1ull << (lim - w * 64)) - 1ull;
                    seen[u] = ind[(uint64_t)t * kWords + w];
                }
            }
            // Bits are only set during the kernel, so bits read as set are
            // set; the test-and-set runs only where some look fresh.
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                nb[u] = 0ull;
                if (g[u]) {
                    n_push += __popcll(g[u]);
                    if (g[u] & ~seen[u]) nb[u] = g[u] & ~atomicOr(&ind[(uint64_t)t * kWords + wsel[u]], g[u]);
                }
            }
            const uint32_t cu1 = (uint32_t)__popcll(nb[0]);
            const uint32_t cu2 = cu1 + (uint32_t)__popcll(nb[1]);
            const uint32_t cu3 = cu2 + (uint32_t)__popcll(nb[2]);
            const uint32_t cnt = cu3 + (uint32_t)__popcll(nb[3]);
            if (cnt) {
                touched = true;
                n_pass += cnt;
            }
            const uint32_t incl = wave_incl_scan(cnt);
            const uint32_t total = __shfl(incl, 63);
            if (total == 0) continue;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                s_nb[wv][u][lane] = nb[u];
                s_keep[wv][u][lane] = 0ull;
            }
            s_ex[wv][lane] = incl - cnt;
            s_cu[wv][lane] = cu1 | (cu2 << 10) | (cu3 << 20);
            wave_lds_sync();
            // Fresh pair x of the wave: owner lane o (the last lane whose
            // prefix is <= x; it has a fresh root), then its word u and root
            // bit b from the lane's per-word running counts.  Lanes of one
            // owner share its tail, so their loads of the tail's line merge.
            // All 64 lanes run every pass (the set phase works across the
            // wave); lanes past the last pair carry nothing.
            for (uint32_t x0 = 0; x0 < total; x0 += 64) {
                const uint32_t x = x0 + lane;
                const bool act = x < total;
                uint32_t o = 0, u = 0, b = 0, w = 0, to = 0;
                unsigned long long cs = 0ull, one = 0ull;
                bool dist_pr = true;
                if (act) {
                    o = item_owner(s_ex[wv], x);
                    const uint32_t rank = x - s_ex[wv][o], cu = s_cu[wv][o];
                    const uint32_t c1 = cu & 1023u, c2 = (cu >> 10) & 1023u, c3 = cu >> 20;
                    u = (rank >= c1 ? 1u : 0u) + (rank >= c2 ? 1u : 0u) + (rank >= c3 ? 1u : 0u);
                    const uint32_t base = u == 0 ? 0u : (u == 1 ? c1 : (u == 2 ? c2 : c3));
                    b = nth_set_bit(s_nb[wv][u][o], rank - base);
                    w = u == 0 ? wsel[0] : (u == 1 ? wsel[1] : (u == 2 ? wsel[2] : wsel[3]));
                    to = s_t[wv][o];
                    if (use_masks) {
                        dist_pr = bp_mask_scan(iter, rmask, w * 64 + b, lmask + (uint64_t)to * kLmWords, cs, one);
                    } else {
                        const uint4 *rl = reinterpret_cast<const uint4 *>(bpd + (uint64_t)(rs + w * 64 + b) * kBpdStride);
                        const uint4 *tl = reinterpret_cast<const uint4 *>(bpd + (uint64_t)to * kBpdStride);
                        const uint4 a0 = rl[0], a1 = rl[1], a2 = rl[2], a3 = rl[3];
                        const uint4 q0 = tl[0], q1 = tl[1], q2 = tl[2], q3 = tl[3];
                        const uint32_t dr[13] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w,
                                                 a2.x, a2.y, a2.z, a2.w, a3.x};
                        const uint32_t dt[13] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                                                 q2.x, q2.y, q2.z, q2.w, q3.x};
                        dist_pr = bp_dist_scan<W>(dr, dt, iter, cs, one);
                    }
                    if (!bp_prune) {
                        // Labels-only pruning (a wide build whose bit-parallel
                        // BFS was cut at level 254: a slot that stops short
                        // cannot vouch for the vertices behind a pruned one).
                        dist_pr = false;
                        cs = 0ull;
                    }
                    if (!dist_pr && cs) {
                        const uint32_t ro = w * 64 + b;
                        const unsigned long long *lt = lmask + (uint64_t)to * kLmWords;
                        cs = bp_nz_filter(cs, one, rmask[kLmNz1 * kBatch + ro], rmask[kLmNz0 * kBatch + ro], lt[kLmNz1],
                                          lt[kLmNz0]);
                    }
                }
                const bool set_pr = bp_set_phase(dist_pr ? 0ull : cs, one, rset, w * 64 + b, bps, to, sc);
                if (act) {
                    if (dist_pr || set_pr)
                        ++n_bp;
                    else
                        atomicOr(&s_keep[wv][u][o], 1ull << b);
                }
            }
            wave_lds_sync();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const unsigned long long pass = s_keep[wv][u][lane];
                if (pass) {
                    atomicOr(&cand[(uint64_t)t * kWords + wsel[u]], pass);
                    n_cand += __popcll(pass);
                    got = true;
                }
            }
            // The next step's writes must not overtake this step's reads.
            wave_lds_sync();
        }
        // vmark[t] = (batch << 16 | iteration) << 1 | candidate, raised with
        // atomicMax: one atomic tells whether t is new to the batch's touched
        // list, new to this iteration, and new to its candidate list.  The
        // plain read (never above the true value) skips the atomic when
        // nothing would change.
        bool new_t = false, new_c = false;
        if (touched) {
            const unsigned long long want = (gen_iter << 1) | (got ? 1ull : 0ull);
            unsigned long long prev = vmark[(uint64_t)t * kLmWords];
            if (prev < want) {
                prev = atomicMax(&vmark[(uint64_t)t * kLmWords], want);
                new_t = (prev >> 17) != gen_batch;
                if ((prev >> 1) != gen_iter) ++n_tail;
                new_c = got && prev < want;
            }
        }
        const uint32_t ti = wave_append(&C->n_touched, new_t);
        if (new_t) tlist[ti] = t;
        const uint32_t ci = wave_append(&C->n_cand[p], new_c);
        if (new_c) clist[ci] = t;
        // s_t is rewritten by the next step.
        wave_lds_sync();
    }
    const unsigned long long t_push = wave_sum_u64(n_push), t_pass = wave_sum_u64(n_pass);
    const unsigned long long t_bp = wave_sum_u64(n_bp), t_cand = wave_sum_u64(n_cand);
    const unsigned long long t_edge = wave_sum_u64(n_edge), t_tail = wave_sum_u64(n_tail);
    if (lane == 0) {
        if (t_edge) atomicAdd(&C->edges, t_edge);
        if (t_tail) atomicAdd(&C->tails, t_tail);
        if (t_push) atomicAdd(&C->pushes, t_push);
        if (t_pass) atomicAdd(&C->ind_pass, t_pass);
        if (t_bp) atomicAdd(&C->bp_pruned, t_bp);
        if (t_cand) atomicAdd(&C->cands, t_cand);
    }
}

template <bool W, uint32_t CH, uint32_t PW>
__global__ __launch_bounds__(PW * 64) __attribute__((amdgpu_waves_per_eu(4))) void k_push(uint32_t iter, const uint64_t *off,
                                              const uint32_t *adj, unsigned long long *abits,
                                              const unsigned long long *ag,
                                              const uint32_t *amask, const uint32_t *aend,
                                              const unsigned long long *chunks, unsigned long long *ind,
                                              unsigned long long *cand, unsigned long long *vmark, uint32_t *tlist,
                                              uint32_t *clist,
                                              const uint8_t *bpd, const unsigned long long *lmask,
                                              const unsigned long long *rmask, const ulonglong2 *bps,
                                              const ulonglong2 *rset, BuildCounters *C, uint32_t bp_prune,
                                              uint32_t pull_ratio)
{
    push_body<W, CH, PW>(blockIdx.x, gridDim.x, iter, off, adj, abits, ag, amask, aend, chunks, ind, cand, vmark, tlist,
                         clist, bpd, lmask, rmask, bps, rset, C, bp_prune, pull_ratio);
}

// Batch b's schedule into the counters: roots [rs, rs + rsz), cnt of them
// not bit-parallel (the first frontier, one self label each).
__device__ inline void batch_begin_write(BuildCounters *C, uint32_t b, uint32_t rs, uint32_t rsz, uint32_t cnt)
{
    C->next_batch = b + 1;
    C->rs = rs;
    C->rsz = rsz;
    C->gen_batch = b + 1;
    C->n_act[0] = 0;
    C->n_act[1] = cnt;
    C->n_pedge[0] = C->n_pedge[1] = 0;
    C->n_chunk[0] = C->n_chunk[1] = 0;
    C->n_cand[0] = C->n_cand[1] = 0;
    C->n_touched = rsz;
    C->labels += cnt;
}

// Start of a batch (construct's loop over batches, PARA:1791-1812): roots
// [b * batch, ...) with b taken from the device-side batch counter.  One
// thread per root offset: the block also counts the batch's
// non-bit-parallel roots, which are the first frontier (active slots and
// touched-list slots [0, rsz) belong to the root offsets, filled by
// k_batch_init) and get one self label each.
__device__ __forceinline__ void batch_begin_body(uint32_t bid, uint32_t nblk, const uint8_t *used, BuildCounters *C, uint32_t batch,
                                                      uint32_t n)
{
    if (C->hold) return;
    const uint32_t b = C->next_batch;
    const uint32_t rs = b * batch;
    const uint32_t rsz = rs < n ? min(batch, n - rs) : 0u;
    const uint32_t o = threadIdx.x;
    const uint32_t cnt = (uint32_t)__syncthreads_count(o < rsz && !used[rs + o]);
    if (o != 0) return;
    batch_begin_write(C, b, rs, rsz, cnt);
}

__global__ __launch_bounds__(1024) void k_batch_begin(const uint8_t *used, BuildCounters *C, uint32_t batch,
                                                      uint32_t n)
{
    batch_begin_body(blockIdx.x, gridDim.x, used, C, batch, n);
}

// add (PARA:1568-1665): one wave per candidate vertex v.  Lane l owns roots
// [16l, 16l+16).  distance_query (PARA:1218-1270) runs over v's labels with
// dist < iter: a hub h can prune only if d_v(h) + min_r d_r(h) <= iter, so a
// per-hub minimum filters most hubs before the 1 KB root-distance row of the
// survivors is read (16 B per lane).  Only hubs below the candidate root
// count (PARA:1250-1254), so v's own label (hub v, above every root that
// pushes to v) never prunes: iteration 1, whose only labels with dist < 1 are
// such self labels, scans nothing, and iteration 2 scans v's distance-1 list
// instead of the whole label array.
constexpr uint32_t kAddHubs = 2;  // root-distance rows k_add loads per step
constexpr uint32_t kGather = 8;   // candidate roots up to which k_add reads single entries

// DEEP compiles in the hub-parallel distance query (see below), which only
// iterations past kNear + 1 use; the launch schedule runs the other
// instantiation for the iterations before, whose code it leaves as lean as
// it was.
template <bool W, uint32_t CH, bool DEEP = true>
__device__ __forceinline__ void add_body(uint32_t bid, uint32_t nblk, uint32_t iter, const uint8_t *used, const uint64_t *off, const uint32_t *adj, const uint32_t *clist,
                                             unsigned long long *cand, LabelStore S, HubList1 H1, HubInfo *hinfo,
                                             typename Hop<W>::Dist *dc,
                                             unsigned long long *abits, unsigned long long *ag, uint32_t *amask,
                                             uint32_t *aend, unsigned long long *chunks, BuildCounters *C)
{
    const uint32_t lane = lane_id();
    const uint32_t wave = (bid * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (nblk * blockDim.x) >> 6;
    // Candidates of iteration `iter` (slot iter & 1); the next frontier goes
    // to the other slot, and block 0 empties the other candidate slot for
    // the next iteration's push.
    if (C->hold) return;
    const uint32_t p = iter & 1u;
    constexpr uint32_t B = Hop<W>::kBits, INF = Hop<W>::kInf;
    const uint32_t ncand = iter < INF ? C->n_cand[p] : 0u;  // 0 past the batch's end
    const uint32_t rs = C->rs, rsz = C->rsz;
    if (bid == 0 && threadIdx.x == 0) C->n_cand[p ^ 1u] = 0;
    // Per-lane (n_dq, n_read) and per-wave (n_lab, n_rows) counts of one launch.
    uint32_t n_dq = 0, n_lab = 0, n_rows = 0, n_read = 0;
    // Next-frontier push descriptors are reserved up to kChunkGrab at a time
    // per wave, so the shared counter sees one atomic per reservation, not per
    // vertex; unused descriptor slots are padded.  The active count and the
    // frontier's push edges are added once per wave at the end.
    uint32_t c_next = 0, c_left = 0, my_act = 0;
    uint32_t my_pedge = 0;  // push edges of this wave's new frontier entries (flushed below 2^31)
    unsigned long long *nbits = abits + (uint64_t)(p ^ 1u) * C->abits_words;
    // The loop over this wave's candidates (grid-stride: i, i + nwaves, ...)
    // is software-pipelined: while candidate i is processed, the loads of
    // the next two are in flight -- for candidate i + nwaves its first labels
    // and first adjacency entries (stage B), which need its stage A, and for
    // candidate i + 2 nwaves its candidate word, label-store entry, flag,
    // offsets and distance-1 count (stage A); the candidate list itself is
    // read three ahead.  The vertices differ, so none of these loads can see
    // a store of an earlier candidate.
    struct StageA {
        unsigned long long word;  // candidate bits of this lane's four words' quarter
        LRec rec;
        uint64_t e0, e1;
        uint32_t v, used, n1;
    };
    struct StageB {
        uint32_t lab0, lab1, probe;
    };
    auto load_a = [&](uint32_t v, bool ok) {
        StageA a{};
        a.v = v;
        if (ok) {
            a.word = cand[(uint64_t)v * kWords + (lane >> 2)];
            a.rec = S.rec[v];
            a.used = used[v];
            a.e0 = off[v];
            a.e1 = off[v + 1];
            a.n1 = iter <= 2u ? H1.len[v] : 0u;
        }
        return a;
    };
    // Per-vertex values are the same in every lane: kept in scalar registers.
    auto uniform_a = [&](const StageA &a) {
        StageA u = a;
        u.rec.len = uniform32(a.rec.len);
        u.rec.base = uniform64(a.rec.base);
        u.rec.cap = uniform32(a.rec.cap);
        u.e0 = uniform64(a.e0);
        u.e1 = uniform64(a.e1);
        u.v = uniform32(a.v);
        u.used = uniform32(a.used);
        u.n1 = uniform32(a.n1);
        return u;
    };
    // The labels the distance query scans (see above): iteration 2 reads the
    // distance-1 list, whose first 128 entries are read within v's slots
    // before n1 is known.
    auto load_b = [&](const StageA &a, bool ok) {
        StageB b{0u, 0u, 0u};
        if (ok) {
            const uint32_t lim0 = iter == 1u ? 0u : (iter == 2u ? (uint32_t)(a.e1 - a.e0) : a.rec.len);
            const uint32_t *lsrc = iter == 2u ? H1.lab + a.e0 : S.pool + a.rec.base;
            if (lane < lim0) b.lab0 = lsrc[lane];
            if (lane + 64 < lim0) b.lab1 = lsrc[lane + 64];
            if (!a.used && lane < a.e1 - a.e0) b.probe = adj[a.e0 + lane];
        }
        return b;
    };
    const uint32_t c0 = wave < ncand ? clist[wave] : 0u;
    const uint32_t c1 = wave + nwaves < ncand ? clist[wave + nwaves] : 0u;
    uint32_t vq = wave + 2 * nwaves < ncand ? clist[wave + 2 * nwaves] : 0u;  // candidate i + 2 nwaves
    StageA A0 = uniform_a(load_a(c0, wave < ncand));                           // candidate i
    StageB B0 = load_b(A0, wave < ncand);
    StageA A1 = load_a(c1, wave + nwaves < ncand);                             // candidate i + nwaves
    for (uint32_t i = wave; i < ncand; i += nwaves) {
        const StageA cu = A0;
        const StageB cbl = B0;
        {
            const bool ok1 = i + nwaves < ncand, ok2 = i + 2 * nwaves < ncand;
            A0 = uniform_a(A1);
            B0 = load_b(A0, ok1);
            A1 = load_a(uniform32(vq), ok2);
            vq = i + 3 * nwaves < ncand ? clist[i + 3 * nwaves] : 0u;
        }
        const uint32_t v = cu.v;
        const unsigned long long word = cu.word;
        const uint32_t len = cu.rec.len;
        const unsigned long long base = cu.rec.base;
        const uint32_t cap = cu.rec.cap;
        const bool v_used = cu.used != 0u;
        const uint64_t e0 = cu.e0, e1 = cu.e1;
        const uint32_t n1 = cu.n1;  // distance-1 labels held
        const uint32_t slen = iter == 1u ? 0u : (iter == 2u ? n1 : len);
        const uint32_t *lsrc = iter == 2u ? H1.lab + e0 : S.pool + base;
        const uint32_t lab0 = cbl.lab0, lab1 = cbl.lab1;
        // Neighbours >= rs: a prefix of the descending adjacency; counted
        // from the loaded entries when they cover it.
        const uint32_t pdeg = v_used ? 0u
                              : e1 - e0 <= 64u
                                  ? (uint32_t)__popcll(__ballot(lane < e1 - e0 && cbl.probe >= rs))
                                  : push_degree_range(adj, e0, e1, rs);
        const uint32_t cb = (uint32_t)(word >> (16 * (lane & 3))) & 0xFFFFu;
        // The clear below is control-dependent on every lane's loaded word
        // (through the ballot), so no lane can clear a word before all lanes
        // have read it.
        if (__ballot(cb != 0) != 0ull && lane < kWords) cand[(uint64_t)v * kWords + lane] = 0ull;
        uint32_t prune = 0;
        const uint32_t my_lo = rs + 16 * lane;  // real id of this lane's first root
        // The passing hubs of one set of 64 labels (h, d, lm per lane): up to
        // kAddHubs per step, their row loads issued together; only lanes of
        // the hub's mask that still hold an unpruned candidate load their
        // slice of a row (16 entries: 16 bytes, 32 wide).
        auto hub_pass = [&](uint32_t h, uint32_t d, unsigned long long lm) {
            // A hub is worth a row only if some lane that still holds an
            // unpruned candidate has a root close enough to it.
            const unsigned long long live = __ballot((cb & ~prune) != 0u);
            unsigned long long m = __ballot((lm & live) != 0ull);
            while (m) {
                int src[kAddHubs];
                bool ok[kAddHubs];
#pragma unroll
                for (int u = 0; u < (int)kAddHubs; ++u) {
                    ok[u] = m != 0ull;
                    src[u] = ok[u] ? __ffsll(m) - 1 : 0;
                    if (ok[u]) m &= m - 1;
                }
                const bool need = (cb & ~prune) != 0u;
                constexpr uint32_t RV = W ? 2 : 1;  // 16-byte vectors of a lane's slice
                uint32_t hh[kAddHubs], thr[kAddHubs];
                uint4 row[kAddHubs][RV];
#pragma unroll
                for (int u = 0; u < (int)kAddHubs; ++u) {
                    hh[u] = __shfl(h, src[u]);
                    // prune root r if d_r(h) <= iter - d_v(h); an unused slot u
                    // (fewer than kAddHubs hubs left) gets threshold 0, which the
                    // all-INF row never meets
                    const uint32_t du = __shfl(d, src[u]);
                    thr[u] = ok[u] ? iter - du : 0u;
                    const uint32_t mlo = __shfl((uint32_t)lm, src[u]), mhi = __shfl((uint32_t)(lm >> 32), src[u]);
                    const bool mine = ((lane < 32 ? mlo : mhi) >> (lane & 31)) & 1u;
#pragma unroll
                    for (uint32_t q = 0; q < RV; ++q) row[u][q] = make_uint4(~0u, ~0u, ~0u, ~0u);
                    if (ok[u] && need && mine) {
                        const uint4 *rp = reinterpret_cast<const uint4 *>(dc + (uint64_t)hh[u] * kBatch) + RV * lane;
#pragma unroll
                        for (uint32_t q = 0; q < RV; ++q) row[u][q] = rp[q];
                    }
                    n_rows += ok[u] ? 1 : 0;
                }
#pragma unroll
                for (int u = 0; u < (int)kAddHubs; ++u) {
                    uint32_t wds[4 * RV];
#pragma unroll
                    for (uint32_t q = 0; q < RV; ++q) {
                        wds[4 * q] = row[u][q].x;
                        wds[4 * q + 1] = row[u][q].y;
                        wds[4 * q + 2] = row[u][q].z;
                        wds[4 * q + 3] = row[u][q].w;
                    }
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        const uint32_t x = W ? (wds[q >> 1] >> ((q & 1) * 16)) & 0xFFFFu
                                             : (wds[q >> 2] >> ((q & 3) * 8)) & 255u;
                        if (x <= thr[u] && hh[u] < my_lo + q) prune |= 1u << q;
                    }
                }
            }
        };
        // What hub h of a label at distance d can do (lanes whose roots it
        // can reach in time; 0 if none).
        auto hub_lanes = [&](uint32_t h, uint32_t d) -> unsigned long long {
            const HubInfo hi = hinfo[h];
            const uint32_t t = iter - d;
            unsigned long long lm = 0ull;
            if (t <= kNear) {
                lm = hi.near[0];
                if (t >= 2u) lm |= hi.near[1];
                if (t >= 3u) lm |= hi.near[2];
            } else if (d + (hi.info & INF) <= iter) {
                lm = ~0ull;
            }
            return lm;
        };
        // Past iteration kNear + 1 the hub summaries stop naming lanes (a
        // threshold above kNear only has the minimum), so most hubs pass and
        // reading a row slice per passing hub costs a round trip per two
        // hubs.  There, a vertex with few candidate roots takes the distance
        // query hub-parallel: lanes hold labels, and each lane reads the one
        // root-distance entry per candidate root it needs, all of a step's
        // reads in flight together.  Otherwise lanes hold roots (16 each) and
        // read row slices of the passing hubs (hub_pass).  croot: the
        // candidate root offsets (wave-uniform); cpr: which of them are
        // pruned.
        uint32_t croot[kGather];
        uint32_t ncand_v = kGather + 1, cpr = 0;
        if (DEEP && iter > kNear + 1 && slen != 0) {
            ncand_v = 0;
            uint32_t rest = cb;
#pragma unroll
            for (uint32_t k = 0; k < kGather; ++k) {
                const unsigned long long m = __ballot(rest != 0u);
                croot[k] = ~0u;
                if (m) {
                    const uint32_t l0 = (uint32_t)__ffsll(m) - 1;
                    const uint32_t bits = uniform32(__shfl(rest, (int)l0));
                    croot[k] = 16 * l0 + (uint32_t)__ffs(bits) - 1;
                    if (lane == l0) rest &= rest - 1;
                    ++ncand_v;
                }
            }
            if (__ballot(rest != 0u) != 0ull) ncand_v = kGather + 1;  // too many: row slices
        }
        const bool gather = ncand_v <= kGather;
        // One label per lane against the listed candidates: prune candidate k
        // if d + d_r(h) <= iter with h below it (PARA:1250-1262).
        auto gather_pass = [&](uint32_t h, uint32_t d, unsigned long long lm) {
            if (__ballot(lm != 0ull) == 0ull) return;
            uint32_t x[kGather];
#pragma unroll
            for (uint32_t k = 0; k < kGather; ++k) {
                const uint32_t ro = croot[k];
                x[k] = ~0u;
                if (k < ncand_v && !((cpr >> k) & 1u) && ((lm >> (ro >> 4)) & 1ull) && h < rs + ro)
                    x[k] = dc[(uint64_t)h * kBatch + ro];
            }
            bool any = false;
#pragma unroll
            for (uint32_t k = 0; k < kGather; ++k) {
                any = any || x[k] != ~0u;
                if (__ballot(x[k] != ~0u && d + x[k] <= iter) != 0ull) cpr |= 1u << k;
            }
            n_rows += __ballot(any) != 0ull ? 1u : 0u;
        };
        // 128 labels per step (two per lane, their hub summaries loaded
        // together); the next step's labels are loaded while this one runs.
        // visit(h, d, lm) per half step; it returns true when every candidate
        // is pruned (the rest of the labels cannot matter).
        auto scan_labels = [&](auto &&visit) {
            uint32_t la = lab0, lb = lab1;
            for (uint32_t j0 = 0; j0 < slen; j0 += 128) {
                const uint32_t ja = j0 + lane, jb = j0 + 64 + lane;
                const uint32_t ca = la, cbl = lb;
                la = ja + 128 < slen ? lsrc[ja + 128] : 0u;
                lb = jb + 128 < slen ? lsrc[jb + 128] : 0u;
                uint32_t ha = 0, da = INF, hb = 0, db = INF;
                if (ja < slen) {
                    ha = ca >> B;
                    da = ca & INF;
                }
                if (jb < slen) {
                    hb = cbl >> B;
                    db = cbl & INF;
                }
                unsigned long long lma = 0ull, lmb = 0ull;
                if (da < iter) {
                    lma = hub_lanes(ha, da);
                    ++n_read;
                }
                if (db < iter) {
                    lmb = hub_lanes(hb, db);
                    ++n_read;
                }
                if (visit(ha, da, lma)) break;
                if (visit(hb, db, lmb)) break;
            }
        };
        if (DEEP && gather) {
            scan_labels([&](uint32_t h, uint32_t d, unsigned long long lm) {
                gather_pass(h, d, lm);
                return cpr == (1u << ncand_v) - 1u;
            });
#pragma unroll
            for (uint32_t k = 0; k < kGather; ++k)
                if (((cpr >> k) & 1u) && lane == (croot[k] >> 4)) prune |= 1u << (croot[k] & 15u);
        } else {
            scan_labels([&](uint32_t h, uint32_t d, unsigned long long lm) {
                hub_pass(h, d, lm);
                return __ballot((cb & ~prune) != 0u) == 0ull;
            });
        }
        const uint32_t surv = cb & ~prune;
        n_dq += __popc(cb & prune);
        const uint32_t cnt = __popc(surv);
        uint32_t incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(incl, o);
            if (lane >= (uint32_t)o) incl += y;
        }
        const uint32_t total = __shfl(incl, 63);
        if (total == 0) continue;
        n_lab += total;
        const unsigned long long nb = store_reserve_at(S, C, v, len, base, cap, total);
        if (nb == ~0ull) continue;
        uint32_t pos = len + incl - cnt;
        for (uint32_t sb = surv; sb; sb &= sb - 1) {
            const uint32_t q = __ffs(sb) - 1;
            S.pool[nb + pos++] = ((my_lo + q) << B) | iter;  // insert_label_only, PARA:1285
        }
        if (lane == 0) S.rec[v].len = len + total;
        if (iter == 1u) {
            // The new labels are at distance 1: v's list gets them too.
            uint32_t p1 = n1 + incl - cnt;
            for (uint32_t sb = surv; sb; sb &= sb - 1) H1.lab[e0 + p1++] = ((my_lo + __ffs(sb) - 1) << B) | 1u;
            if (lane == 0) H1.len[v] = n1 + total;
        }
        // Roots of this batch also record the new label in the root table
        // (dist_matrix[v_root][cand_real] = iter, PARA:1287-1290).
        if (v >= rs && v < rs + rsz) {
            for (uint32_t sb = surv; sb; sb &= sb - 1) {
                const uint32_t h = my_lo + __ffs(sb) - 1;
                dc[(uint64_t)h * kBatch + (v - rs)] = (typename Hop<W>::Dist)iter;
                if (iter < (hinfo[h].info & INF)) atomicMin(&hinfo[h].info, ~INF | iter);
                if (iter <= kNear) atomicOr(&hinfo[h].near[iter - 1], 1ull << ((v - rs) >> 4));
            }
        }
        // v becomes active with the new distance group as its last group
        // (row v of ag) and, unless it is a bit-parallel vertex (whose
        // pushes the reference skips, PARA:1117-1119), joins the next
        // frontier's bitmap.  Descriptor reservations are capped by the
        // vertices this wave has left in the launch (grid-stride), so light
        // launches reserve exactly.
        const uint32_t rem = (ncand - 1 - i) / nwaves + 1;
        ++my_act;
        my_pedge += __builtin_amdgcn_readfirstlane(pdeg);
        if (my_pedge >= 0x80000000u) {
            if (lane == 0) atomicAdd(&C->n_pedge[p ^ 1u], (unsigned long long)my_pedge);
            my_pedge = 0;
        }
        reinterpret_cast<uint16_t *>(ag + (uint64_t)v * kWords)[lane] = (uint16_t)surv;
        const unsigned long long nzm = __ballot(surv != 0);
        uint32_t wmask = 0;
#pragma unroll
        for (int w = 0; w < (int)kWords; ++w)
            if ((nzm >> (4 * w)) & 0xFull) wmask |= 1u << w;
        if (lane == 0) {
            amask[v] = wmask;
            aend[v] = pdeg;
            if (!v_used) atomicOr(&nbits[v >> 6], 1ull << (v & 63));
        }
        const uint32_t nch = (pdeg + CH - 1) / CH;
        if (nch > c_left) {
            for (uint32_t q = lane; q < c_left; q += 64) chunks[c_next + q] = kChunkPad;
            const uint64_t guess = (uint64_t)nch * rem;  // chunks left if v is typical
            const uint32_t cap = guess < kChunkGrab ? (uint32_t)guess : kChunkGrab;
            const uint32_t want = nch > cap ? nch : cap;
            uint32_t b0 = 0;
            if (lane == 0) b0 = atomicAdd(&C->n_chunk[p ^ 1u], want);
            c_next = __shfl(b0, 0);
            c_left = want;
        }
        for (uint32_t q = lane; q < nch; q += 64) chunks[c_next + q] = ((unsigned long long)(q * CH) << 32) | v;
        c_next += nch;
        c_left -= nch;
    }
    for (uint32_t q = lane; q < c_left; q += 64) chunks[c_next + q] = kChunkPad;
    if (lane == 0 && my_act) atomicAdd(&C->n_act[p ^ 1u], my_act);
    if (lane == 0 && my_pedge) atomicAdd(&C->n_pedge[p ^ 1u], (unsigned long long)my_pedge);
    const unsigned long long t_dq = wave_sum_u64(n_dq), t_read = wave_sum_u64(n_read);
    if (lane == 0) {
        if (t_read) atomicAdd(&C->dq_labels, t_read);
        if (t_dq) atomicAdd(&C->dq_pruned, t_dq);
        if (n_lab) atomicAdd(&C->labels, (unsigned long long)n_lab);
        if (n_rows) atomicAdd(&C->dq_rows, n_rows);
    }
}

template <bool W, uint32_t CH, bool DEEP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_add(uint32_t iter, const uint8_t *used, const uint64_t *off, const uint32_t *adj, const uint32_t *clist,
                                             unsigned long long *cand, LabelStore S, HubList1 H1, HubInfo *hinfo,
                                             typename Hop<W>::Dist *dc,
                                             unsigned long long *abits, unsigned long long *ag, uint32_t *amask,
                                             uint32_t *aend, unsigned long long *chunks, BuildCounters *C)
{
    add_body<W, CH, DEEP>(blockIdx.x, gridDim.x, iter, used, off, adj, clist, cand, S, H1, hinfo, dc, abits, ag, amask,
                          aend, chunks, C);
}

// The graph's shape in the counters (once, at allocation).
__global__ void k_ctr_shape(BuildCounters *C, uint32_t n, uint32_t nw)
{
    C->n_vert = n;
    C->abits_words = nw;
}

__global__ void k_unhold(BuildCounters *C) { C->hold = 0u; }

// reset_at_end (PARA:1328-1410) and the next batch's indicator reset
// (PARA:877-901): clear the indicator rows of the touched vertices (the
// root-distance table is reset by hub_table_body), and frontier bitmap slot 1,
// which the next batch's first frontier goes into.
__device__ __forceinline__ void batch_end_body(uint32_t bid, uint32_t nblk, unsigned long long *ind, const uint32_t *tlist,
                                                   unsigned long long *abits, const BuildCounters *C)
{
    if (C->hold) return;
    const uint32_t nt = C->n_touched, nw = C->abits_words;
    const uint64_t stride = (uint64_t)nblk * blockDim.x;
    const uint64_t tot = (uint64_t)nt * kWords;
    for (uint64_t i = (uint64_t)bid * blockDim.x + threadIdx.x; i < tot; i += stride)
        ind[(uint64_t)tlist[i / kWords] * kWords + (i % kWords)] = 0ull;
    for (uint64_t i = (uint64_t)bid * blockDim.x + threadIdx.x; i < nw; i += stride) abits[nw + i] = 0ull;
}

__global__ __launch_bounds__(256) void k_batch_end(unsigned long long *ind, const uint32_t *tlist,
                                                   unsigned long long *abits, const BuildCounters *C)
{
    batch_end_body(blockIdx.x, gridDim.x, ind, tlist, abits, C);
}

// First kernel of a batch's prologue: the gate -- if the previous batch
// still has a frontier (its next iteration's count in slot (iter_last + 1) &
// 1), this batch was enqueued too early and everything up to the host's
// k_unhold is a no-op (a no-op before the first batch: all counts zero) --
// then the previous batch's reset -- its roots' root-distance entries (one wave per
// root in the first blocks) and batch_end_body -- which are independent of
// each other.  Every block takes the gate's decision itself from counters
// that nothing in this kernel changes.
template <bool W>
__global__ __launch_bounds__(256) void k_batch_close(const uint8_t *used, LabelStore S, HubInfo *hinfo,
                                                     typename Hop<W>::Dist *dc, unsigned long long *ind,
                                                     const uint32_t *tlist, unsigned long long *abits, BuildCounters *C)
{
    if (C->hold || C->n_act[(C->iter_last + 1u) & 1u] != 0u) {
        if (blockIdx.x == 0 && threadIdx.x == 0) C->hold = 1u;
        return;
    }
    hub_table_body<W>(blockIdx.x, gridDim.x, 0u, used, S, hinfo, dc, C);
    batch_end_body(blockIdx.x, gridDim.x, ind, tlist, abits, C);
}

// One graph's whole batch phase in one workgroup (no reference
// counterpart: the reference runs batch_process from the host, PARA:1413-1736
// and construct's loop over batches, PARA:1791-1812).  For graphs whose
// iterations are many and small -- road-like graphs and grids, hundreds of
// BFS levels per batch -- the launch-per-iteration schedule is bound by
// launch and host-check latency; here the same bodies run back to back with
// workgroup barriers in place of kernel boundaries, and a launch carries
// one workgroup per graph, so a set of such graphs is labeled by one launch.
template <bool W>
struct BatchPlan {
    uint32_t n, batch, bp_prune, pull_ratio;
    const uint8_t *used;
    const uint64_t *off;
    const uint32_t *adj;
    LabelStore S;
    HubList1 H1;
    unsigned long long *ind, *cand, *vmark;
    uint32_t *tlist, *clist;
    unsigned long long *abits, *ag;
    uint32_t *amask, *aend;
    unsigned long long *chunks;
    const uint8_t *bpd;
    const unsigned long long *lmask;
    unsigned long long *rmask;
    const ulonglong2 *bps;
    ulonglong2 *rset;
    HubInfo *hinfo;
    typename Hop<W>::Dist *dc;
    BuildCounters *C;
};

constexpr uint32_t kLoopWaves = 8;  // waves of a k_batch_loop workgroup

// Every phase ends at a workgroup barrier; the bodies' virtual blocks are
// looped over (bid) with nblk = 1, so their grid-stride loops cover the work.
// Leaves early (every thread, after a barrier) on a full pool or when an
// iteration would leave the hop domain; the counters say which.
template <bool W>
__global__ __launch_bounds__(kLoopWaves * 64) void k_batch_loop(const BatchPlan<W> *plans)
{
    constexpr uint32_t kEnd = Hop<W>::kInf;  // the last iteration ever run
    __shared__ uint32_t s_cnt;
    const BatchPlan<W> &P = plans[blockIdx.x];
    BuildCounters *C = P.C;
    const uint32_t nb = (P.n + P.batch - 1) / P.batch;
    for (uint32_t b = 0; b < nb; ++b) {
        // The previous batch's root-distance entries and indicator rows.
        const uint32_t prsz = C->rsz;
        for (uint32_t vb = 0; vb * kLoopWaves < prsz; ++vb)
            hub_table_body<W>(vb, 1, 0u, P.used, P.S, P.hinfo, P.dc, C);
        batch_end_body(0, 1, P.ind, P.tlist, P.abits, C);
        // Batch b: its non-bit-parallel roots are the first frontier.
        const uint32_t rs = b * P.batch, rsz = min(P.batch, P.n - rs);
        uint32_t c = 0;
        for (uint32_t o = threadIdx.x; o < rsz; o += blockDim.x) c += P.used[rs + o] ? 0u : 1u;
        c = wave_sum_u32(c);
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        if ((threadIdx.x & 63) == 0) atomicAdd(&s_cnt, c);
        __syncthreads();
        if (threadIdx.x == 0) batch_begin_write(C, b, rs, rsz, s_cnt);
        __syncthreads();
        for (uint32_t vb = 0; vb * kLoopWaves < rsz; ++vb)
            batch_init_body<W, kChunk>(vb, 1, P.used, P.off, P.adj, P.S, P.hinfo, P.dc, P.ind, P.vmark, P.tlist,
                                       P.abits, P.ag, P.amask, P.aend, P.chunks, P.bps, P.rset, P.lmask, P.rmask, C);
        __syncthreads();
        for (uint32_t iter = 1;; ++iter) {
            push_body<W, kChunk, kLoopWaves, false>(0, 1, iter, P.off, P.adj, P.abits, P.ag, P.amask, P.aend, P.chunks,
                                             P.ind, P.cand, P.vmark, P.tlist, P.clist, P.bpd, P.lmask, P.rmask,
                                             P.bps, P.rset, C, P.bp_prune, P.pull_ratio);
            __syncthreads();
            add_body<W, kChunk>(0, 1, iter, P.used, P.off, P.adj, P.clist, P.cand, P.S, P.H1, P.hinfo, P.dc, P.abits,
                                P.ag, P.amask, P.aend, P.chunks, C);
            __syncthreads();
            const bool stop = C->overflow != 0u || C->domain != 0u;
            const bool done = C->n_act[(iter + 1) & 1u] == 0u;
            __syncthreads();  // every thread has read the counters before any writes them
            if (stop) return;
            if (done) break;
            if (iter >= kEnd) {
                if (threadIdx.x == 0) C->domain = 1u;
                return;
            }
        }
    }
}

// Pool compaction: copy each vertex's labels into a fresh pool.
__global__ __launch_bounds__(256) void k_pool_move(uint32_t n, const uint32_t *src, LRec *rec,
                                                   const uint64_t *nbase, uint32_t *dst, const uint32_t *cap_new)
{
    const uint32_t lane = lane_id();
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < n; v += nwaves) {
        const LRec r = rec[v];
        const uint64_t nb = nbase[v];
        for (uint32_t j = lane; j < r.len; j += 64) dst[nb + j] = src[r.base + j];
        if (lane == 0) {
            rec[v].base = nb;
            rec[v].cap = cap_new[v];
        }
    }
}

// Per-vertex pool slots for a repack (exact = 0) or the label counts (1).
__global__ void k_cap_for_len(uint32_t n, const LRec *rec, uint32_t *cap, uint32_t exact)
{
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        uint32_t L = rec[v].len;
        cap[v] = exact ? L : (L + (L >> 2) + 4);
    }
}

// Final CSR gather (order is fixed by the segmented sort that follows).
__global__ __launch_bounds__(256) void k_gather_labels(uint32_t n, const uint32_t *pool, const LRec *rec,
                                                       const uint64_t *loff, uint32_t *lab)
{
    const uint32_t lane = lane_id();
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < n; v += nwaves) {
        const uint64_t o = loff[v];
        const uint32_t L = (uint32_t)(loff[v + 1] - o);
        const unsigned long long b = rec[v].base;
        for (uint32_t j = lane; j < L; j += 64) lab[o + j] = pool[b + j];
    }
}

}  // namespace

// Grows (or just compacts) the label pool between batches.
static pado_gpu_status pool_repack(hipStream_t st, uint32_t n, DevBuf<uint32_t> &pool, DevBuf<LRec> &rec,
                                   unsigned long long new_size, unsigned long long *new_top)
{
    DevBuf<uint32_t> ncap;
    DevBuf<uint64_t> nbase;
    PADO_TRY(ncap.alloc(n));
    PADO_TRY(nbase.alloc((uint64_t)n + 1));
    hipLaunchKernelGGL(k_cap_for_len, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, rec.p, ncap.p, 0u);
    PADO_TRY(scan_u32_to_u64(ncap.p, nbase.p, n, st));
    uint64_t need = 0;
    PADO_HIP(copy_sync(&need, nbase.p + n, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if (new_size < need * 2) new_size = need * 2;
    DevBuf<uint32_t> npool;
    PADO_TRY(npool.alloc(new_size));
    hipLaunchKernelGGL(k_pool_move, dim3(grid_for((uint64_t)n * 64, 256, 8192)), dim3(256), 0, st, n, pool.p, rec.p,
                       nbase.p, npool.p, ncap.p);
    PADO_HIP(hipStreamSynchronize(st));
    std::swap(pool.p, npool.p);
    std::swap(pool.n, npool.n);
    *new_top = need;
    return PADO_GPU_OK;
}

// The device buffers of one batch phase (layout at the top of this file).
template <bool W>
struct BatchBuffers {
    DevBuf<unsigned long long> ind, cand, ag, chunks;
    DevBuf<unsigned long long> abits;  // frontier bitmaps, two slots of ceil(n / 64) words
    DevBuf<unsigned long long> lmask;  // bit-parallel mask rows, kLmWords words per vertex
    DevBuf<unsigned long long> rmask;  // batch roots' mask words 0..kLmNz0, word-major [10][kBatch]
    DevBuf<uint32_t> tlist, clist, amask, aend, pool;
    DevBuf<HubInfo> hinfo;
    DevBuf<uint32_t> d1, d1len;  // distance-1 lists (HubList1): m slots, n counts
    DevBuf<LRec> lrec;
    DevBuf<typename Hop<W>::Dist> dc;
    DevBuf<ulonglong2> rset;  // batch roots' bit-parallel sets, slot-major [kBP][kBatch]
    DevBuf<BuildCounters> ctr;
    unsigned long long pool_size = 0;
    uint32_t dc_rows = 0;
    void release()
    {
        for (auto *b : {&ind, &cand, &ag, &chunks, &abits, &lmask, &rmask}) b->release();
        for (auto *b : {&tlist, &clist, &amask, &aend, &pool, &d1, &d1len}) b->release();
        hinfo.release();
        lrec.release();
        dc.release();
        rset.release();
        ctr.release();
    }
};

// Allocates and initialises them on the graph's stream.  add_waves bounds
// the padding of k_add's per-wave descriptor reservations: padded
// descriptors are fewer than the real ones plus kChunkGrab per wave.  The
// pool holds pool_size label words, the root-distance table dc_rows 1 KB
// rows (255).
template <bool W>
static pado_gpu_status batch_alloc(const pado_gpu_graph *g, const pado_gpu_index *idx, uint64_t add_waves,
                                   uint32_t chunk, unsigned long long pool_size, uint32_t dc_rows,
                                   BatchBuffers<W> &B)
{
    using DT = typename Hop<W>::Dist;
    hipStream_t st = g->ctx->stream;
    const uint32_t n = g->n;
    const uint32_t nw = (n + 63) / 64;  // words of one frontier bitmap slot
    PADO_TRY(B.ind.alloc((uint64_t)n * kWords));
    PADO_TRY(B.cand.alloc((uint64_t)n * kWords));
    PADO_TRY(B.ag.alloc((uint64_t)n * kWords));
    PADO_TRY(B.abits.alloc(2ull * nw));
    PADO_TRY(B.chunks.alloc(2 * (g->m / chunk + (uint64_t)n) + kChunkGrab * add_waves + 64));
    PADO_TRY(B.tlist.alloc(n));
    PADO_TRY(B.clist.alloc(n));
    PADO_TRY(B.amask.alloc(n));
    PADO_TRY(B.aend.alloc(n));
    PADO_TRY(B.hinfo.alloc(n));
    PADO_TRY(B.d1.alloc(g->m ? g->m : 1));
    PADO_TRY(B.d1len.alloc(n));
    PADO_TRY(B.lrec.alloc(n));
    PADO_TRY(B.rset.alloc((uint64_t)kBP * kBatch));
    PADO_TRY(B.lmask.alloc((uint64_t)n * kLmWords));
    PADO_TRY(B.rmask.alloc((kLmNz0 + 1ull) * kBatch));
    PADO_TRY(B.ctr.alloc(1));
    PADO_TRY(B.pool.alloc(pool_size));
    PADO_TRY(B.dc.alloc((uint64_t)dc_rows * kBatch));
    B.pool_size = pool_size;
    B.dc_rows = dc_rows;
    const uint32_t fg = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, st, reinterpret_cast<uint32_t *>(B.dc.p),
                       (uint64_t)dc_rows * kBatch * sizeof(DT) / 4, ~0u);
    hipLaunchKernelGGL(k_zero_u64, dim3(fg), dim3(256), 0, st, B.ind.p, (uint64_t)n * kWords);
    hipLaunchKernelGGL(k_zero_u64, dim3(fg), dim3(256), 0, st, B.cand.p, (uint64_t)n * kWords);
    hipLaunchKernelGGL(k_bp_masks, dim3(fg), dim3(256), 0, st, n, idx->bpd.p, (const ulonglong2 *)idx->bps.p,
                       B.lmask.p);
    hipLaunchKernelGGL(k_hub_info_init, dim3(fg), dim3(256), 0, st, B.hinfo.p, (uint64_t)n);
    hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, st, B.d1len.p, (uint64_t)n, 0u);
    hipLaunchKernelGGL(k_zero_u64, dim3(fg), dim3(256), 0, st, reinterpret_cast<unsigned long long *>(B.lrec.p),
                       2ull * n);
    hipLaunchKernelGGL(k_zero_u64, dim3(fg), dim3(256), 0, st, B.abits.p, 2ull * nw);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t *>(B.ctr.p),
                       (uint64_t)(sizeof(BuildCounters) / 4), 0u);
    hipLaunchKernelGGL(k_ctr_shape, dim3(1), dim3(1), 0, st, B.ctr.p, n, nw);
    PADO_HIP(hipGetLastError());
    return PADO_GPU_OK;
}

// End of the batch phase: clear the indicator rows still marked, keep the
// counters, and build the canonical rank-space CSR -- each vertex's labels
// gathered and sorted by hub (order_labels' per-vertex std::sort, PARA:2240).
template <bool W>
static pado_gpu_status batch_finish(pado_gpu_graph *g, pado_gpu_index *idx, const pado_gpu_build_cfg &cfg,
                                    BatchBuffers<W> &B, const KernelTimes &kt)
{
    hipStream_t st = g->ctx->stream;
    const uint32_t n = g->n;
    hipLaunchKernelGGL(k_batch_end, dim3(1024), dim3(256), 0, st, B.ind.p, B.tlist.p, B.abits.p, B.ctr.p);
    BuildCounters c{};
    PADO_HIP(copy_sync(&c, B.ctr.p, sizeof(BuildCounters), hipMemcpyDeviceToHost, st));
    idx->counters = c;
    idx->iterations = c.iterations;
    idx->ktimes = kt;
    PADO_TRY(idx->loff.alloc((uint64_t)n + 1));
    DevBuf<uint32_t> lens;  // label counts (lives until the end: the scan is asynchronous)
    PADO_TRY(lens.alloc(n));
    hipLaunchKernelGGL(k_cap_for_len, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, B.lrec.p, lens.p, 1u);
    PADO_TRY(scan_u32_to_u64(lens.p, idx->loff.p, n, st));
    uint64_t total = 0;
    PADO_HIP(copy_sync(&total, idx->loff.p + n, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    idx->num_labels = total;
    PADO_TRY(idx->lab.alloc(total));
    hipLaunchKernelGGL(k_gather_labels, dim3(grid_for((uint64_t)n * 64, 256, 8192)), dim3(256), 0, st, n, B.pool.p,
                       B.lrec.p, idx->loff.p, idx->lab.p);
    PADO_TRY(segmented_sort_u32(idx->lab.p, total, idx->loff.p, n, st, cfg.sort_max_items));
    return PADO_GPU_OK;
}

// The batch kernels of one (chunk, push waves) choice.
template <bool W>
struct BatchKernels {
    decltype(&k_batch_init<W, kChunk>) init;
    decltype(&k_push<W, kChunk, kPushWaves>) push;
    decltype(&k_add<W, kChunk, false>) add;       // iterations up to kNear + 1
    decltype(&k_add<W, kChunk, true>) add_deep;   // the later ones
    uint32_t chunk, waves;
};

template <bool W, uint32_t CH>
static BatchKernels<W> batch_kernels_ch(uint32_t waves)
{
    BatchKernels<W> K{k_batch_init<W, CH>, k_push<W, CH, kPushWaves>, k_add<W, CH, false>, k_add<W, CH, true>, CH,
                      kPushWaves};
    if (waves == 2) K.push = k_push<W, CH, 2>;
    if (waves == 8) K.push = k_push<W, CH, 8>;
    if (waves == 2 || waves == 8) K.waves = waves;
    return K;
}

template <bool W>
static BatchKernels<W> batch_kernels(uint32_t chunk, uint32_t waves)
{
    if (chunk == 4) return batch_kernels_ch<W, 4>(waves);
    if (chunk == 8) return batch_kernels_ch<W, 8>(waves);
    if (chunk == 32) return batch_kernels_ch<W, 32>(waves);
    return batch_kernels_ch<W, kChunk>(waves);
}

template <bool W>
static pado_gpu_status build_batches_t(pado_gpu_graph *g, pado_gpu_index *idx, const DevBuf<uint8_t> &d_used,
                                       const pado_gpu_build_cfg &cfg, unsigned long long pool_min,
                                       unsigned long long *pool_exhausted)
{
    using DT = typename Hop<W>::Dist;
    constexpr uint32_t kIterEnd = Hop<W>::kInf;  // the first iteration that is never run
    hipStream_t st = g->ctx->stream;
    const uint32_t n = g->n;
    const uint32_t batch = cfg.batch_size ? cfg.batch_size : kBatch;
    if (batch == 0 || batch > kBatch) {
        set_error("batch_size must be in [1, 1024]");
        return PADO_GPU_ERR_ARG;
    }
    if (n > Hop<W>::kMaxN) {
        set_error("graph too large for the wide hop domain (at most 2^20 vertices)");
        return PADO_GPU_ERR_DOMAIN;
    }
    const uint32_t chunk = cfg.push_chunk ? cfg.push_chunk : kPushChunk;
    const uint32_t waves = cfg.push_waves ? cfg.push_waves : kPushWaves;
    if ((chunk != 4 && chunk != 8 && chunk != 16 && chunk != 32) || (waves != 2 && waves != 4 && waves != 8)) {
        set_error("push_chunk must be 4, 8, 16 or 32 and push_waves 2, 4 or 8");
        return PADO_GPU_ERR_ARG;
    }
    const BatchKernels<W> K = batch_kernels<W>(chunk, waves);
    const std::string domain_msg =
        W ? "batch iteration reached 4095: distances outside the wide (12-bit) hop domain"
          : "batch iteration reached 255: distances outside the uint8 domain of the reference";

    // Per-wave reservations in k_add leave holes in the active list and the
    // descriptor array (see batch_alloc).
    const uint64_t add_waves = 4ull * (cfg.add_blocks ? cfg.add_blocks : default_add_grid(n));
    // The label pool starts at 64 words per vertex and doubles whenever a
    // host check finds it over 40% full; the checks come every
    // iters_per_sync iterations, so the headroom covers several iterations.
    // The root-distance table covers the hubs of the first 2^16 ids and
    // grows on the host before a batch that could exceed it.
    BatchBuffers<W> B;
    PADO_TRY(batch_alloc<W>(g, idx, add_waves, chunk,
                            std::max<unsigned long long>(std::max<unsigned long long>(1ull << 22, 64ull * n), pool_min),
                            std::min<uint32_t>(n, 1u << 16), B));
    auto &ind = B.ind, &cand = B.cand, &ag = B.ag, &chunks = B.chunks, &abits = B.abits;
    unsigned long long *vmark = B.lmask.p + kLmMark;  // word kLmMark of each mask row (stride kLmWords)
    auto &lmask = B.lmask, &rmask = B.rmask;
    auto &tlist = B.tlist, &clist = B.clist, &amask = B.amask, &aend = B.aend, &pool = B.pool;
    auto &hinfo = B.hinfo;
    auto &lrec = B.lrec;
    auto &dc = B.dc;
    auto &rset = B.rset;
    auto &ctr = B.ctr;
    unsigned long long pool_size = B.pool_size;
    uint32_t dc_rows = B.dc_rows;
    const uint32_t fg = 4096;

    BuildCounters *h_ctr = nullptr;
    PADO_HIP(hipHostMalloc((void **)&h_ctr, 4 * sizeof(BuildCounters), hipHostMallocDefault));
    struct PinnedFree {
        BuildCounters *p;
        ~PinnedFree()
        {
            std::lock_guard<std::recursive_mutex> g(capture_mutex());
            (void)hipHostFree(p);
        }
    } pin{h_ctr};

    const bool use_graph = (cfg.flags & PADO_GPU_BUILD_STREAM) == 0;
    const uint32_t U = cfg.iters_per_sync ? std::min<uint32_t>(cfg.iters_per_sync, 64u) : 8u;
    const uint32_t push_grid = cfg.push_blocks ? cfg.push_blocks : default_push_grid(n);
    const uint32_t add_grid = cfg.add_blocks ? cfg.add_blocks : default_add_grid(n);
    const uint32_t root_blocks = (batch * 64 + 255) / 256;
    const uint32_t pull_ratio = cfg.pull_ratio ? cfg.pull_ratio : kPullRatio;

    // Stream mode brackets every push and add launch with HIP events (read
    // after the host check that follows each block of U iterations).
    std::vector<hipEvent_t> ev;
    struct EventFree {
        std::vector<hipEvent_t> *e;
        ~EventFree()
        {
            for (auto x : *e) (void)hipEventDestroy(x);
        }
    } evf{&ev};
    if (!use_graph) {
        ev.resize(4 * U);
        for (auto &e : ev) PADO_HIP(hipEventCreate(&e));
    }
    KernelTimes kt{};

    LabelStore S{pool.p, lrec.p, pool_size};
    const HubList1 H1{B.d1.p, B.d1len.p};
    auto enqueue_prologue = [&]() {
        // Previous batch's reset_at_end first (a no-op before the first batch).
        hipLaunchKernelGGL(k_batch_close<W>, dim3(std::max<uint32_t>(1024u, root_blocks)), dim3(256), 0, st, d_used.p,
                           S, hinfo.p, dc.p, ind.p, tlist.p, abits.p, ctr.p);
        hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(batch), 0, st, d_used.p, ctr.p, batch, n);
        hipLaunchKernelGGL(K.init, dim3((batch + 3) / 4), dim3(256), 0, st, d_used.p, g->off.p, g->adj.p, S,
                           hinfo.p, dc.p, ind.p, vmark, tlist.p, abits.p, ag.p, amask.p, aend.p, chunks.p,
                           (const ulonglong2 *)idx->bps.p, rset.p, lmask.p, rmask.p, ctr.p);
    };
    // Iterations first + 1 .. first + count of the current batch.
    auto enqueue_iterations = [&](uint32_t first, uint32_t count, bool timed) -> pado_gpu_status {
        for (uint32_t u = 0; u < count; ++u) {
            const uint32_t iter = first + 1 + u;
            if (timed) PADO_HIP(hipEventRecord(ev[4 * u], st));
            hipLaunchKernelGGL(K.push, dim3(push_grid), dim3(64 * K.waves), 0, st, iter, g->off.p, g->adj.p, abits.p,
                               ag.p, amask.p, aend.p, chunks.p, ind.p, cand.p, vmark, tlist.p, clist.p,
                               idx->bpd.p, lmask.p, rmask.p, (const ulonglong2 *)idx->bps.p, rset.p, ctr.p,
                               idx->bp_pass_through ? 0u : 1u, pull_ratio);
            if (timed) PADO_HIP(hipEventRecord(ev[4 * u + 1], st));
            if (timed) PADO_HIP(hipEventRecord(ev[4 * u + 2], st));
            hipLaunchKernelGGL(iter > kNear + 1 ? K.add_deep : K.add, dim3(add_grid), dim3(256), 0, st, iter, d_used.p,
                               g->off.p, g->adj.p, clist.p,
                               cand.p, S, H1, hinfo.p, dc.p, abits.p, ag.p, amask.p, aend.p, chunks.p, ctr.p);
            if (timed) PADO_HIP(hipEventRecord(ev[4 * u + 3], st));
        }
        return PADO_GPU_OK;
    };

    // A block of k <= U iterations is enqueued per host check.  A batch's
    // first block has as many iterations as the recent batches needed (see
    // iters_hist below); a batch that needs more continues in blocks of 2,
    // 4, 8, ... iterations.
    // Graph mode: one captured graph per (first iteration, count), the batch
    // prologue included when first == 0; captured lazily and replayed for
    // every batch.  Kernel arguments are fixed per capture, so all graphs are
    // dropped when the pool or the root-distance table moves.
    std::map<std::pair<uint32_t, uint32_t>, hipGraphExec_t> graphs;
    auto drop_graphs = [&]() {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    };
    struct GraphFree {
        std::function<void()> f;
        ~GraphFree() { f(); }
    } gxf{drop_graphs};
    auto capture = [&](uint32_t first, uint32_t count, hipGraphExec_t *out) -> pado_gpu_status {
        // No other thread of this process allocates or frees while the
        // stream captures (see capture_mutex).
        std::lock_guard<std::recursive_mutex> cap_lock(capture_mutex());
        hipGraph_t gr = nullptr;
        PADO_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        if (first == 0) enqueue_prologue();
        pado_gpu_status s = enqueue_iterations(first, count, false);
        hipError_t e = hipStreamEndCapture(st, &gr);
        if (s != PADO_GPU_OK) return s;
        if (e != hipSuccess) {
            set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            return PADO_GPU_ERR_HIP;
        }
        e = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
            set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            return PADO_GPU_ERR_HIP;
        }
        return PADO_GPU_OK;
    };
    auto first_block = [&](uint32_t prev) { return std::max<uint32_t>(1u, std::min<uint32_t>(prev, U)); };

    // Counter snapshots: each enqueued block of iterations is followed by a
    // copy of the counters into its own pinned slot and an event, so the host
    // reads a block's state while later blocks run.
    constexpr uint32_t kSnaps = 4;  // pinned counter slots in h_ctr
    std::vector<hipEvent_t> sev(kSnaps, nullptr);
    struct SnapEventFree {
        std::vector<hipEvent_t> *e;
        ~SnapEventFree()
        {
            for (auto x : *e)
                if (x) (void)hipEventDestroy(x);
        }
    } sevf{&sev};
    for (auto &e : sev) PADO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    uint32_t snap_next = 0;
    auto snapshot = [&](uint32_t *slot) -> pado_gpu_status {
        *slot = snap_next;
        snap_next = (snap_next + 1) % kSnaps;
        PADO_HIP(hipMemcpyAsync(&h_ctr[*slot], ctr.p, sizeof(BuildCounters), hipMemcpyDeviceToHost, st));
        PADO_HIP(hipEventRecord(sev[*slot], st));
        return PADO_GPU_OK;
    };
    auto launch_block = [&](uint32_t first, uint32_t count, bool timed) -> pado_gpu_status {
        if (use_graph) {
            hipGraphExec_t &gx = graphs[std::make_pair(first, count)];
            if (!gx) PADO_TRY(capture(first, count, &gx));
            PADO_HIP(hipGraphLaunch(gx, st));
        } else {
            if (first == 0) enqueue_prologue();
            PADO_TRY(enqueue_iterations(first, count, timed));
        }
        return PADO_GPU_OK;
    };

    // Graph mode runs one batch ahead when the label pool has ample room: the
    // next batch's first block is enqueued before the host has checked the
    // current one, and its gate makes it a no-op if the current batch turns
    // out to need more iterations (the host then clears the gate, runs them,
    // and enqueues the next batch again).
    const uint32_t nb = (n + batch - 1) / batch;
    uint64_t iterations_seen = 0;
    uint32_t prev_batch_iters = 1;
    constexpr uint32_t kIterHistory = 8;
    uint32_t iters_hist[kIterHistory] = {1, 1, 1, 1, 1, 1, 1, 1}, hist_pos = 0;
    bool ahead = false;               // batch b's first block is already enqueued
    uint32_t ahead_count = 0, ahead_slot = 0;
    unsigned long long top_prev = 0, growth = ~0ull;  // pool use after the last batch, and its growth
    // Pool words taken per iteration, from the last checked block (a recent
    // maximum), and the pool top at that check: the headroom kept before the
    // next blocks scales with it, so a fast-growing phase (labels-only
    // pruning, early batches) cannot run the pool dry inside a block.
    double rate = 0.0;
    unsigned long long snap_top_prev = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t batch_end = std::min<uint64_t>((uint64_t)(b + 1) * batch, n);
        if (batch_end > dc_rows) {
            // Never with a block ahead: a batch is only enqueued early when the
            // table already covers it.  hipFree/hipMalloc synchronise.  The new
            // table starts all 255, so the previous batch's reset is a no-op on it.
            PADO_HIP(hipStreamSynchronize(st));
            dc_rows = std::min<uint32_t>(n, std::max<uint32_t>(batch_end, dc_rows * 2));
            PADO_TRY(dc.alloc((uint64_t)dc_rows * kBatch));
            hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, st, reinterpret_cast<uint32_t *>(dc.p),
                               (uint64_t)dc_rows * kBatch * sizeof(DT) / 4, ~0u);
            drop_graphs();
        }
        const uint64_t batch_first_iter = iterations_seen;
        uint32_t next_count = std::min<uint32_t>(2u, U);  // a continuation block's size
        bool repacked = false;
        unsigned long long last_top = 0;
        uint32_t launched, cur_slot, cur_count;  // cur_count: iterations of the block being checked
        if (ahead) {
            launched = cur_count = ahead_count;
            cur_slot = ahead_slot;
            ahead = false;
        } else {
            // Iteration kIterEnd is the last one ever launched: k_push flags
            // it if the frontier is still live there.
            launched = cur_count = std::min<uint32_t>(first_block(prev_batch_iters), kIterEnd);
            PADO_TRY(launch_block(0, launched, true));
            PADO_TRY(snapshot(&cur_slot));
        }
        const uint32_t next_end = std::min<uint64_t>((uint64_t)(b + 2) * batch, n);
        bool spec = use_graph && b + 1 < nb && next_end <= dc_rows && growth != ~0ull &&
                    (top_prev + 4 * growth) * 5 < pool_size * 2;
        if (spec) {
            ahead_count = std::min<uint32_t>(first_block(prev_batch_iters), kIterEnd);
            PADO_TRY(launch_block(0, ahead_count, false));
            PADO_TRY(snapshot(&ahead_slot));
        }
        for (;;) {
            PADO_HIP(hipEventSynchronize(sev[cur_slot]));
            PADO_HIP(hipGetLastError());
            const BuildCounters snap = h_ctr[cur_slot];
            last_top = snap.pool_top;
            if (!use_graph) {
                const uint64_t live = snap.iterations - iterations_seen;
                for (uint64_t u = 0; u < live && u < cur_count; ++u) {
                    float t_push = 0.f, t_add = 0.f;
                    PADO_HIP(hipEventElapsedTime(&t_push, ev[4 * u], ev[4 * u + 1]));
                    PADO_HIP(hipEventElapsedTime(&t_add, ev[4 * u + 2], ev[4 * u + 3]));
                    kt.ms_push += t_push;
                    kt.ms_add += t_add;
                    kt.n_push += 1;
                    kt.n_add += 1;
                }
            }
            iterations_seen = snap.iterations;
            // Work enqueued after this block (a block ahead) still uses the
            // buffers: it drains before an error return frees them.
            if (snap.overflow) {
                (void)hipStreamSynchronize(st);
                if (pool_exhausted) *pool_exhausted = pool_size;
                set_error("label pool exhausted");
                return PADO_GPU_ERR_NOMEM;
            }
            if (snap.domain) {
                (void)hipStreamSynchronize(st);
                set_error(domain_msg);
                return PADO_GPU_ERR_DOMAIN;
            }
            if (snap.pool_top >= snap_top_prev && cur_count) {
                const double r = (double)(snap.pool_top - snap_top_prev) / cur_count;
                rate = std::max(r, rate * 0.75);
            }
            snap_top_prev = snap.pool_top;
            // Keep at least 60% of the pool free, and room for four times the
            // recent growth over the next blocks (this batch's next block and
            // a block run ahead: U iterations each at most); a block ahead
            // runs to completion first.
            const double need = rate * 4.0 * (spec ? 2.0 * U : (double)U);
            if (snap.pool_top * 5 > pool_size * 2 || (double)snap.pool_top + need > 0.8 * (double)pool_size) {
                PADO_HIP(hipStreamSynchronize(st));
                unsigned long long top = 0;
                const unsigned long long want =
                    std::max<unsigned long long>(pool_size * 2, (unsigned long long)(2.0 * (snap.pool_top + need)));
                PADO_TRY(pool_repack(st, n, pool, lrec, want, &top));
                snap_top_prev = top;
                pool_size = pool.n;
                S.pool = pool.p;
                S.pool_size = pool_size;
                PADO_HIP(hipMemcpyAsync(&ctr.p->pool_top, &top, sizeof(top), hipMemcpyHostToDevice, st));
                PADO_HIP(hipStreamSynchronize(st));
                drop_graphs();
                repacked = true;
                // The block ahead's snapshot predates the repack: take it again.
                if (spec) PADO_TRY(snapshot(&ahead_slot));
            }
            // The frontier of the next iteration is in slot (launched + 1) & 1.
            if (snap.n_act[(launched + 1) & 1u] == 0) break;
            if (launched >= kIterEnd) {
                (void)hipStreamSynchronize(st);
                set_error(domain_msg);
                return PADO_GPU_ERR_DOMAIN;
            }
            // The batch continues: a block enqueued ahead was gated off.
            if (spec) {
                hipLaunchKernelGGL(k_unhold, dim3(1), dim3(1), 0, st, ctr.p);
                spec = false;
            }
            cur_count = std::min<uint32_t>(next_count, kIterEnd - launched);
            next_count = std::min<uint32_t>(2 * next_count, U);  // the batch runs long: double the block
            PADO_TRY(launch_block(launched, cur_count, true));
            PADO_TRY(snapshot(&cur_slot));
            launched += cur_count;
        }
        // The first block of the next batches: the most iterations any of the
        // last kIterHistory batches needed.  A batch that needs more than its
        // first block waits for a host round trip (and its successor, enqueued
        // ahead, is gated off and enqueued again); an iteration past a batch's
        // end costs two empty launches.  On R-MAT 20 this policy takes 9 such
        // round trips and 282 empty iterations where the previous batch's count
        // alone takes 81 and 86.
        iters_hist[hist_pos++ % kIterHistory] = (uint32_t)(iterations_seen - batch_first_iter);
        prev_batch_iters = *std::max_element(iters_hist, iters_hist + kIterHistory);
        // Pool growth of this batch (unknown across a repack, which blocks
        // running ahead until a batch has been measured again).
        if (repacked) {
            growth = ~0ull;
            top_prev = 0;
        } else {
            const unsigned long long top_now = last_top;
            growth = top_now > top_prev ? top_now - top_prev : 0ull;
            top_prev = top_now;
        }
        ahead = spec;
    }
    return batch_finish<W>(g, idx, cfg, B, kt);
}

template <bool W>
static BatchPlan<W> batch_plan(const pado_gpu_graph *g, const pado_gpu_index *idx, const uint8_t *used, uint32_t batch,
                               BatchBuffers<W> &B)
{
    BatchPlan<W> P{};
    P.n = g->n;
    P.batch = batch;
    P.bp_prune = idx->bp_pass_through ? 0u : 1u;
    P.used = used;
    P.off = g->off.p;
    P.adj = g->adj.p;
    P.S = LabelStore{B.pool.p, B.lrec.p, B.pool_size};
    P.H1 = HubList1{B.d1.p, B.d1len.p};
    P.ind = B.ind.p;
    P.cand = B.cand.p;
    P.vmark = B.lmask.p + kLmMark;
    P.tlist = B.tlist.p;
    P.clist = B.clist.p;
    P.abits = B.abits.p;
    P.pull_ratio = ~0u;  // the one-workgroup loop always pushes (its push_body has no pull form)
    P.ag = B.ag.p;
    P.amask = B.amask.p;
    P.aend = B.aend.p;
    P.chunks = B.chunks.p;
    P.bpd = idx->bpd.p;
    P.lmask = B.lmask.p;
    P.rmask = B.rmask.p;
    P.bps = (const ulonglong2 *)idx->bps.p;
    P.rset = B.rset.p;
    P.hinfo = B.hinfo.p;
    P.dc = B.dc.p;
    P.C = B.ctr.p;
    return P;
}

// The batch phase of jobs[i] for the i with jobs[i].idx->wide == W, all in
// one k_batch_loop launch on st.  The pool cannot grow inside the loop: a
// job whose pool runs dry ends with PADO_GPU_ERR_NOMEM and pool_exhausted.
template <bool W>
static pado_gpu_status batch_loop_t(std::vector<LoopJob> &jobs, hipStream_t st)
{
    std::vector<size_t> mine;
    for (size_t i = 0; i < jobs.size(); ++i)
        if ((jobs[i].idx->wide != 0) == W && jobs[i].status == PADO_GPU_OK) mine.push_back(i);
    if (mine.empty()) return PADO_GPU_OK;
    std::vector<BatchBuffers<W>> bufs(mine.size());
    std::vector<BatchPlan<W>> plans(mine.size());
    for (size_t j = 0; j < mine.size(); ++j) {
        LoopJob &J = jobs[mine[j]];
        const uint32_t n = J.g->n;
        if (J.batch == 0 || J.batch > kBatch || n > Hop<W>::kMaxN || n == 0) {
            set_error("device-loop batch phase: batch size or vertex count out of range");
            return PADO_GPU_ERR_ARG;
        }
        // Labels per vertex of deep graphs stay in the low hundreds; the
        // growable per-vertex arrays take up to four times their length.
        const unsigned long long pool =
            std::max<unsigned long long>(std::max<unsigned long long>(1ull << 22, 384ull * n), J.pool_min);
        PADO_TRY(batch_alloc<W>(J.g, J.idx, kLoopWaves, kChunk, pool, n, bufs[j]));
        plans[j] = batch_plan<W>(J.g, J.idx, J.used, J.batch, bufs[j]);
    }
    DevBuf<BatchPlan<W>> dplans;
    PADO_TRY(dplans.alloc(plans.size()));
    PADO_HIP(hipMemcpyAsync(dplans.p, plans.data(), sizeof(BatchPlan<W>) * plans.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_batch_loop<W>, dim3((uint32_t)plans.size()), dim3(kLoopWaves * 64), 0, st, dplans.p);
    PADO_HIP(hipGetLastError());
    PADO_HIP(hipStreamSynchronize(st));
    for (size_t j = 0; j < mine.size(); ++j) {
        LoopJob &J = jobs[mine[j]];
        BuildCounters c{};
        PADO_HIP(copy_sync(&c, bufs[j].ctr.p, sizeof(BuildCounters), hipMemcpyDeviceToHost, st));
        if (c.overflow) {
            J.status = PADO_GPU_ERR_NOMEM;
            J.pool_exhausted = bufs[j].pool_size;
            continue;
        }
        if (c.domain) {
            J.status = PADO_GPU_ERR_DOMAIN;
            continue;
        }
        pado_gpu_build_cfg cfg{};
        cfg.batch_size = J.batch;
        J.status = batch_finish<W>(J.g, J.idx, cfg, bufs[j], KernelTimes{});
        bufs[j].release();  // this graph's buffers go before the next finish allocates
    }
    return PADO_GPU_OK;
}

pado_gpu_status build_batches_loop(std::vector<LoopJob> &jobs, hipStream_t st)
{
    PADO_TRY(batch_loop_t<false>(jobs, st));
    return batch_loop_t<true>(jobs, st);
}

pado_gpu_status build_batches(pado_gpu_graph *g, pado_gpu_index *idx, const DevBuf<uint8_t> &d_used,
                              const pado_gpu_build_cfg &cfg)
{
    // A pool that still runs dry inside a block (growth far above the recent
    // rate) costs one rerun of the batch phase from scratch with a larger
    // pool; no label is ever dropped.
    unsigned long long exhausted = 0, pool_min = 0;
    if ((cfg.flags & PADO_GPU_BUILD_DEVICE_LOOP) && !(cfg.flags & PADO_GPU_BUILD_STREAM)) {
        // One workgroup runs the whole batch phase; a pool that runs dry
        // there falls back to the launch schedule, which grows its pool.
        std::vector<LoopJob> job(1);
        job[0].g = g;
        job[0].idx = idx;
        job[0].used = d_used.p;
        job[0].batch = cfg.batch_size ? cfg.batch_size : kBatch;
        PADO_TRY(build_batches_loop(job, g->ctx->stream));
        if (job[0].status != PADO_GPU_ERR_NOMEM) return job[0].status;
        pool_min = job[0].pool_exhausted * 2;
    }
    for (int attempt = 0;; ++attempt) {
        exhausted = 0;
        const pado_gpu_status s = idx->wide ? build_batches_t<true>(g, idx, d_used, cfg, pool_min, &exhausted)
                                            : build_batches_t<false>(g, idx, d_used, cfg, pool_min, &exhausted);
        if (s != PADO_GPU_ERR_NOMEM || !exhausted || attempt == 2) return s;
        pool_min = exhausted * 4;
        (void)hipStreamSynchronize(g->ctx->stream);
    }
}

}  // namespace pado
