// query.hip -- exact distance queries over a PADO index on gfx950.
//
// Reference: query_distance, includes/pado_unw_unv_para.201912311607.
// parallel_bp_labeling.h:2283-2416:
//   d = min over the 50 bit-parallel slots of  da + db - {2 | 1 | 0}
//       (2 if S^-1 sets intersect, 1 if S^-1 meets S^0, else 0),
//   d = min(d, min over common hubs h of  La(h) + Lb(h)),
//   d >= 253 -> 255;  ids out of range -> (a == b ? 0 : 255).
// The reference gates each slot with `int td; inti d; if (td - 2 <= d)`,
// an unsigned comparison that skips slots with td < 2 and so over-estimates
// distances from a bit-parallel root to itself and to its neighbours.  The
// default here is the exact (signed) comparison; PADO_GPU_QUERY_REF_COMPAT
// reproduces the reference's answers bit for bit.  Either way the slot order
// does not matter: the result is a minimum over slots.
//
// Mapping: one wave per pair.  Lanes 0..49 take one bit-parallel slot each:
// first the distance bytes only, then, after a wave minimum m of the slot
// sums, the set pairs of just the slots that could still go below m (sum
// <= m + 1).  The wave intersects the two hub-sorted label lists: the longer
// list is copied into the wave's LDS (up to kQLds labels) and lanes take 64
// labels of the shorter list at a time and binary-search them there, so the
// search's dependent steps are LDS reads instead of global ones.  A wave of
// the batch kernels answers pairs i, i + W, i + 2W, ... (W = waves in the
// grid) as a software pipeline: while pair i is answered, pair i + W's label
// offsets and pair i + 2W's ids are already in flight (1.37 -> 1.45 G
// queries/s on R-MAT 20 uniform pairs; loading the next pair's bit-parallel
// bytes early as well measured slower).  Within a pair, every read that needs
// only the label offsets is issued first (distance bytes, the shorter list's
// first 64 labels, the longer list's first 256), the set pairs are requested
// next and the label search runs while they are in flight.  k_query is held
// to 64 VGPRs (8 waves per SIMD instead of 7): 1.45 -> 1.98 G queries/s for
// both changes (without the cap, 128 early labels per wave measured best).
#include "internal.hpp"

namespace pado {
namespace {

__device__ inline uint32_t wave_min_u32(uint32_t x)
{
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t y = __shfl_xor(x, o);
        x = y < x ? y : x;
    }
    return x;
}

// Orders LDS writes of some lanes of a wave before reads by other lanes.
__device__ inline void q_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

constexpr uint32_t kQLds = 1024;   // labels of the longer list staged in LDS per wave
constexpr uint32_t kServeLds = 768;  // the same in the resident serving kernel
constexpr uint32_t kQWaves = 4;    // waves per query block
constexpr uint32_t kQPre = 4;      // longer-list loads per lane issued with the pair's first reads
constexpr uint32_t kWPre = 4;      // the same for weighted pairs (8 measured slower)

}  // namespace

// Index of the first entry of the sorted list L[0..len) whose hub is >= key
// (labels compared by hub, the bits above the distance field).
template <bool W>
__device__ inline uint32_t lower_hub(const uint32_t *L, uint32_t len, uint32_t key)
{
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((L[mid] >> Hop<W>::kBits) < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// One pair whose label offsets (oa, la), (ob, lb) are already loaded.
// sbuf: the wave's kQLds-entry LDS buffer.  Hop<false>: the reference's
// answers (255 for unreachable or >= 253); Hop<true>: exact hops, 0xFFFF for
// unreachable, using only bit-parallel slots reached on both sides.
template <bool W, uint32_t LDS = kQLds>
__device__ inline uint32_t pair_distance_at(uint32_t n, const uint8_t *bpd, const ulonglong2 *bps,
                                            const uint32_t *lab, uint32_t a, uint32_t b, uint64_t oa, uint32_t la,
                                            uint64_t ob, uint32_t lb, bool ref_compat, uint32_t *sbuf)
{
    constexpr uint32_t B = Hop<W>::kBits, MASK = Hop<W>::kInf;
    constexpr uint32_t UNREACHED = W ? 0xFFFFu : 255u;
    const uint32_t lane = threadIdx.x & 63;
    if (a >= n || b >= n) return a == b ? 0u : UNREACHED;
    // Every load that needs only the offsets goes out first: the slots'
    // distance bytes, the shorter list's first 64 labels and the longer
    // list's first kQPre * 64 labels (into registers; written to LDS after
    // the set pairs are requested), so the pair's chain of dependent global
    // reads is bytes -> set pairs, with the label search in between.
    uint32_t da = 0, db = 0;
    if (lane < kBP) {
        da = bpd[(uint64_t)a * kBpdStride + lane];
        db = bpd[(uint64_t)b * kBpdStride + lane];
    }
    const uint32_t *S = lab + (la <= lb ? oa : ob);
    const uint32_t *T = lab + (la <= lb ? ob : oa);
    const uint32_t ls = la <= lb ? la : lb, lt = la <= lb ? lb : la;
    const bool staged = lt <= LDS;
    const uint32_t s0 = lane < ls ? S[lane] : 0u;
    uint32_t tv[kQPre];
    if (staged) {
#pragma unroll
        for (uint32_t u = 0; u < kQPre; ++u) {
            const uint32_t j = u * 64 + lane;
            tv[u] = j < lt ? T[j] : 0u;
        }
    }
    // Bit-parallel slots.  A slot counts if the reference's gate passes; its
    // result is td - {2, 1, 0}, so only slots with td <= m + 1 (m = the
    // smallest counted td) can beat m.
    int td = 0x7FFF;
    bool gate = false;
    if (lane < kBP) {
        td = (int)(da + db);
        if (W)
            gate = da != kInf8 && db != kInf8;  // 255: the slot's BFS did not reach this vertex
        else
            gate = !(ref_compat && td < 2) && td - 2 <= 255;
    }
    const uint32_t m = wave_min_u32(gate ? (uint32_t)td : 0xFFFFu);
    const bool use_sets = gate && (uint32_t)td <= m + 1u;
    ulonglong2 sa = {0, 0}, sb = {0, 0};
    if (use_sets) {
        sa = bps[(uint64_t)a * kBP + lane];
        sb = bps[(uint64_t)b * kBP + lane];
    }
    if (staged) {
#pragma unroll
        for (uint32_t u = 0; u < kQPre; ++u) {
            const uint32_t j = u * 64 + lane;
            if (j < lt) sbuf[j] = tv[u];
        }
        for (uint32_t j = kQPre * 64 + lane; j < lt; j += 64) sbuf[j] = T[j];
        q_lds_sync();
    }
    uint32_t d = UNREACHED;
    for (uint32_t j0 = 0; j0 < ls; j0 += 64) {
        const uint32_t j = j0 + lane;
        if (j < ls) {
            const uint32_t s = j0 == 0 ? s0 : S[j];
            const uint32_t key = s >> B;
            const uint32_t *L = staged ? sbuf : T;
            const uint32_t lo = lower_hub<W>(L, lt, key);
            if (lo < lt && (L[lo] >> B) == key) {
                const uint32_t t = (s & MASK) + (L[lo] & MASK);
                if (t < d) d = t;
            }
        }
    }
    // The next pair's copy must not overtake this pair's reads.
    if (staged) q_lds_sync();
    if (use_sets) {
        int t = td;
        if (sa.x & sb.x)
            t -= 2;
        else if ((sa.x & sb.y) | (sa.y & sb.x))
            t -= 1;
        if ((uint32_t)t < d) d = (uint32_t)t;
    }
    d = wave_min_u32(d);
    if (W) return d;
    return d >= 253u ? 255u : d;
}

template <bool W>
__device__ inline uint32_t pair_distance(uint32_t n, const uint8_t *bpd, const ulonglong2 *bps, const uint64_t *loff,
                                         const uint32_t *lab, uint32_t a, uint32_t b, bool ref_compat,
                                         uint32_t *sbuf)
{
    if (a >= n || b >= n) return a == b ? 0u : (W ? 0xFFFFu : 255u);
    const uint64_t oa = loff[a], ob = loff[b];
    return pair_distance_at<W>(n, bpd, bps, lab, a, b, oa, (uint32_t)(loff[a + 1] - oa), ob,
                               (uint32_t)(loff[b + 1] - ob), ref_compat, sbuf);
}

// A pair's ids and label offsets, loaded ahead of its turn.
struct QPair {
    uint32_t a, b;
    uint64_t oa, ea, ob, eb;
};

__device__ inline void q_load_ids(QPair &p, const uint32_t *qa, const uint32_t *qb, uint64_t i, uint64_t k)
{
    p.a = 0;
    p.b = 0;
    if (i < k) {
        p.a = qa[i];
        p.b = qb[i];
    }
}

__device__ inline void q_load_offsets(QPair &p, uint32_t n, const uint64_t *loff)
{
    p.oa = p.ea = p.ob = p.eb = 0;
    if (p.a < n && p.b < n) {
        p.oa = loff[p.a];
        p.ea = loff[p.a + 1];
        p.ob = loff[p.b];
        p.eb = loff[p.b + 1];
    }
}

// OT: the answer type (u8, or u16 for a wide index or widened answers).
template <bool W, class OT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_query(uint32_t n, const uint8_t *bpd, const ulonglong2 *bps,
                                               const uint64_t *loff, const uint32_t *lab, const uint32_t *qa,
                                               const uint32_t *qb, uint64_t k, OT *out, int ref_compat)
{
    __shared__ uint32_t s_q[kQWaves][kQLds];
    uint32_t *sbuf = s_q[threadIdx.x >> 6];
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= k) return;
    // Software pipeline over the wave's pairs i, i + nw, ...: while pair i is
    // answered, pair i + nw's label offsets and pair i + 2nw's ids are in
    // flight, so a pair's dependent chain starts at its bit-parallel bytes.
    QPair cur, nxt, far;
    q_load_ids(cur, qa, qb, i, k);
    q_load_offsets(cur, n, loff);
    q_load_ids(nxt, qa, qb, i + nw, k);
    for (;;) {
        q_load_offsets(nxt, n, loff);
        q_load_ids(far, qa, qb, i + 2 * nw, k);
        const uint32_t d = pair_distance_at<W>(n, bpd, bps, lab, cur.a, cur.b, cur.oa, (uint32_t)(cur.ea - cur.oa),
                                               cur.ob, (uint32_t)(cur.eb - cur.ob), ref_compat != 0, sbuf);
        if ((threadIdx.x & 63) == 0) out[i] = (OT)d;
        i += nw;
        if (i >= k) break;
        cur = nxt;
        nxt = far;
    }
}

// Weighted query_distance (WPARA:2125-2243): min over common hubs of
// La(h) + Lb(h).  Exact u32 answers, 0xFFFFFFFF when no hub is shared; with
// ref_compat the reference's int16 domain (>= 32765 -> 32767, ids out of
// range -> a == b ? 0 : 32767).  One wave per pair, the shorter list's hubs
// binary-searched in the longer one as for unweighted labels.
template <uint32_t LDS = kQLds>
__device__ inline uint32_t wpair_distance_at(uint32_t n, const uint32_t *hub, const uint32_t *dist, uint32_t a,
                                             uint32_t b, uint64_t oa, uint32_t la, uint64_t ob, uint32_t lb,
                                             bool ref_compat, uint32_t *sbuf)
{
    const uint32_t lane = threadIdx.x & 63;
    if (a >= n || b >= n) return a == b ? 0u : (ref_compat ? 32767u : 0xFFFFFFFFu);
    const uint64_t os = la <= lb ? oa : ob, ot = la <= lb ? ob : oa;
    const uint32_t ls = la <= lb ? la : lb, lt = la <= lb ? lb : la;
    // The longer list's hubs go to the wave's LDS (distances are read only
    // for matches); its first kWPre * 64 hubs and the shorter list's first 64
    // are requested together.
    const bool staged = lt <= LDS;
    const uint32_t k0 = lane < ls ? hub[os + lane] : 0u;
    if (staged) {
        uint32_t tv[kWPre];
#pragma unroll
        for (uint32_t u = 0; u < kWPre; ++u) {
            const uint32_t j = u * 64 + lane;
            tv[u] = j < lt ? hub[ot + j] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kWPre; ++u) {
            const uint32_t j = u * 64 + lane;
            if (j < lt) sbuf[j] = tv[u];
        }
        for (uint32_t j = kWPre * 64 + lane; j < lt; j += 64) sbuf[j] = hub[ot + j];
        q_lds_sync();
    }
    const uint32_t *H = staged ? sbuf : hub + ot;
    uint32_t d = 0xFFFFFFFFu;
    for (uint32_t j0 = 0; j0 < ls; j0 += 64) {
        const uint32_t j = j0 + lane;
        if (j < ls) {
            const uint32_t key = j0 == 0 ? k0 : hub[os + j];
            uint32_t lo = 0, hi = lt;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (H[mid] < key)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo < lt && H[lo] == key) {
                const uint64_t td = (uint64_t)dist[os + j] + dist[ot + lo];
                if (td < d) d = (uint32_t)td;
            }
        }
    }
    // The next pair's copy must not overtake this pair's reads.
    if (staged) q_lds_sync();
    d = wave_min_u32(d);
    if (ref_compat && d >= 32765u) d = 32767u;
    return d;
}

__device__ inline uint32_t wpair_distance(uint32_t n, const uint64_t *loff, const uint32_t *hub, const uint32_t *dist,
                                          uint32_t a, uint32_t b, bool ref_compat, uint32_t *sbuf)
{
    if (a >= n || b >= n) return a == b ? 0u : (ref_compat ? 32767u : 0xFFFFFFFFu);
    const uint64_t oa = loff[a], ob = loff[b];
    return wpair_distance_at(n, hub, dist, a, b, oa, (uint32_t)(loff[a + 1] - oa), ob, (uint32_t)(loff[b + 1] - ob),
                             ref_compat, sbuf);
}

__global__ __launch_bounds__(256) void k_wquery(uint32_t n, const uint64_t *loff, const uint32_t *hub,
                                                const uint32_t *dist, const uint32_t *qa, const uint32_t *qb,
                                                uint64_t k, uint32_t *out, int ref_compat)
{
    __shared__ uint32_t s_q[kQWaves][kQLds];
    uint32_t *sbuf = s_q[threadIdx.x >> 6];
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= k) return;
    // The same software pipeline as k_query.
    QPair cur, nxt, far;
    q_load_ids(cur, qa, qb, i, k);
    q_load_offsets(cur, n, loff);
    q_load_ids(nxt, qa, qb, i + nw, k);
    for (;;) {
        q_load_offsets(nxt, n, loff);
        q_load_ids(far, qa, qb, i + 2 * nw, k);
        const uint32_t d = wpair_distance_at(n, hub, dist, cur.a, cur.b, cur.oa, (uint32_t)(cur.ea - cur.oa), cur.ob,
                                             (uint32_t)(cur.eb - cur.ob), ref_compat != 0, sbuf);
        if ((threadIdx.x & 63) == 0) out[i] = d;
        i += nw;
        if (i >= k) break;
        cur = nxt;
        nxt = far;
    }
}

// Persistent label-merge kernels for the serving loop: a grid sized to the
// chip's resident waves; each wave takes kGrab pairs at a time from a
// per-launch work counter until the launch's k pairs are exhausted, so waves
// that draw long label lists do not hold back the launch.  Every wave leaves
// once the counter passes k.
constexpr uint32_t kGrab = 16;

__global__ __launch_bounds__(256) void k_query_ws(uint32_t n, const uint8_t *bpd, const ulonglong2 *bps,
                                                  const uint64_t *loff, const uint32_t *lab, const uint32_t *qa,
                                                  const uint32_t *qb, uint64_t k, uint8_t *out, int ref_compat,
                                                  unsigned long long *work)
{
    __shared__ uint32_t s_q[kQWaves][kQLds];
    uint32_t *sbuf = s_q[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63;
    for (;;) {
        unsigned long long c0 = 0;
        if (lane == 0) c0 = atomicAdd(work, (unsigned long long)kGrab);
        c0 = __shfl(c0, 0);
        if (c0 >= k) break;
        const uint64_t c1 = c0 + kGrab < k ? c0 + kGrab : k;
        for (uint64_t i = c0; i < c1; ++i) {
            const uint32_t d = pair_distance<false>(n, bpd, bps, loff, lab, qa[i], qb[i], ref_compat != 0, sbuf);
            if (lane == 0) out[i] = (uint8_t)d;
        }
    }
}

__global__ __launch_bounds__(256) void k_wquery_ws(uint32_t n, const uint64_t *loff, const uint32_t *hub,
                                                   const uint32_t *dist, const uint32_t *qa, const uint32_t *qb,
                                                   uint64_t k, uint32_t *out, int ref_compat,
                                                   unsigned long long *work)
{
    __shared__ uint32_t s_q[kQWaves][kQLds];
    uint32_t *sbuf = s_q[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63;
    for (;;) {
        unsigned long long c0 = 0;
        if (lane == 0) c0 = atomicAdd(work, (unsigned long long)kGrab);
        c0 = __shfl(c0, 0);
        if (c0 >= k) break;
        const uint64_t c1 = c0 + kGrab < k ? c0 + kGrab : k;
        for (uint64_t i = c0; i < c1; ++i) {
            const uint32_t d = wpair_distance(n, loff, hub, dist, qa[i], qb[i], ref_compat != 0, sbuf);
            if (lane == 0) out[i] = d;
        }
    }
}

__global__ void k_zero_work(unsigned long long *work) { *work = 0ull; }

// Resident workgroups per CU of the persistent kernel: the label merge waits
// on dependent loads, so the grid fills every slot the kernel's registers and
// LDS leave (a fixed 4 per CU left serving at 70% of the grid-stride kernel).
uint32_t query_ws_blocks_per_cu(bool weighted)
{
    int nb = 0;
    const hipError_t e = weighted ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_wquery_ws, 256, 0)
                                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_query_ws, 256, 0);
    return (e == hipSuccess && nb > 0) ? (uint32_t)nb : 4u;
}

// ---------------------------------------------------------------------------
// Resident serving kernel.  One launch stays on the device while requests
// flow: the host writes a request's pairs into the mapped ring, cuts them
// into descriptors of at most kServeGroup pairs spread over kServeQueues
// queues, and raises each queue's published count in mapped memory.  Wave 0
// of workgroup 0 polls those counts (one host read per queue per round) and
// mirrors them into device memory; every other wave takes tickets from its
// queue's device counter (queues are picked by workgroup, so the eight
// counters and their traffic spread over the XCDs) and waits, sleeping, on
// the mirrored count in L2.  A wave answers its descriptor's pairs with the
// label merge of the batch kernel: one host read brings in the up-to-64
// pairs' ids (lane j holds pair j), the lanes load their own pair's label
// offsets together, and the pairs are then answered in turn from those
// registers; the answers go back as one contiguous store into the mapped
// answer ring.  The wave that completes a request's last descriptor writes
// the request's sequence number into its slot's completion flag.  The poller
// stops taking in new descriptors once asked to (host stop flag), after
// max_ticks in any case, or after idle_ticks without work; when every
// descriptor taken in is done it raises the exit flag, and every wave leaves.
// The host relaunches the kernel when work arrives after an exit.
namespace {

// Relaxed device-coherent load.  No acquire anywhere in the serving loop: an
// acquire at device scope invalidates the L2, and thousands of polling waves
// doing that would strip the cache from the waves that answer queries.
// Nothing a wave reads after a poll can be stale without it: the descriptors,
// pairs and answers are in uncached host memory, read and written with
// system-scope accesses, and the index does not change while it is served.
__device__ inline unsigned long long ld_agent(unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ inline T ld_sys(const T *p)
{
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <class T>
__device__ inline void st_sys(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Waits until this wave's earlier memory operations (stores included) are done.
__device__ inline void wait_memory() { __builtin_amdgcn_s_waitcnt(0); }

// The value of the wave's first active lane, as a wave-uniform (scalar)
// value.  Every control decision of the serving loop goes through it: the
// compiler must see the ticket, the published count, the exit flag and the
// descriptor as uniform, or it structures the loops for lanes that could
// part ways -- and a lane set that re-enters the wait without lane 0 would
// broadcast a stale ticket.
// (uniform64 / uniform32: common.hpp)

__device__ void serve_poll(const ServeKernelArgs &A)
{
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long *S = A.state;
    uint64_t *pub = const_cast<uint64_t *>(A.ctl->pub);
    uint32_t *stop_p = const_cast<uint32_t *>(&A.ctl->stop);
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    uint64_t last = t0;
    bool frozen = false;
    for (;;) {
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        const uint32_t stop = uniform32(ld_sys(stop_p));
        bool moved = false, busy = false;
        if (lane < kServeQueues) {
            unsigned long long seen = ld_agent(&S[serve_seen(lane)]);
            if (!frozen) {
                const uint64_t p = ld_sys(pub + lane);
                if (p > seen) {
                    __hip_atomic_store(&S[serve_seen(lane)], (unsigned long long)p, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    seen = p;
                    moved = true;
                }
            }
            busy = ld_agent(&S[serve_done(lane)]) != seen;
        }
        const bool any_busy = __ballot(busy) != 0ull;
        if (__ballot(moved) != 0ull || any_busy) last = now;
        if (stop || now - t0 > A.max_ticks || (!any_busy && now - last > A.idle_ticks)) frozen = true;
        if (frozen && !any_busy) {
            __hip_atomic_store(&S[kServeExit], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(16);
    }
}

// K: 0 = u8 hop answers, 1 = wide index (u16 exact hops), 2 = weighted (u32).
template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_serve(uint32_t n, const uint8_t *bpd, const ulonglong2 *bps,
                                               const uint64_t *loff, const uint32_t *lab, const uint32_t *wdist,
                                               ServeKernelArgs A)
{
    // A smaller label buffer than the batch kernel's leaves room for the
    // descriptor's pairs within 20 KB of LDS per workgroup (eight per CU).
    __shared__ uint32_t s_q[kQWaves][kServeLds];
    // The descriptor's pairs: ids and label offsets, lane j's pair at [j].
    __shared__ uint32_t s_ab[kQWaves][2][kServeGroup], s_len[kQWaves][2][kServeGroup];
    __shared__ uint64_t s_off[kQWaves][2][kServeGroup];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (blockIdx.x == 0 && wv == 0) {
        serve_poll(A);
        return;
    }
    uint32_t *sbuf = s_q[wv];
    unsigned long long *S = A.state;
    const uint32_t q = blockIdx.x % kServeQueues;
    for (;;) {
        // Every lane takes part in the loop's atomics with lane 0 alone
        // contributing (no lane-0-only branch for the loop structure to split
        // on); lane 0 is the first active lane, so its result is the ticket.
        const uint64_t c = uniform64(atomicAdd(&S[serve_claim(q)], lane == 0 ? 1ull : 0ull));
        bool leave = false;
        for (;;) {
            if (uniform64(ld_agent(&S[serve_seen(q)])) > c) break;
            if (uniform64(ld_agent(&S[kServeExit])) != 0) {
                leave = true;
                break;
            }
            // About a microsecond between polls: waiting waves stay off the
            // memory system the working ones use.
            __builtin_amdgcn_s_sleep(40);
        }
        if (leave) return;
        // The descriptor and the pairs live in mapped host memory that is
        // not cached on the device.
        ServeDesc *dp = const_cast<ServeDesc *>(A.desc) + (uint64_t)q * A.dcap + c % A.dcap;
        const uint64_t pos = uniform64(ld_sys(&dp->pos));
        const uint32_t len = uniform32(ld_sys(&dp->len));
        const uint32_t slot = uniform32(ld_sys(&dp->slot));
        const uint64_t seq = uniform64(ld_sys(&dp->seq));
        const uint32_t rk = uniform32(ld_sys(&dp->k));
        const uint64_t phys = pos % A.cap;
        {
            uint32_t ia = 0, ib = 0;
            if (lane < len) {
                ia = ld_sys(A.ra + phys + lane);
                ib = ld_sys(A.rb + phys + lane);
            }
            uint64_t oa = 0, ob = 0;
            uint32_t la = 0, lb = 0;
            if (lane < len && ia < n && ib < n) {
                oa = loff[ia];
                ob = loff[ib];
                la = (uint32_t)(loff[ia + 1] - oa);
                lb = (uint32_t)(loff[ib + 1] - ob);
            }
            s_ab[wv][0][lane] = ia;
            s_ab[wv][1][lane] = ib;
            s_off[wv][0][lane] = oa;
            s_off[wv][1][lane] = ob;
            s_len[wv][0][lane] = la;
            s_len[wv][1][lane] = lb;
            q_lds_sync();
        }
        uint32_t mine = 0;
        for (uint32_t j = 0; j < len; ++j) {
            const uint32_t a = s_ab[wv][0][j], b = s_ab[wv][1][j];
            const uint64_t xa = s_off[wv][0][j], xb = s_off[wv][1][j];
            const uint32_t ya = s_len[wv][0][j], yb = s_len[wv][1][j];
            uint32_t d;
            if (K == 2)
                d = wpair_distance_at<kServeLds>(n, lab, wdist, a, b, xa, ya, xb, yb, A.ref_compat != 0, sbuf);
            else
                d = pair_distance_at<K == 1, kServeLds>(n, bpd, bps, lab, a, b, xa, ya, xb, yb, A.ref_compat != 0,
                                                        sbuf);
            if (lane == j) mine = d;
        }
        // The next descriptor's writes must not overtake these reads.
        q_lds_sync();
        if (lane < len) {
            if (K == 0) st_sys(static_cast<uint8_t *>(A.out) + phys + lane, (uint8_t)mine);
            if (K == 1) st_sys(static_cast<uint16_t *>(A.out) + phys + lane, (uint16_t)mine);
            if (K == 2) st_sys(static_cast<uint32_t *>(A.out) + phys + lane, mine);
        }
        // The answers have reached host memory before this wave's count
        // moves (the stores write through to the host; waiting for them is
        // all the ordering they need).
        wait_memory();
        const uint32_t prev = uniform32(__hip_atomic_fetch_add(&A.rdone[slot], lane == 0 ? len : 0u,
                                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (prev + len == rk) {
            // Last descriptor of the request: reset the slot's count (the
            // host reuses the slot only after it sees the flag), then raise
            // the flag.  Every lane stores the same value.
            __hip_atomic_store(&A.rdone[slot], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wait_memory();
            st_sys(A.flags + slot, seq);
            wait_memory();
        }
        __hip_atomic_fetch_add(&S[serve_done(q)], lane == 0 ? 1ull : 0ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Start of a launch: the previous launch exited only once every descriptor
// it took in was done, so tickets restart at the done count.
__global__ void k_serve_reset(unsigned long long *S)
{
    const uint32_t q = threadIdx.x;
    if (q < kServeQueues) S[serve_claim(q)] = S[serve_done(q)];
    if (q == 0) S[kServeExit] = 0ull;
}

}  // namespace

uint32_t serve_resident_blocks_per_cu(const pado_gpu_index *idx)
{
    int nb = 0;
    hipError_t e;
    if (idx->weighted)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_serve<2>, 256, 0);
    else if (idx->wide)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_serve<1>, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_serve<0>, 256, 0);
    return (e == hipSuccess && nb > 0) ? (uint32_t)nb : 4u;
}

pado_gpu_status launch_serve_resident(const pado_gpu_index *idx, const ServeKernelArgs &a, uint32_t blocks,
                                      hipStream_t st)
{
    hipLaunchKernelGGL(k_serve_reset, dim3(1), dim3(64), 0, st, a.state);
    if (idx->weighted)
        hipLaunchKernelGGL(k_serve<2>, dim3(blocks), dim3(256), 0, st, idx->n, nullptr, nullptr, idx->qoff.p,
                           idx->qlab.p, idx->qwdist.p, a);
    else if (idx->wide)
        hipLaunchKernelGGL(k_serve<1>, dim3(blocks), dim3(256), 0, st, idx->n, idx->qbpd.p,
                           (const ulonglong2 *)idx->qbps.p, idx->qoff.p, idx->qlab.p, nullptr, a);
    else
        hipLaunchKernelGGL(k_serve<0>, dim3(blocks), dim3(256), 0, st, idx->n, idx->qbpd.p,
                           (const ulonglong2 *)idx->qbps.p, idx->qoff.p, idx->qlab.p, nullptr, a);
    PADO_HIP(hipGetLastError());
    return PADO_GPU_OK;
}

// One pair, ids passed by value, the answer written into mapped host memory
// with a system-scope store: a lone query costs one launch and one stream
// synchronisation, no copies (K: 0 u8 hops, 1 wide, 2 weighted).
template <int K>
__global__ __launch_bounds__(64) void k_query_pair(uint32_t n, const uint8_t *bpd, const ulonglong2 *bps,
                                                   const uint64_t *loff, const uint32_t *lab, const uint32_t *wdist,
                                                   uint32_t a, uint32_t b, int ref_compat, uint32_t *out)
{
    __shared__ uint32_t sbuf[kQLds];
    uint32_t d;
    if (K == 2)
        d = wpair_distance(n, loff, lab, wdist, a, b, ref_compat != 0, sbuf);
    else
        d = pair_distance<K == 1>(n, bpd, bps, loff, lab, a, b, ref_compat != 0, sbuf);
    if (threadIdx.x == 0) __hip_atomic_store(out, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

pado_gpu_status launch_query_pair(const pado_gpu_index *idx, bool rank_space, uint32_t a, uint32_t b,
                                  bool ref_compat, uint32_t *out_dev, hipStream_t st)
{
    const uint64_t *loff = rank_space ? idx->loff.p : idx->qoff.p;
    if (idx->weighted)
        hipLaunchKernelGGL(k_query_pair<2>, dim3(1), dim3(64), 0, st, idx->n, nullptr, nullptr, loff,
                           rank_space ? idx->whub.p : idx->qlab.p, rank_space ? idx->wdist.p : idx->qwdist.p, a, b,
                           ref_compat ? 1 : 0, out_dev);
    else if (idx->wide)
        hipLaunchKernelGGL(k_query_pair<1>, dim3(1), dim3(64), 0, st, idx->n, rank_space ? idx->bpd.p : idx->qbpd.p,
                           (const ulonglong2 *)(rank_space ? idx->bps.p : idx->qbps.p), loff,
                           rank_space ? idx->lab.p : idx->qlab.p, nullptr, a, b, 0, out_dev);
    else
        hipLaunchKernelGGL(k_query_pair<0>, dim3(1), dim3(64), 0, st, idx->n, rank_space ? idx->bpd.p : idx->qbpd.p,
                           (const ulonglong2 *)(rank_space ? idx->bps.p : idx->qbps.p), loff,
                           rank_space ? idx->lab.p : idx->qlab.p, nullptr, a, b, ref_compat ? 1 : 0, out_dev);
    PADO_HIP(hipGetLastError());
    return PADO_GPU_OK;
}

pado_gpu_status launch_wquery(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                              uint64_t k, uint32_t *out, bool ref_compat, hipStream_t st);
pado_gpu_status launch_query(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                             uint64_t k, uint8_t *out, bool ref_compat, hipStream_t st);
pado_gpu_status launch_query_wide(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                                  uint64_t k, uint16_t *out, bool ref_compat, hipStream_t st);

// Query launch on device pairs for either index type: out holds u8
// (unweighted) or u32 (weighted) answers.  work == nullptr selects the plain
// grid-stride kernel; otherwise the persistent kernel with `blocks`
// workgroups, its counter cleared on the same stream first.
pado_gpu_status launch_query_any(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                                 uint64_t k, void *out, bool ref_compat, hipStream_t st,
                                 unsigned long long *work, uint32_t blocks)
{
    if (k == 0) return PADO_GPU_OK;
    const uint64_t *loff = rank_space ? idx->loff.p : idx->qoff.p;
    // A wide index always takes the grid-stride kernel (u16 answers).
    if (idx->wide) return launch_query_wide(idx, rank_space, qa, qb, k, (uint16_t *)out, false, st);
    if (!work) {
        if (idx->weighted) return launch_wquery(idx, rank_space, qa, qb, k, (uint32_t *)out, ref_compat, st);
        return launch_query(idx, rank_space, qa, qb, k, (uint8_t *)out, ref_compat, st);
    }
    const uint64_t need = (k + kGrab - 1) / kGrab;  // grabs; one wave per grab at most
    uint32_t grid = blocks;
    if ((uint64_t)grid * 4 > need) grid = (uint32_t)((need + 3) / 4);
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_zero_work, dim3(1), dim3(1), 0, st, work);
    if (idx->weighted) {
        hipLaunchKernelGGL(k_wquery_ws, dim3(grid), dim3(256), 0, st, idx->n, loff,
                           rank_space ? idx->whub.p : idx->qlab.p, rank_space ? idx->wdist.p : idx->qwdist.p, qa, qb,
                           k, (uint32_t *)out, ref_compat ? 1 : 0, work);
    } else {
        hipLaunchKernelGGL(k_query_ws, dim3(grid), dim3(256), 0, st, idx->n, rank_space ? idx->bpd.p : idx->qbpd.p,
                           (const ulonglong2 *)(rank_space ? idx->bps.p : idx->qbps.p), loff,
                           rank_space ? idx->lab.p : idx->qlab.p, qa, qb, k, (uint8_t *)out, ref_compat ? 1 : 0,
                           work);
    }
    PADO_HIP(hipGetLastError());
    return PADO_GPU_OK;
}

pado_gpu_status launch_wquery(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                              uint64_t k, uint32_t *out, bool ref_compat, hipStream_t st)
{
    if (k == 0) return PADO_GPU_OK;
    const uint64_t *loff = rank_space ? idx->loff.p : idx->qoff.p;
    const uint32_t *hub = rank_space ? idx->whub.p : idx->qlab.p;
    const uint32_t *dist = rank_space ? idx->wdist.p : idx->qwdist.p;
    uint32_t grid = grid_for(k * 64, 256, 65536);
    hipLaunchKernelGGL(k_wquery, dim3(grid), dim3(256), 0, st, idx->n, loff, hub, dist, qa, qb, k, out,
                       ref_compat ? 1 : 0);
    PADO_HIP(hipGetLastError());
    return PADO_GPU_OK;
}

pado_gpu_status launch_query(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                             uint64_t k, uint8_t *out, bool ref_compat, hipStream_t st)
{
    if (k == 0) return PADO_GPU_OK;
    const uint8_t *bpd = rank_space ? idx->bpd.p : idx->qbpd.p;
    const ulonglong2 *bps = (const ulonglong2 *)(rank_space ? idx->bps.p : idx->qbps.p);
    const uint64_t *loff = rank_space ? idx->loff.p : idx->qoff.p;
    const uint32_t *lab = rank_space ? idx->lab.p : idx->qlab.p;
    uint32_t grid = grid_for(k * 64, 256, 65536);
    hipLaunchKernelGGL((k_query<false, uint8_t>), dim3(grid), dim3(256), 0, st, idx->n, bpd, bps, loff, lab, qa, qb,
                       k, out, ref_compat ? 1 : 0);
    PADO_HIP(hipGetLastError());
    return PADO_GPU_OK;
}

// u16 answers on device pairs: exact hops for a wide index, the reference's
// answers widened for one in the u8 domain.
pado_gpu_status launch_query_wide(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                                  uint64_t k, uint16_t *out, bool ref_compat, hipStream_t st)
{
    if (k == 0) return PADO_GPU_OK;
    const uint8_t *bpd = rank_space ? idx->bpd.p : idx->qbpd.p;
    const ulonglong2 *bps = (const ulonglong2 *)(rank_space ? idx->bps.p : idx->qbps.p);
    const uint64_t *loff = rank_space ? idx->loff.p : idx->qoff.p;
    const uint32_t *lab = rank_space ? idx->lab.p : idx->qlab.p;
    uint32_t grid = grid_for(k * 64, 256, 65536);
    if (idx->wide)
        hipLaunchKernelGGL((k_query<true, uint16_t>), dim3(grid), dim3(256), 0, st, idx->n, bpd, bps, loff, lab, qa,
                           qb, k, out, 0);
    else
        hipLaunchKernelGGL((k_query<false, uint16_t>), dim3(grid), dim3(256), 0, st, idx->n, bpd, bps, loff, lab,
                           qa, qb, k, out, ref_compat ? 1 : 0);
    PADO_HIP(hipGetLastError());
    return PADO_GPU_OK;
}

}  // namespace pado
