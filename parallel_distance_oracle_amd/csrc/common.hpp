// common.hpp -- shared definitions of the PADO GPU library (gfx950 / MI355X).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "pado_gpu.h"

namespace pado {

constexpr uint32_t kBP = 50;          // BITPARALLEL_SIZE (PARA:45)
constexpr uint32_t kBpdStride = 64;   // bytes per vertex in the bp_dist array (50 used)
constexpr uint32_t kBatch = 1024;     // BATCH_SIZE (PARA:42)
constexpr uint32_t kWords = kBatch / 64;  // 64-bit words per 1024-root bitmap row
constexpr uint8_t kInf8 = 255;        // SMALLI_MAX (GLB:40)
constexpr uint32_t kHubBits = 24;     // packed label = (hub << 8) | dist
constexpr uint32_t kMaxVertices = 1u << kHubBits;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kChunk = 16;       // edges per push work descriptor

// Distance domains of the labeling and query kernels (one templated path):
//  * Hop<false> -- the reference's: u8 hop distances (smalli, GLB:29-40),
//    255 = INF, packed label (hub << 8) | dist, hubs below 2^24, iterations
//    below 255 (the reference's u8 counters wrap there);
//  * Hop<true>  -- wide: 12-bit hop distances in (hub << 12) | dist, 4095 =
//    INF, hubs below 2^20, for graphs whose BFS depth reaches 255 (large
//    grids, road-like graphs).  Bit-parallel distances stay u8 with the BFS
//    stopped at level 254 (farther vertices keep 255: no information in that
//    slot), and the bit-parallel checks skip slots with 255 on either side.
template <bool W>
struct Hop {
    static constexpr uint32_t kBits = W ? 12u : 8u;        // distance bits of a packed label
    static constexpr uint32_t kInf = (1u << kBits) - 1u;   // INF distance; also the iteration limit
    static constexpr uint32_t kMaxN = 1u << (32u - kBits); // vertex ids must stay below this
    using Dist = typename std::conditional<W, uint16_t, uint8_t>::type;  // root-distance table entry
    using Out = typename std::conditional<W, uint16_t, uint8_t>::type;   // query answer
};
constexpr uint32_t kWideMaxVertices = Hop<true>::kMaxN;
// Default grids of the labeling kernels (256-thread blocks): one resident
// round of waves at their occupancy, so a launch with little work ends fast.
constexpr uint32_t kPushGrid = 1024;
constexpr uint32_t kAddGrid = 1024;
// Graphs up to 2^17 vertices rarely have more than a few thousand active
// vertices or candidates per iteration: 256 blocks serve them faster (a 40K
// grid: 0.18 s vs 0.22 s with 1024), and concurrent builds share the chip.
inline uint32_t default_grid(uint32_t n) { return n <= (1u << 17) ? 256u : 1024u; }
// The batch labeling's push and add grids and push descriptor size by
// default.  Measured on R-MAT 20 ef16 (median of 3 builds): push 768 blocks /
// add 512 blocks / 8-edge descriptors 0.326 s; 1024 / 1024 / 16: 0.356 s;
// 4-edge descriptors tie with 8; other grids within 1%.
inline uint32_t default_push_grid(uint32_t n) { return n <= (1u << 17) ? 256u : 768u; }
inline uint32_t default_add_grid(uint32_t n) { return n <= (1u << 17) ? 256u : 512u; }
constexpr uint32_t kPushChunk = 8;

// Thread-local error text behind pado_gpu_last_error().
void set_error(const std::string &msg);
void clear_error();

// Status-returning HIP check for the C ABI implementation.
#define PADO_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::pado::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
            return PADO_GPU_ERR_HIP;                                                \
        }                                                                           \
    } while (0)

#define PADO_TRY(call)                                  \
    do {                                                \
        pado_gpu_status s_ = (call);                    \
        if (s_ != PADO_GPU_OK) return s_;               \
    } while (0)

// Serialises device allocation and release with hipGraph capture: builds of a
// many-graph job run on several threads, and a release may synchronise the
// device while another thread's stream is capturing.
inline std::recursive_mutex &capture_mutex()
{
    static std::recursive_mutex m;
    return m;
}

// Device buffer owned by a handle.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    pado_gpu_status alloc(size_t count)
    {
        release();
        if (count == 0) count = 1;
        std::lock_guard<std::recursive_mutex> g(capture_mutex());
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(" + std::to_string(count * sizeof(T)) + " B): " + hipGetErrorString(e));
            return PADO_GPU_ERR_NOMEM;
        }
        n = count;
        return PADO_GPU_OK;
    }
    void release()
    {
        if (p) {
            std::lock_guard<std::recursive_mutex> g(capture_mutex());
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
};

// The first active lane's value, in scalar registers: for values that are the
// same in every lane (loaded from one address, or broadcast), so that the
// compiler treats them as wave-uniform.
__device__ inline uint64_t uniform64(uint64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint32_t uniform32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Append one item per lane with `pred` to a list counted by *counter: the
// active lanes of the wave take consecutive slots with a single atomic, so a
// shared counter sees one update per wave instead of one per item.  Returns
// the lane's slot (meaningless where !pred).
__device__ inline uint32_t wave_append(uint32_t *counter, bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return 0u;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t leader = (uint32_t)__ffsll(m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// Device-side counters shared by the labeling kernels (one struct per build).
struct BuildCounters {
    // Frontier and candidate counts of batch iteration i live in slot i & 1:
    // each kernel reads its own slot and empties the other one for the next
    // kernel to fill, so no control launch runs between iterations.
    uint32_t n_act[2];     // active vertices of the iteration
    uint32_t n_vert;       // vertices of the graph
    uint32_t abits_words;  // 64-bit words of one frontier bitmap slot
    uint32_t n_chunk[2];   // push work descriptors of the iteration
    uint32_t n_cand[2];    // candidate vertices of the iteration
    uint32_t n_touched;    // vertices whose indicator row was touched this batch
    // Batch gate: a batch enqueued before the host has seen the previous one
    // finish runs as no-ops if the previous batch still has a frontier.
    uint32_t hold;         // set: every batch kernel returns at once
    uint32_t iter_last;    // last iteration k_push ran (not held)
    uint32_t overflow;     // label pool exhausted
    uint32_t dq_rows;      // root-distance rows scanned (statistic)
    unsigned long long pool_top;  // next free label slot in the pool
    unsigned long long n_pedge[2];  // push edges of the iteration's frontier (sum of push degrees)
    unsigned long long pulls;     // iterations run in pull form (statistic)
    unsigned long long pushes;    // (tail, root) pairs offered by push
    unsigned long long ind_pass;  // pairs that set a fresh indicator bit
    unsigned long long bp_pruned; // pairs pruned by the bit-parallel check
    unsigned long long cands;     // (vertex, root) candidates after the BP check
    unsigned long long dq_pruned; // candidates pruned by the distance query
    unsigned long long labels;    // labels appended
    unsigned long long edges;     // push edges scanned (tails >= roots_start)
    unsigned long long tails;     // distinct tails with a fresh indicator bit, summed over iterations
    unsigned long long dq_labels; // labels with dist < iter read by distance queries
    unsigned long long actives;   // active vertices, summed over iterations
    // Device-side schedule: the batch step reads its parameters from here, so
    // one launch sequence (or one captured hipGraph) serves every batch.
    uint32_t next_batch;          // batch index the next k_batch_begin starts
    uint32_t rs, rsz, gen_batch;  // current batch: roots [rs, rs + rsz)
    uint32_t domain;              // an iteration reached 255 (u8 domain)
    unsigned long long iterations;  // live iterations, summed over batches
};

// Host-side timing of the labeling kernels (HIP events on the build stream).
struct KernelTimes {
    double ms_push = 0, ms_add = 0, ms_bp = 0;
    uint64_t n_push = 0, n_add = 0, n_bp = 0;
};

}  // namespace pado

// Opaque handle definitions (C ABI types).
struct pado_gpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;       // second stream of chunked host-array queries
    hipEvent_t side_ev = nullptr;     // orders the side stream after work queued before a query
    int num_cu = 0;
};

struct pado_gpu_graph {
    pado_gpu_ctx *ctx = nullptr;
    uint32_t n = 0;
    uint64_t m = 0;  // directed edge count (2x undirected)
    pado::DevBuf<uint64_t> off;
    pado::DevBuf<uint32_t> adj;
    pado::DevBuf<uint8_t> w;  // edge weights (weighted graphs), else empty
    bool weighted = false;
    std::vector<uint64_t> h_off;  // host copy: root selection is a serial pass
    std::vector<uint32_t> h_adj;
    std::vector<uint8_t> h_w;
};

struct pado_gpu_index {
    pado_gpu_ctx *ctx = nullptr;
    uint32_t n = 0;
    uint32_t weighted = 0;
    uint32_t wide = 0;  // labels in the wide hop domain (Hop<true>)
    uint32_t bp_pass_through = 0;  // wide build whose bit-parallel BFS was cut at level 254:
                                   // its bit-parallel vertices were labeled as ordinary hubs
    // Bit-parallel labels, rank space: bpd[v*64 + i], bps[(v*50 + i)*2 + {0,1}].
    pado::DevBuf<uint8_t> bpd;
    pado::DevBuf<uint64_t> bps;
    // Canonical labels, rank space: CSR of packed (hub << 8) | dist, hubs ascending.
    pado::DevBuf<uint64_t> loff;
    pado::DevBuf<uint32_t> lab;
    uint64_t num_labels = 0;
    // Weighted labels (hub u32, dist u32), rank space.
    pado::DevBuf<uint32_t> whub;
    pado::DevBuf<uint32_t> wdist;
    // Query copy in original-id order (order_labels, PARA:2198-2281).
    int ordered = 0;
    pado::DevBuf<uint8_t> qbpd;
    pado::DevBuf<uint64_t> qbps;
    pado::DevBuf<uint64_t> qoff;
    pado::DevBuf<uint32_t> qlab;
    pado::DevBuf<uint32_t> qwhub;
    pado::DevBuf<uint32_t> qwdist;
    // Staging for host-pointer queries.
    pado::DevBuf<uint32_t> qa, qb;
    pado::DevBuf<uint8_t> qout;
    pado::DevBuf<uint32_t> qout32;
    // One-pair queries: the answer word in mapped host memory, the kernel's
    // result slot (pado_gpu_query_pair); its lock serialises callers.
    uint32_t *pair_out = nullptr;
    uint32_t *pair_out_dev = nullptr;
    std::mutex pair_mu;
    ~pado_gpu_index()
    {
        if (pair_out) {
            std::lock_guard<std::recursive_mutex> g(pado::capture_mutex());
            (void)hipHostFree(pair_out);
        }
    }
    // Build statistics.
    double ms_bp = 0, ms_batch = 0, ms_final = 0;
    uint32_t bp_roots[pado::kBP] = {0};
    uint32_t bp_max_level = 0;
    uint64_t iterations = 0;
    pado::BuildCounters counters{};
    pado::KernelTimes ktimes{};
};
