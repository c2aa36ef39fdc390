// internal.hpp -- declarations shared between the translation units of the
// PADO GPU library (not part of the C ABI).
#pragma once

#include "common.hpp"

namespace pado {

// Bit-parallel labeling (PARA:544-840).  Fills idx->bpd / idx->bps and the
// used-root flags (device + host copies).
pado_gpu_status build_bp(pado_gpu_graph *g, pado_gpu_index *idx, DevBuf<uint8_t> &d_used,
                         std::vector<uint8_t> &h_used);
// Same result, all 50 slots in one multi-source sweep (bp_ms.hip).
pado_gpu_status build_bp_multi(pado_gpu_graph *g, pado_gpu_index *idx, DevBuf<uint8_t> &d_used,
                               std::vector<uint8_t> &h_used);

// Seeds of the multi-source sweep: slot roots (j = 255) and their selected
// neighbours (j = neighbour index), as vertex ids of the graph being labeled.
struct BpSeeds {
    std::vector<uint32_t> v, slot, j;
};
// Root / neighbour selection of PARA:670-709 on one rank-space CSR; `map`
// (nullable) renames the chosen vertices (e.g. into a union graph's ids).
void select_bp_roots(uint32_t n, const uint64_t *off, const uint32_t *adj, std::vector<uint8_t> &used,
                     uint32_t *roots, BpSeeds &seeds, const uint32_t *map);
// Multi-source sweep from given seeds; h_used = the used flags (n entries).
pado_gpu_status build_bp_seeded(pado_gpu_graph *g, pado_gpu_index *idx, DevBuf<uint8_t> &d_used,
                                const std::vector<uint8_t> &h_used, const BpSeeds &seeds);

// An unweighted index up to its batch phase: the bit-parallel labels
// (idx->wide chooses the sweep) and the used flags -- all zero when the
// wide domain's slot BFS was cut at level 254 (pass-through).
pado_gpu_status build_bp_phase(pado_gpu_graph *g, pado_gpu_index *idx, const pado_gpu_build_cfg &c,
                               DevBuf<uint8_t> &used);

// Batched pruned labeling (PARA:848-1836).  Produces the canonical rank-space
// label CSR idx->loff / idx->lab.
pado_gpu_status build_batches(pado_gpu_graph *g, pado_gpu_index *idx, const DevBuf<uint8_t> &d_used,
                              const pado_gpu_build_cfg &cfg);

// The same batch phase run by one workgroup per graph (k_batch_loop), for
// several graphs in one launch on st: jobs with status PADO_GPU_OK on entry
// take part; on return each has its own status (PADO_GPU_ERR_NOMEM with
// pool_exhausted when its fixed pool ran dry -- rerun it with build_batches).
struct LoopJob {
    pado_gpu_graph *g = nullptr;
    pado_gpu_index *idx = nullptr;
    const uint8_t *used = nullptr;   // device bit-parallel flags of the graph
    uint32_t batch = kBatch;
    unsigned long long pool_min = 0;
    pado_gpu_status status = PADO_GPU_OK;
    unsigned long long pool_exhausted = 0;
};
pado_gpu_status build_batches_loop(std::vector<LoopJob> &jobs, hipStream_t st);

// Weighted labeling (WPARA).
pado_gpu_status build_weighted(pado_gpu_graph *g, pado_gpu_index *idx, uint32_t batch_size);

// Exclusive scan of n u32 values into n+1 u64 offsets (out[n] = total).
pado_gpu_status scan_u32_to_u64(const uint32_t *in, uint64_t *out, uint32_t n, hipStream_t s);
// Sort every CSR segment of u32 keys ascending in place, in vertex ranges of
// at most max_items keys per hipCUB call (0 = kSortMaxItems; hipCUB counts
// items in int).
constexpr uint64_t kSortMaxItems = 0x7FFFFFFFull - 1024;
pado_gpu_status segmented_sort_u32(uint32_t *keys, uint64_t total, const uint64_t *off, uint32_t n,
                                   hipStream_t s, uint64_t max_items = 0);

// A blocking copy on `s` (never the legacy default stream, which would have
// to wait for -- or fail against -- other threads' streams and captures).
inline hipError_t copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

// Resident serving (serve.hip drives it, query.hip holds the kernel).
// Work reaches the resident kernel as descriptors of at most kServeGroup
// pairs each, spread over kServeQueues queues; the host publishes a queue's
// descriptor count in mapped memory and the kernel reports each request's
// completion by writing its sequence number into the request slot's flag.
constexpr uint32_t kServeQueues = 8;
constexpr uint32_t kServeGroup = 64;
struct alignas(32) ServeDesc {
    uint64_t pos;   // ring position of the first pair (monotonic)
    uint32_t len;   // pairs, 1..kServeGroup
    uint32_t slot;  // request slot
    uint64_t seq;   // the request's sequence number (its id)
    uint32_t k;     // the request's pairs
    uint32_t pad;
};
// Host-written control words, mapped for the device.
struct ServeCtl {
    uint64_t pub[kServeQueues];  // descriptors published per queue
    uint32_t stop;               // the kernel drains and exits
    uint32_t pad[15];
};
// Device-side queue state, one 128-byte line per counter: claim (tickets
// handed out), seen (published descriptors the kernel has taken in), done
// (descriptors completed), then the exit flag.
constexpr uint32_t kServeLine = 16;  // u64 per line
inline constexpr uint32_t serve_claim(uint32_t q) { return q * kServeLine; }
inline constexpr uint32_t serve_seen(uint32_t q) { return (kServeQueues + q) * kServeLine; }
inline constexpr uint32_t serve_done(uint32_t q) { return (2 * kServeQueues + q) * kServeLine; }
constexpr uint32_t kServeExit = 3 * kServeQueues * kServeLine;
constexpr uint32_t kServeStateWords = kServeExit + kServeLine;

struct ServeKernelArgs {
    const uint32_t *ra, *rb;  // the pair ring (mapped host memory)
    void *out;                // the answer ring (mapped host memory)
    uint64_t cap;             // ring capacity in pairs
    const ServeCtl *ctl;
    const ServeDesc *desc;    // kServeQueues rings of dcap descriptors
    uint32_t dcap;
    uint64_t *flags;          // per request slot (mapped host memory)
    unsigned long long *state;  // kServeStateWords (device memory)
    uint32_t *rdone;          // pairs answered per request slot (device memory)
    uint64_t idle_ticks;      // exit after this long without work (100 MHz ticks)
    uint64_t max_ticks;       // drain and exit after this long in any case
    int ref_compat;
};
// Enqueues the state reset and the resident kernel on `st`.
pado_gpu_status launch_serve_resident(const pado_gpu_index *idx, const ServeKernelArgs &a, uint32_t blocks,
                                      hipStream_t st);
uint32_t serve_resident_blocks_per_cu(const pado_gpu_index *idx);

// Launch geometry helpers.
inline uint32_t grid_for(uint64_t threads, uint32_t block, uint32_t cap = 1u << 20)
{
    uint64_t g = (threads + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

}  // namespace pado
