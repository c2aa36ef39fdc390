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

// Batched pruned labeling (PARA:848-1836).  Produces the canonical rank-space
// label CSR idx->loff / idx->lab.
pado_gpu_status build_batches(pado_gpu_graph *g, pado_gpu_index *idx, const DevBuf<uint8_t> &d_used,
                              const pado_gpu_build_cfg &cfg);

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

// Launch geometry helpers.
inline uint32_t grid_for(uint64_t threads, uint32_t block, uint32_t cap = 1u << 20)
{
    uint64_t g = (threads + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

}  // namespace pado
