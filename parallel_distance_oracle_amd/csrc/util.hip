// util.hip -- device scan / segmented sort helpers (hipCUB) used outside the
// hot loops: label-store compaction and final canonical ordering.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace pado {
namespace {

__global__ void k_widen(const uint32_t *in, uint64_t *out, uint32_t n)
{
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) out[i] = (i < n) ? in[i] : 0;
}

// Plain device copy, 16 B per thread per step (grid-stride), tail by bytes.
__global__ __launch_bounds__(256) void k_copy_bytes(uint8_t *dst, const uint8_t *src, uint64_t bytes)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n16 = bytes / 16;
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) d4[i] = s4[i];
    for (uint64_t i = n16 * 16 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes; i += stride)
        dst[i] = src[i];
}

}  // namespace

extern "C" pado_gpu_status pado_gpu_copy_bytes(pado_gpu_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    clear_error();
    if (!ctx || (bytes && (!dst || !src))) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    if (((uintptr_t)dst | (uintptr_t)src) & 15u) {
        set_error("device copy needs 16-byte aligned buffers");
        return PADO_GPU_ERR_ARG;
    }
    if (!bytes) return PADO_GPU_OK;
    PADO_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_copy_bytes, dim3(grid_for(bytes / 16 + 1, 256, 16384)), dim3(256), 0, ctx->stream,
                       (uint8_t *)dst, (const uint8_t *)src, bytes);
    PADO_HIP(hipGetLastError());
    return PADO_GPU_OK;
}

pado_gpu_status scan_u32_to_u64(const uint32_t *in, uint64_t *out, uint32_t n, hipStream_t s)
{
    DevBuf<uint64_t> tmp;
    PADO_TRY(tmp.alloc((uint64_t)n + 1));
    hipLaunchKernelGGL(k_widen, dim3(grid_for((uint64_t)n + 1, 256, 8192)), dim3(256), 0, s, in, tmp.p, n);
    size_t bytes = 0;
    PADO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, tmp.p, out, (int)(n + 1), s));
    DevBuf<uint8_t> ws;
    PADO_TRY(ws.alloc(bytes));
    PADO_HIP(hipcub::DeviceScan::ExclusiveSum(ws.p, bytes, tmp.p, out, (int)(n + 1), s));
    PADO_HIP(hipStreamSynchronize(s));
    return PADO_GPU_OK;
}

// Relative offsets of vertices [v0, v1]: rel[i] = off[v0 + i] - off[v0].
__global__ void k_rel_offsets(const uint64_t *off, uint32_t v0, uint32_t cnt, uint64_t *rel)
{
    uint32_t stride = gridDim.x * blockDim.x;
    const uint64_t b = off[v0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= cnt; i += stride) rel[i] = off[v0 + i] - b;
}

pado_gpu_status segmented_sort_u32(uint32_t *keys, uint64_t total, const uint64_t *off, uint32_t n,
                                   hipStream_t s, uint64_t max_items)
{
    if (total == 0 || n == 0) return PADO_GPU_OK;
    // hipCUB takes the item and segment counts as int: an index with more than
    // INT_MAX labels is sorted in vertex ranges of at most max_items labels
    // (one vertex holds fewer than n < 2^24 labels, so every range is valid).
    if (max_items == 0 || max_items > kSortMaxItems) max_items = kSortMaxItems;
    std::vector<uint32_t> cuts{0u, n};
    if (total > max_items) {
        std::vector<uint64_t> h_off((uint64_t)n + 1);
        PADO_HIP(hipMemcpyAsync(h_off.data(), off, 8ull * (n + 1), hipMemcpyDeviceToHost, s));
        PADO_HIP(hipStreamSynchronize(s));
        cuts.resize((uint64_t)n + 1);
        const uint32_t nr = pado_sort_plan(h_off.data(), n, max_items, cuts.data(), n + 1);
        if (nr == 0) {
            set_error("a vertex holds more labels than one sort call can take");
            return PADO_GPU_ERR_DOMAIN;
        }
        cuts.resize(nr + 1);
    }
    std::vector<uint64_t> h_base(cuts.size());
    if (cuts.size() > 2) {
        for (size_t r = 0; r < cuts.size(); ++r) {
            PADO_HIP(hipMemcpyAsync(&h_base[r], off + cuts[r], 8, hipMemcpyDeviceToHost, s));
        }
        PADO_HIP(hipStreamSynchronize(s));
    } else {
        h_base[0] = 0;
        h_base[1] = total;
    }
    uint64_t biggest = 0;
    for (size_t r = 0; r + 1 < cuts.size(); ++r) biggest = std::max<uint64_t>(biggest, h_base[r + 1] - h_base[r]);
    DevBuf<uint32_t> out;
    PADO_TRY(out.alloc(biggest));
    DevBuf<uint64_t> rel;
    if (cuts.size() > 2) PADO_TRY(rel.alloc((uint64_t)n + 1));
    DevBuf<uint8_t> ws;
    for (size_t r = 0; r + 1 < cuts.size(); ++r) {
        const uint32_t v0 = cuts[r], v1 = cuts[r + 1];
        const uint64_t items = h_base[r + 1] - h_base[r];
        if (items == 0) continue;
        const uint64_t *o = off;
        if (cuts.size() > 2) {
            hipLaunchKernelGGL(k_rel_offsets, dim3(grid_for((uint64_t)(v1 - v0) + 1, 256, 4096)), dim3(256), 0, s,
                               off, v0, v1 - v0, rel.p);
            o = rel.p;
        }
        uint32_t *kr = keys + h_base[r];
        size_t bytes = 0;
        PADO_HIP(hipcub::DeviceSegmentedSort::SortKeys(nullptr, bytes, kr, out.p, (int)items, (int)(v1 - v0), o,
                                                       o + 1, s));
        if (bytes > ws.n) PADO_TRY(ws.alloc(bytes));
        PADO_HIP(hipcub::DeviceSegmentedSort::SortKeys(ws.p, bytes, kr, out.p, (int)items, (int)(v1 - v0), o, o + 1,
                                                       s));
        PADO_HIP(hipMemcpyAsync(kr, out.p, items * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        // The relative offsets and the workspace are reused by the next range.
        PADO_HIP(hipStreamSynchronize(s));
    }
    return PADO_GPU_OK;
}

}  // namespace pado

// Vertex ranges for the canonical label sort: cuts[0] = 0 < cuts[1] < ... <
// cuts[r] = n with at most max_items labels in each range [cuts[i],
// cuts[i+1]).  Greedy: a range grows while it fits.  Returns r, or 0 if one
// vertex alone holds more than max_items labels or max_cuts is too small.
extern "C" uint32_t pado_sort_plan(const uint64_t *off, uint32_t n, uint64_t max_items, uint32_t *cuts,
                                   uint32_t max_cuts)
{
    if (!off || !cuts || max_cuts < 2 || max_items == 0) return 0;
    uint32_t r = 0;
    cuts[0] = 0;
    uint32_t v0 = 0;
    while (v0 < n) {
        uint32_t v = v0;
        while (v < n && off[v + 1] - off[v0] <= max_items) ++v;
        if (v == v0) return 0;  // vertex v0 alone exceeds the cap
        if (r + 2 > max_cuts) return 0;
        cuts[++r] = v;
        v0 = v;
    }
    if (n == 0) cuts[++r] = 0;
    return r;
}
