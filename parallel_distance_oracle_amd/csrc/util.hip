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

pado_gpu_status segmented_sort_u32(uint32_t *keys, uint64_t total, const uint64_t *off, uint32_t n,
                                   hipStream_t s)
{
    if (total == 0 || n == 0) return PADO_GPU_OK;
    DevBuf<uint32_t> out;
    PADO_TRY(out.alloc(total));
    size_t bytes = 0;
    PADO_HIP(hipcub::DeviceSegmentedSort::SortKeys(nullptr, bytes, keys, out.p, (int)total, (int)n, off, off + 1,
                                                   s));
    DevBuf<uint8_t> ws;
    PADO_TRY(ws.alloc(bytes));
    PADO_HIP(hipcub::DeviceSegmentedSort::SortKeys(ws.p, bytes, keys, out.p, (int)total, (int)n, off, off + 1, s));
    PADO_HIP(hipMemcpyAsync(keys, out.p, total * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    PADO_HIP(hipStreamSynchronize(s));
    return PADO_GPU_OK;
}

}  // namespace pado
