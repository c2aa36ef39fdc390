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

}  // namespace

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
