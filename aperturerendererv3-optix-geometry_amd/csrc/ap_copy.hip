// Streaming device-to-device copy (include/aperture/ap_device.h).
//
// 16 B per lane per iteration, grid-stride, 1024 workgroups of 256 lanes:
// every byte of the source is read once and every byte of the destination
// written once, so its HBM counter readings are the known-size reference
// against which other kernels' FETCH_SIZE / WRITE_SIZE are scaled.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "aperture/ap_device.h"
#include "ap_internal.h"

namespace ap {
namespace {

__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                          size_t n16) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ void tail_copy_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace
}  // namespace ap

extern "C" int ap_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
    using namespace ap;
    if (bytes == 0) return AP_OK;
    if (!dst || !src) return fail(AP_ERR_INVALID_ARG, "ap_device_copy: NULL pointer");
    if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u)
        return fail(AP_ERR_INVALID_ARG, "ap_device_copy: pointers must be 16-byte aligned");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t n16 = bytes / 16, tail = bytes - n16 * 16;
    if (n16) {
        const size_t blocks = std::min<size_t>(1024, (n16 + 255) / 256);
        hipLaunchKernelGGL(stream_copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                           static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
    }
    if (tail) {
        hipLaunchKernelGGL(tail_copy_kernel, dim3(1), dim3(64), 0, s, static_cast<const uint8_t*>(src) + n16 * 16,
                           static_cast<uint8_t*>(dst) + n16 * 16, tail);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_device_copy");
}
