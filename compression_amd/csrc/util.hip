// Utility kernels: a streaming device copy.  It is the calibration kernel
// for HBM counters (a known number of bytes read and written) and plain
// device-to-device plumbing for callers that do not link HIP.
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

__global__ __launch_bounds__(256) void copy16_kernel(const u32x4* __restrict__ src,
                                                      u32x4* __restrict__ dst, uint64_t n16) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ __launch_bounds__(256) void copy1_kernel(const uint8_t* __restrict__ src,
                                                     uint8_t* __restrict__ dst, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

int dev_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
    if (!bytes) return MDZ_OK;
    if (!dst || !src) return MDZ_ERR_INVALID;
    if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0) {
        uint64_t n16 = bytes / 16;
        if (n16) {
            uint64_t blocks = (n16 + 255) / 256;
            if (blocks > 256 * 8 * 4) blocks = 256 * 8 * 4;     // ~32 waves per CU, grid-stride
            hipLaunchKernelGGL(copy16_kernel, dim3((uint32_t)blocks), dim3(256), 0, s,
                               (const u32x4*)src, (u32x4*)dst, n16);
        }
        uint64_t done = n16 * 16;
        if (done < bytes)
            hipLaunchKernelGGL(copy1_kernel, dim3(1), dim3(256), 0, s, (const uint8_t*)src + done,
                               (uint8_t*)dst + done, bytes - done);
    } else {
        hipLaunchKernelGGL(copy1_kernel, dim3((uint32_t)((bytes + 255) / 256)), dim3(256), 0, s,
                           (const uint8_t*)src, (uint8_t*)dst, bytes);
    }
    return record_hip(hipGetLastError(), "copy kernel launch");
}

}  // namespace mdz
