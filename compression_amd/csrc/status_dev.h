// Per-chunk status combined across the workgroups that share a chunk.
//
// A chunk decoded by several workgroups (LZSS lane groups, RLE v2 lane
// parts) gets its status as an atomic minimum of (lane << 8 | status byte)
// over the workgroups that found an error: the lowest failing lane decides,
// as in a decoder that checks lanes in order.  status_fill_kernel sets the
// "no error" value before the decode, status_finish_kernel turns the minimum
// into the status code after it.
#pragma once
#include <climits>

#include "mdz_device.h"

namespace mdz {
namespace {

__global__ void status_fill_kernel(int32_t* st, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        st[i] = INT_MAX;
}

__global__ void status_finish_kernel(int32_t* st, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t v = st[i];
        st[i] = v == INT_MAX ? MDZ_ST_OK : (int32_t)(int8_t)(v & 0xff);
    }
}

inline dim3 status_grid(uint64_t n) {
    const uint64_t b = (n + 255) / 256;
    return dim3((uint32_t)(b < 4096 ? (b ? b : 1) : 4096));
}

}  // namespace
}  // namespace mdz
