// Scan and interleave kernels shared by the lane-transposed encoders (RLE v1,
// RLE v2, LZSS): chunk lengths -> offsets, and the 4-byte round-robin
// transpose of 32 lane streams (LZSS: in its sorted order) into a chunk (lane t's word r lands at word
// sum_t' min(words_t', r) + #{t' < t : words_t' > r}), tail zeroed to the
// chunk's 128-byte length.
#pragma once

#include "mdz_device.h"

namespace mdz {
namespace {

constexpr uint32_t kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void lanes_scan_kernel(const uint64_t* __restrict__ sizes, uint64_t n,
                                                                  uint64_t* __restrict__ off, uint64_t* __restrict__ total,
                                                                  uint64_t last) {
    __shared__ uint64_t part[kScanThreads / kWave];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (n + kScanThreads - 1) / kScanThreads;
    const uint64_t b = min(n, per * t), e = min(n, b + per);
    uint64_t s = 0;
    for (uint64_t i = b; i < e; ++i) s += sizes[i];
    const uint64_t incl = wave_incl_scan_u64(s);
    if (lane_id() == kWave - 1) part[t / kWave] = incl;
    __syncthreads();
    uint64_t run = 0;
    for (uint32_t w = 0; w < t / kWave; ++w) run += part[w];
    run += incl - s;
    if (t == 0) off[0] = 0;
    for (uint64_t i = b; i < e; ++i) {
        run += sizes[i];
        off[i + 1] = i + 1 == n && last != ~0ull ? last : run;
        if (i + 1 == n && total) *total = run;
    }
}

constexpr uint32_t kIlvThreads = 256;

constexpr uint32_t kRoundTile = 1024;

// Word r of slot t lands at base(r) + #{t' < t : words_t' > r}, with
// base(r) = sum_t' min(words_t', r).  Per tile of rounds, base and the mask of
// slots still holding words are computed once per round (LDS); each word then
// costs a popcount.
__global__ __launch_bounds__(kIlvThreads) void lanes_interleave_kernel(const uint8_t* __restrict__ work,
                                                                       uint64_t lane_slot,
                                                                       const uint64_t* __restrict__ col_len,
                                                                       const uint64_t* __restrict__ blk_off,
                                                                       const uint64_t* __restrict__ chunk_len,
                                                                       const uint8_t* __restrict__ lane_map,
                                                                       uint8_t* __restrict__ payload) {
    __shared__ uint32_t words[32];
    __shared__ uint32_t rbase[kRoundTile], rmask[kRoundTile];
    const uint32_t tid = threadIdx.x;
    const uint64_t c = blockIdx.x;
    if (tid < 32) words[tid] = (uint32_t)((col_len[c * 32 + tid] + 3) / 4);
    __syncthreads();
    uint32_t* dst = reinterpret_cast<uint32_t*>(payload + blk_off[c]);   // blk_off is a multiple of 128
    uint32_t total = 0, maxw = 0;
    for (uint32_t t = 0; t < 32; ++t) {
        total += words[t];
        maxw = max(maxw, words[t]);
    }
    for (uint32_t r0 = 0; r0 < maxw; r0 += kRoundTile) {
        const uint32_t rn = min(kRoundTile, maxw - r0);
        for (uint32_t i = tid; i < rn; i += kIlvThreads) {
            const uint32_t r = r0 + i;
            uint32_t base = 0, mask = 0;
            for (uint32_t u = 0; u < 32; ++u) {
                const uint32_t wu = words[u];
                base += wu < r ? wu : r;
                mask |= (wu > r ? 1u : 0u) << u;
            }
            rbase[i] = base;
            rmask[i] = mask;
        }
        __syncthreads();
        for (uint32_t j = tid; j < 32 * rn; j += kIlvThreads) {
            const uint32_t t = j / rn, i = j - t * rn, r = r0 + i;
            if (r >= words[t]) continue;
            const uint32_t pos = rbase[i] + __builtin_popcount(rmask[i] & ((1u << t) - 1u));
            // slot t holds lane_map[t]'s stream (identity without a map)
            const uint8_t* src = work + (c * 32 + (lane_map ? lane_map[c * 32 + t] : t)) * lane_slot;
            const uint64_t len = col_len[c * 32 + t];
            uint32_t v = *reinterpret_cast<const uint32_t*>(src + 4ull * r);
            const uint64_t have = len - 4ull * r;       // bytes of this word inside the stream
            if (have < 4) v &= (1u << (8 * have)) - 1u;
            dst[pos] = v;
        }
        __syncthreads();
    }
    const uint64_t clen = chunk_len[c];
    for (uint64_t q = total + tid; q < clen / 4; q += kIlvThreads) dst[q] = 0;
}

}  // namespace
}  // namespace mdz
