// Plain RLE v2 decoder for gfx950: one run stream per chunk of int64 values.
//
// Replaces the reference's one-thread-per-chunk kernel_decompress and
// block_decode (include/rlev2/decoder.h:190-226) and its host entry
// rlev2::decompress_gpu (:227-262).  Container: [u32 n+1][u64 ptr[n+1]]
// [u64 raw_bytes][streams]; chunk c's stream is data[ptr[c], ptr[c+1]) and
// its values are the int64s at byte c * chunk of the output.
//
// One 64-thread workgroup (one wave) per chunk:
//   * the wave stages its stream in 5 KiB windows of LDS with 16-byte loads
//     (a window is restaged at the run that would cross its end; a run is at
//     most 4.4 KB), small enough for seven waves per SIMD: the walk is bound
//     by LDS and memory latency, which more resident chunks hide (8 KiB
//     windows and four waves per SIMD took 1.15 ms on the 1 GiB column,
//     5 KiB and seven 0.84 ms; one chunk per workgroup frees each chunk's
//     LDS as soon as it ends);
//   * the header walk is wave-uniform: every lane reads the same header
//     bytes (an LDS broadcast), which readfirstlane moves to scalar
//     registers, so the run's fields and the control flow are scalar;
//   * each run is checked as the oracle's mdz_ref_rlev2_plain_chunk_decode
//     checks it, in the same order, before any of its values is written; the
//     first failing run ends the chunk with its status;
//   * the run's values are written by all 64 lanes, one value per lane, as
//     8-byte stores of consecutive addresses:
//       SHORT_REPEAT  lanes below the count write the value;
//       DIRECT        lane i unpacks the i-th MSB-first field;
//       DELTA         base + i * delta, or lane i unpacks delta i and a wave
//                     prefix sum adds (or subtracts, for a negative delta
//                     base) them to the second value;
//       PATCHED_BASE  lane j reads patch entry j, a prefix sum of the gaps
//                     places it, and each lane ORs in the patches that land on
//                     its value (full pw-bit mask) before adding the base.
// The reference's 16-bit patch mask and the two values its packed-delta loop
// rewrites past a run's end are not reproduced (the oracle states why).
#include "rlev2_dev.h"

namespace mdz {
namespace {

constexpr uint32_t kPlainWaves = 1;              // chunks (waves) per workgroup
constexpr uint32_t kPlainWin = 5120;             // staged stream bytes per wave (a multiple of 1 KiB)
constexpr uint32_t kPlainSlack = 64;             // room for the field readers' over-reads
constexpr uint32_t kPlainHdr = 40;               // bytes a header may span (two varints)

struct PlainV2Lds {
    uint8_t win[kPlainWaves][kPlainWin + kPlainSlack];
};
static_assert((kPlainWin + kPlainSlack) % 16 == 0, "16-byte aligned windows");
static_assert(kPlainWin % (16 * kWave) == 0, "whole 16-byte loads per lane");
// the longest run (PATCHED_BASE: 4 header + 8 base bytes, 512 64-bit values,
// 31 64-bit patch entries) fits a window staged 15 bytes before it
static_assert(kPlainWin >= 15 + 4 + 8 + 512 * 8 + 31 * 8, "a run fits a window");

// Copies the 16-byte-aligned absolute range [gb, gb + kPlainWin) of the
// stream buffer [lo, hi) into the wave's window (bytes outside it read 0).
MDZ_DEV void plain_stage(uint8_t* win, uintptr_t gb, uintptr_t lo, uintptr_t hi) {
    const uint32_t l = lane_id();
    u32x4 v[kPlainWin / 16 / kWave];
#pragma unroll
    for (uint32_t j = 0; j < kPlainWin / 16 / kWave; ++j) {
        const uintptr_t a = gb + 16 * (j * kWave + l);
        if (a >= lo && a + 16 <= hi) {
            v[j] = *reinterpret_cast<const u32x4*>(a);
        } else {
            uint32_t w[4] = {0, 0, 0, 0};
            for (uint32_t b = 0; b < 16; ++b)
                if (a + b >= lo && a + b < hi) w[b >> 2] |= (uint32_t)*reinterpret_cast<const uint8_t*>(a + b)
                                                            << (8 * (b & 3));
            v[j] = u32x4{w[0], w[1], w[2], w[3]};
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < kPlainWin / 16 / kWave; ++j)
        reinterpret_cast<u32x4*>(win)[j * kWave + l] = v[j];
    wave_sync();
}

// One value of the chunk: an 8-byte store, or the bytes below `limit`.
MDZ_DEV void plain_put(uint8_t* out, uint64_t limit, uint32_t q, uint64_t v) {
    const uint64_t off = 8ull * q;
    if (off + 8 <= limit) {
        *reinterpret_cast<uint64_t*>(out + off) = v;
    } else {
        for (uint64_t b = off; b < limit; ++b) out[b] = (uint8_t)(v >> (8 * (b - off)));
    }
}

__global__ __launch_bounds__(kPlainWaves * kWave) __attribute__((amdgpu_waves_per_eu(7))) void rlev2_plain_kernel(
    const uint8_t* __restrict__ data, uint64_t data_bytes, const uint64_t* __restrict__ ptr, uint64_t n_chunks,
    uint64_t chunk, uint8_t* __restrict__ dst, uint64_t dst_bytes, int32_t* __restrict__ status) {
    __shared__ __align__(16) PlainV2Lds S;
    const uint32_t wave = threadIdx.x / kWave, l = lane_id();
    const uint64_t c = (uint64_t)blockIdx.x * kPlainWaves + wave;
    if (c >= n_chunks) return;
    uint8_t* const win = S.win[wave];
    const uint64_t a0 = ptr[c], a1 = ptr[c + 1];
    if (a0 > a1 || a1 > data_bytes || a1 - a0 > 0x7fffffffull) {
        if (l == 0 && status) status[c] = MDZ_ST_BAD_HEADER;
        return;
    }
    const uintptr_t lo = reinterpret_cast<uintptr_t>(data), hi = lo + data_bytes;
    const uintptr_t s0 = lo + a0;                           // the stream's first byte
    const uint32_t slen = (uint32_t)(a1 - a0);
    const uint32_t cap = (uint32_t)(chunk / 8);
    uint8_t* const out = dst + c * chunk;
    const uint64_t limit = dst_bytes > c * chunk ? min(dst_bytes - c * chunk, chunk) : 0;

    uintptr_t gb = 0;                                       // window base (absolute), none yet
    bool staged = false;
    // the window holds stream bytes [p, p + n)
    auto covers = [&](uint32_t p, uint32_t n) {
        return staged && s0 + p >= gb && s0 + p + n <= gb + kPlainWin;
    };
    auto restage = [&](uint32_t p) {
        gb = (s0 + p) & ~(uintptr_t)15;
        staged = true;
        plain_stage(win, gb, lo, hi);
    };

    uint32_t p = 0, k = 0;
    int32_t st = MDZ_ST_OK;
    while (p < slen) {
        const uint32_t avail = slen - p;
        if (!covers(p, min(kPlainHdr, avail))) restage(p);
        uint32_t lp = (uint32_t)(s0 + p - gb);                // p's LDS offset
        // header bytes are wave-uniform: keep them (and all that derives
        // from them) in scalar registers
        const uint32_t first = __builtin_amdgcn_readfirstlane(win[lp]);
        const uint32_t kind = first >> 6;
        uint32_t n = 0, size = 0, w = 0, code = 0;
        uint32_t nb = 0, n1 = 0, n2 = 0, bw = 0, pw = 0, pll = 0, cfb = 0;
        uint64_t va = 0, vb = 0;
        if (kind == 0) {                                    // SHORT_REPEAT
            nb = ((first >> 3) & 7) + 1;
            n = (first & 7) + 3;
            size = 1 + nb;
            if (size > avail) st = MDZ_ST_INPUT_OVERRUN;
            else if (n > cap - k) st = MDZ_ST_OUTPUT_OVERRUN;
        } else if (avail < 2) {
            st = MDZ_ST_INPUT_OVERRUN;
        } else {
            code = (first >> 1) & 31;
            w = decode_width(code);
            n = (((first & 1) << 8) | __builtin_amdgcn_readfirstlane(win[lp + 1])) + 1;
            if (kind == 1) {                                // DIRECT
                size = 2 + (n * w + 7) / 8;
                if (n > cap - k) st = MDZ_ST_OUTPUT_OVERRUN;
                else if (size > avail) st = MDZ_ST_INPUT_OVERRUN;
            } else if (kind == 3) {                         // DELTA
                const uint32_t hend = lp + min(kPlainHdr, avail);
                int32_t e = MDZ_ST_OK;
                if (n < 2) {
                    st = MDZ_ST_BAD_SYMBOL;
                } else if (!(n1 = __builtin_amdgcn_readfirstlane(uvarint(win, lp + 2, hend, va, e)))) {
                    st = __builtin_amdgcn_readfirstlane(e);
                } else if (!(n2 = __builtin_amdgcn_readfirstlane(uvarint(win, lp + 2 + n1, hend, vb, e)))) {
                    st = __builtin_amdgcn_readfirstlane(e);
                } else {
                    const uint32_t nd = code ? n - 2 : 0;
                    size = 2 + n1 + n2 + (nd * w + 7) / 8;
                    if (n > cap - k) st = MDZ_ST_OUTPUT_OVERRUN;
                    else if (size > avail) st = MDZ_ST_INPUT_OVERRUN;
                }
            } else {                                        // PATCHED_BASE
                if (avail < 4) {
                    st = MDZ_ST_INPUT_OVERRUN;
                } else {
                    const uint32_t third = __builtin_amdgcn_readfirstlane(win[lp + 2]),
                                   fourth = __builtin_amdgcn_readfirstlane(win[lp + 3]);
                    bw = ((third >> 5) & 7) + 1;
                    pw = decode_width(third & 31);
                    pll = fourth & 31;
                    cfb = closest_bits(pw + ((fourth >> 5) & 7) + 1);
                    size = 4 + bw + (n * w + 7) / 8 + (pll * cfb + 7) / 8;
                    if (n > cap - k) st = MDZ_ST_OUTPUT_OVERRUN;
                    else if (size > avail) st = MDZ_ST_INPUT_OVERRUN;
                }
            }
        }
        if (st != MDZ_ST_OK) break;
        if (!covers(p, size)) {
            restage(p);
            lp = (uint32_t)(s0 + p - gb);
        }
        if (kind == 0) {
            const uint64_t v = lds_bits_be32(win, lp + 1, 0, 8 * nb);
            if (l < n) plain_put(out, limit, k + l, v);
        } else if (kind == 1) {
            for (uint32_t i0 = 0; i0 < n; i0 += kWave) {
                const uint32_t i = i0 + l;
                if (i < n) plain_put(out, limit, k + i, lds_bits_be32(win, lp + 2, i * w, w));
            }
        } else if (kind == 3) {
            const uint64_t base = va, bd = (vb >> 1) ^ (0 - (vb & 1));
            if (code == 0) {
                for (uint32_t i0 = 0; i0 < n; i0 += kWave) {
                    const uint32_t i = i0 + l;
                    if (i < n) plain_put(out, limit, k + i, base + (uint64_t)i * bd);
                }
            } else {
                const uint32_t nd = n - 2, pd = lp + 2 + n1 + n2;
                const uint64_t v1 = base + bd;
                const bool up = (int64_t)bd > 0;            // decoder.h:115
                if (l < 2) plain_put(out, limit, k + l, l ? v1 : base);
                uint64_t carry = v1;
                for (uint32_t j0 = 0; j0 < nd; j0 += kWave) {
                    const uint32_t j = j0 + l;
                    uint64_t d = j < nd ? lds_bits_be32(win, pd, j * w, w) : 0;
                    if (!up) d = 0 - d;
                    const uint64_t incl = wave_incl_scan_u64(d);
                    if (j < nd) plain_put(out, limit, k + 2 + j, carry + incl);
                    carry += last_lane_u64(incl);
                }
            }
        } else {
            const uint64_t base = lds_bits_be32(win, lp + 4, 0, 8 * bw);
            const uint32_t pb = lp + 4 + bw, pp = pb + (n * w + 7) / 8;
            const uint64_t e = l < pll ? lds_bits_be32(win, pp, l * cfb, cfb) : 0;
            const uint32_t gap = (l < pll && pw < 64) ? (uint32_t)(e >> pw) : 0;
            const uint32_t pos = wave_incl_scan_u32(gap);  // gaps are below 2^8 each
            if (pll && (uint32_t)__builtin_amdgcn_readlane(pos, pll - 1) >= n) {
                st = MDZ_ST_BAD_SYMBOL;
                break;
            }
            const uint64_t pv = pw >= 64 ? e : e & ((1ull << pw) - 1);
            for (uint32_t i0 = 0; i0 < n; i0 += kWave) {
                const uint32_t i = i0 + l;
                uint64_t v = i < n ? lds_bits_be32(win, pb, i * w, w) : 0;
                for (uint32_t j = 0; j < pll; ++j) {
                    const uint32_t pj = __builtin_amdgcn_readlane(pos, j);
                    const uint32_t lo32 = (uint32_t)__builtin_amdgcn_readlane((uint32_t)pv, j);
                    const uint32_t hi32 = (uint32_t)__builtin_amdgcn_readlane((uint32_t)(pv >> 32), j);
                    if (pj == i && w < 64) v |= ((uint64_t)hi32 << 32 | lo32) << w;
                }
                if (i < n) plain_put(out, limit, k + i, v + base);
            }
        }
        p += size;
        k += n;
    }
    if (l == 0 && status) status[c] = st;
}

}  // namespace

int rlev2_plain_decompress(const void* data, uint64_t data_bytes, const uint64_t* ptr, uint64_t n_chunks,
                           uint64_t chunk, void* dst, uint64_t dst_bytes, int32_t* status, hipStream_t s) {
    if (chunk == 0 || chunk % 8 || chunk / 8 > 0x7fffffffull)
        return set_error(MDZ_ERR_INVALID, "rlev2 plain: chunk %llu bytes (a multiple of 8 below 16 GiB)",
                         (unsigned long long)chunk);
    if (n_chunks == 0) return MDZ_OK;
    if (!data || !ptr || (!dst && dst_bytes)) return set_error(MDZ_ERR_INVALID, "rlev2 plain: null pointer");
    if (reinterpret_cast<uintptr_t>(dst) & 7)
        return set_error(MDZ_ERR_INVALID, "rlev2 plain: output not 8-byte aligned");
    const uint64_t grid = (n_chunks + kPlainWaves - 1) / kPlainWaves;
    if (grid > 0x7fffffffull) return set_error(MDZ_ERR_INVALID, "rlev2 plain: %llu chunks",
                                               (unsigned long long)n_chunks);
    hipLaunchKernelGGL(rlev2_plain_kernel, dim3((uint32_t)grid), dim3(kPlainWaves * kWave), 0, s,
                       static_cast<const uint8_t*>(data), data_bytes, ptr, n_chunks, chunk,
                       static_cast<uint8_t*>(dst), dst_bytes, status);
    return record_hip(hipGetLastError(), "rlev2_plain_kernel launch");
}

}  // namespace mdz
