// Lane-transposed container geometry shared by the RLE v1 and RLE v2
// decoders: the per-lane setup of a chunk and the scatter of its
// interleaved 4-byte words into contiguous per-lane LDS regions.
//
// Container shape (include/rle_v1/encoder.h:540-588,
// include/rlev2/encoder_trans.h:291-349): chunk c's payload starts at
// blk_off[c] and holds 32 lane streams of col_len[32c+t] bytes; round r
// stores word r of every lane with more than r words, in lane order.
#pragma once
#include "mdz_device.h"

namespace mdz {
namespace {

constexpr uint32_t kNoLane = 0xffffffffu;      // lbase / dest of a lane outside the part
constexpr uint32_t kSegBlocks = 160 * 1024 / 128;  // 32-word blocks of the largest LDS payload

// Per-lane setup of chunk c, run by the 64 lanes of wave 0 (lanes 32..63
// mirror 0..31).  Lanes [t0, t0 + NL) are the part being decoded.  Fills
//   lbase[t]  LDS byte offset of lane t's region (16-byte aligned; kNoLane
//             outside the part), cl[t] its stream bytes, lerr[t] = ~0;
//   sW, S, M  the interleave in segments: lane word counts ascending; segment
//             k holds rounds sW[k-1] .. sW[k]-1 with the 32 - k lanes of mask
//             M[k] active and starts at raw word S[k];
//   total     words of all lanes; extent_ok: they lie inside the chunk's
//             payload extent [c_off, c_end) and c_off is 16-byte aligned;
//   fits      the part's regions fit data_cap bytes.
// Returns the bytes the part's regions occupy (every lane rounded to 16).
template <uint32_t NL, class Tables>
MDZ_DEV uint32_t lane_setup(Tables& tb, const uint64_t* __restrict__ col_len, uint64_t c,
                            uint32_t t0, uint32_t data_cap, uint64_t c_off, uint64_t c_end) {
    const uint32_t l = lane_id(), t = l & 31;
    const uint32_t cl = (uint32_t)col_len[c * 32 + t];
    const bool own = t - t0 < NL;
    const uint32_t W = (cl + 3) / 4, r16 = own ? (cl + 15) & ~15u : 0u;
    uint32_t incl = r16, tot = W;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, kWave);
        if (t >= (uint32_t)o) incl += y;
        tot += __shfl_xor(tot, o, kWave);
    }
    // sort the lane lengths: rank = lanes shorter, ties to the lower lane
    uint32_t rk = 0;
#pragma unroll
    for (int u = 0; u < 32; ++u) {
        const uint32_t wu = __builtin_amdgcn_readlane(W, u);
        rk += (wu < W) || (wu == W && (uint32_t)u < t);
    }
    if (l < 32) {
        tb.lbase[t] = own ? incl - r16 : kNoLane;
        tb.cl[t] = cl;
        tb.lerr[t] = 0xffffffffu;
        tb.sW[rk] = W;
    }
    // segment k's active lanes: rank >= k
    uint32_t mk = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        const uint32_t m = (uint32_t)__ballot(l < 32 && rk >= (uint32_t)k);
        mk = t == (uint32_t)k ? m : mk;
    }
    wave_sync();
    // S[k] = raw words before round sW[k-1] = sum over lanes of min(W, sW[k-1])
    const uint32_t prev = t ? tb.sW[t - 1] : 0;
    uint32_t sk = 0;
#pragma unroll
    for (int u = 0; u < 32; ++u) sk += min((uint32_t)__builtin_amdgcn_readlane(W, u), prev);
    if (l < 32) { tb.S[t] = sk; tb.M[t] = mk; }
    if (l == 31) tb.fits = incl <= data_cap;
    if (l == 0) {
        tb.total = tot;
        tb.extent_ok = !(c_off & 15) && c_end >= c_off && (uint64_t)tot * 4 <= c_end - c_off;
    }
    return (uint32_t)__builtin_amdgcn_readlane(incl, 31);
}

// Segment of raw word w: the number of segment starts S[1..31] at or below
// w (the starts ascend, so this is the last segment starting at or below w).
MDZ_DEV uint32_t segment_of(const uint32_t* S, uint32_t w) {
    uint32_t k = 0;
#pragma unroll
    for (int j = 1; j < 32; ++j) k += w >= S[j];
    return k;
}

// Destination table and segment-of-block table for stage_lanes (whole
// chunks), by all NT threads of the workgroup; barriers must separate it
// from lane_setup and from stage_lanes.
template <class Tables>
MDZ_DEV void stage_tables(Tables& tb, uint32_t tid, uint32_t NT) {
    for (uint32_t x = tid; x < 32 * 32; x += NT) {
        const uint32_t k = x >> 5, jj = x & 31;
        if (jj < 32 - k) {
            const uint32_t lb = tb.lbase[select_bit(tb.M[k], jj)];
            tb.st.dest[k][jj] = lb == kNoLane ? kNoLane : (lb >> 2) + (k ? tb.sW[k - 1] : 0);
        }
    }
    uint32_t S[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) S[j] = __builtin_amdgcn_readfirstlane(tb.S[j]);
    for (uint32_t b = tid; b < (tb.total + 31) / 32; b += NT)
        tb.st.segof[b] = (uint8_t)segment_of(S, 32 * b);
}

// Scatters the chunk's interleaved words into the lane regions of the part
// with 16-byte loads: raw word w of segment k is the j-th active lane's word of round
// sW[k-1] + (w - S[k]) / (32 - k), j = (w - S[k]) mod (32 - k), stored at
// dest[k][j] + (w - S[k]) / (32 - k).  A group's segment comes from the
// table of 32-word blocks; the quotient is found once per group and then
// stepped word by word.
template <class Tables>
MDZ_DEV void stage_lanes(const Tables& tb, uint32_t* d32, const u32x4* __restrict__ src,
                         uint32_t tid, uint32_t nthreads) {
    const uint32_t total = tb.total, ngroups = (total + 3) / 4;
    auto next_start = [&](uint32_t k) { return k < 31 ? tb.S[k + 1] : total; };
    for (uint32_t g0 = tid; g0 < ngroups; g0 += 4 * nthreads) {
        u32x4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t g = g0 + u * nthreads;
            if (g < ngroups) v[u] = __builtin_nontemporal_load(src + g);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t g = g0 + u * nthreads;
            if (g >= ngroups) break;
            const uint32_t w0 = 4 * g;
            uint32_t k = tb.st.segof[g >> 3];
            uint32_t sn = next_start(k);
            while (w0 >= sn) sn = next_start(++k);
            uint32_t nk = 32 - k, off = w0 - tb.S[k];
            uint32_t r = (uint32_t)((float)off * __builtin_amdgcn_rcpf((float)nk));
            if (r * nk > off) --r;
            else if ((r + 1) * nk <= off) ++r;
            uint32_t j = off - r * nk;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t w = w0 + e;
                if (w >= total) break;
                if (w >= sn) {                       // the group crosses a segment start
                    while (w >= sn) sn = next_start(++k);
                    nk = 32 - k;
                    off = w - tb.S[k];
                    r = (uint32_t)((float)off * __builtin_amdgcn_rcpf((float)nk));
                    if (r * nk > off) --r;
                    else if ((r + 1) * nk <= off) ++r;
                    j = off - r * nk;
                }
                const uint32_t dw = tb.st.dest[k][j];
                if (dw != kNoLane) d32[dw + r] = v[u][e];     // words of other parts are dropped
                if (++j == nk) { j = 0; ++r; }
            }
        }
    }
}

// Stages only the part's lanes (NL < 32).  Thread (j, x) takes a contiguous
// block of rounds of lane t0 + j.  Round r lies in segment
// k = #{lanes with W <= r} (sW ascending), whose active lanes are M[k], and
// the word's raw index is base_k + r * (32 - k) with
// base_k = S[k] - sW[k-1] * (32 - k) + popc(M[k] below the lane): affine in r
// inside a segment, so the tables are read only where a segment starts.  A
// thread issues up to 16 loads before their LDS stores.
template <uint32_t NL, uint32_t NT, class Tables>
MDZ_DEV void stage_part(const Tables& tb, uint32_t* d32, const uint32_t* __restrict__ src, uint32_t t0,
                        uint32_t tid) {
    constexpr uint32_t XS = NT / NL;              // threads per lane
    const uint32_t j = tid % NL, x = tid / NL, t = t0 + j;
    const uint32_t cl = tb.cl[t], Wt = (cl + 3) / 4, lb = tb.lbase[t] >> 2;
    const uint32_t below = (1u << t) - 1u;
    const uint32_t rb = (Wt + XS - 1) / XS;
    uint32_t r = x * rb;
    const uint32_t rend = min(Wt, r + rb);
    if (r >= rend) return;
    // the segment of round r: count of sorted lengths <= r (scalar copies)
    uint32_t k = 0;
#pragma unroll
    for (int i = 0; i < 32; ++i) k += __builtin_amdgcn_readfirstlane(tb.sW[i]) <= r;
    uint32_t nbnd, base, stride;
    auto enter = [&]() {                          // k < 32 while lane t is active
        nbnd = k < 31 ? tb.sW[k] : 0xffffffffu;
        stride = 32 - k;
        base = tb.S[k] - (k ? tb.sW[k - 1] : 0) * stride + __popc(tb.M[k] & below);
    };
    enter();
    while (r < rend) {
        uint32_t v[16], at[16];
        uint32_t nb = 0;
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u) {
            if (r < rend) {
                if (r >= nbnd) {                  // next segment (a shorter lane ended)
                    while (k < 31 && tb.sW[k] <= r) ++k;
                    enter();
                }
                v[u] = __builtin_nontemporal_load(src + base + r * stride);
                at[u] = lb + r;
                ++nb;
                ++r;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u)
            if (u < nb) d32[at[u]] = v[u];
    }
}

}  // namespace
}  // namespace mdz
