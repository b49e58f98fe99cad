// mdz device helpers for gfx950 (CDNA4, wave64).
//
// Building blocks shared by the lane-transposed decoders (RLE v1, RLE v2,
// LZSS): the per-chunk lane-address table that locates lane t's r-th 4-byte
// word inside the compacted round-robin interleave, the LDS staging ring, a
// byte buffer fed from that ring, and a line emitter that turns a lane's
// value stream into 16-byte (or narrower) global stores.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mdz.h"

#define MDZ_DEV __device__ __forceinline__

// Phase timing for tuning builds.  With MDZ_PHASE_PROF defined (never in the
// library build), thread 0 of each workgroup adds the shader-clock cycles
// spent between consecutive marks to mdz_phase_cycles[k] (a __device__ array
// the profiling translation unit defines); otherwise the marks vanish.
#ifdef MDZ_PHASE_PROF
extern __device__ unsigned long long mdz_phase_cycles[16];
#define MDZ_PHASE_START() uint64_t mdz_phase_t_ = clock64()
#define MDZ_PHASE(k)                                                              \
    do {                                                                          \
        const uint64_t mdz_now_ = clock64();                                      \
        if (threadIdx.x == 0) atomicAdd(&mdz_phase_cycles[k], mdz_now_ - mdz_phase_t_); \
        mdz_phase_t_ = mdz_now_;                                                  \
    } while (0)
#else
#define MDZ_PHASE_START() ((void)0)
#define MDZ_PHASE(k) ((void)0)
#endif

namespace mdz {

constexpr int kWave = 64;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kLanes = 32;   // sub-streams per chunk, fixed by the container format

MDZ_DEV uint32_t lane_id() { return __lane_id(); }

// v of the lane the DPP control names (row_shr:n = 0x110 + n within each row
// of 16 lanes, row_bcast:15 = 0x142, row_bcast:31 = 0x143); kFill where the
// control names no lane or the row is masked off.  The scans and reductions
// below are built from these, in registers, with every lane active.
template <int kCtrl, int kRowMask, uint32_t kFill>
MDZ_DEV uint32_t dpp_src(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)kFill, (int)v, kCtrl, kRowMask, 0xf, false);
}

// Inclusive min-scan; lane 63 holds the wave minimum.
MDZ_DEV uint32_t wave_incl_min_u32(uint32_t v) {
    v = min(v, dpp_src<0x111, 0xf, 0xffffffffu>(v));
    v = min(v, dpp_src<0x112, 0xf, 0xffffffffu>(v));
    v = min(v, dpp_src<0x114, 0xf, 0xffffffffu>(v));
    v = min(v, dpp_src<0x118, 0xf, 0xffffffffu>(v));
    v = min(v, dpp_src<0x142, 0xa, 0xffffffffu>(v));
    v = min(v, dpp_src<0x143, 0xc, 0xffffffffu>(v));
    return v;
}

MDZ_DEV uint32_t wave_incl_scan_u32(uint32_t v);

MDZ_DEV uint32_t wave_min_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_min_u32(v), kWave - 1);
}

MDZ_DEV uint32_t wave_sum_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan_u32(v), kWave - 1);
}

MDZ_DEV uint32_t alignbyte(uint32_t hi, uint32_t lo, uint32_t k) {
    return __builtin_amdgcn_alignbyte(hi, lo, k);
}

// ------------------------------------------------------------------------
// Lane-address table.  In a chunk, round r stores word r of every lane that
// has more than r words, in lane order.  Sort the 32 word counts W ascending
// (ties by lane id): sW[0] <= ... <= sW[31].  Rounds [sW[k-1], sW[k]) form
// segment k, in which exactly the lanes of sorted rank >= k are active
// (mask M[k], 32-k of them).  S[k] = raw word index where segment k starts.
// Lane t's word r, inside segment k, is at
//     S[k] + (r - sW[k-1]) * (32-k) + popc(M[k] & lanes_below(t)).
struct LaneTable {
    uint32_t sW[kLanes];
    uint32_t S[kLanes];
    uint32_t M[kLanes];
    uint32_t W[kLanes];
    uint32_t rank[kLanes];
    uint32_t total;     // sum of W: words the chunk's lanes occupy
};

// Orders LDS traffic among the lanes of one wave (no cross-wave effect).
MDZ_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Builds the table.  Called by all 64 lanes of ONE wave (lanes 32..63 mirror
// 0..31); `W` is the calling lane's word count.  Uses only wave-level
// synchronisation, so other waves of the workgroup need not take part.
MDZ_DEV void lane_table_build(LaneTable& T, uint32_t W) {
    const uint32_t l = lane_id(), t = l & 31;
    if (l < 32) T.W[t] = W;
    wave_sync();
    uint32_t rk = 0;
#pragma unroll 8
    for (uint32_t u = 0; u < 32; ++u) {
        uint32_t Wu = T.W[u];
        rk += (Wu < W) || (Wu == W && u < t);
    }
    if (l < 32) { T.sW[rk] = W; T.rank[t] = rk; }
    wave_sync();
    if (l < 32) {
        uint32_t m = 0;
#pragma unroll 8
        for (uint32_t u = 0; u < 32; ++u) m |= (T.rank[u] >= t ? 1u : 0u) << u;
        uint32_t s = 0, prev = 0;
        for (uint32_t j = 0; j < t; ++j) {
            uint32_t w = T.sW[j];
            s += (w - prev) * (32 - j);
            prev = w;
        }
        T.M[t] = m;
        T.S[t] = s;
    }
    uint32_t tot = wave_sum_u32(l < 32 ? W : 0);
    if (l == 0) T.total = tot;
    wave_sync();
}

// Position (0-based) of the q-th set bit of m (q < popc(m)).
MDZ_DEV uint32_t select_bit(uint32_t m, uint32_t q) {
    uint32_t pos = 0, c;
    c = __popc(m & 0xffffu); if (q >= c) { q -= c; m >>= 16; pos += 16; }
    c = __popc(m & 0xffu);   if (q >= c) { q -= c; m >>= 8;  pos += 8; }
    c = __popc(m & 0xfu);    if (q >= c) { q -= c; m >>= 4;  pos += 4; }
    c = __popc(m & 0x3u);    if (q >= c) { q -= c; m >>= 2;  pos += 2; }
    c = m & 1u;              if (q >= c) { pos += 1; }
    return pos;
}

// Inverse of the lane-address map: raw word w (< T.total) -> (lane, round).
MDZ_DEV void lane_of_word(const LaneTable& T, uint32_t w, uint32_t& lane, uint32_t& round) {
    if (w < T.S[1]) {
        lane = w & 31; round = w >> 5;          // segment 0: all 32 lanes active
        return;
    }
    uint32_t lo = 0, hi = 31;
    while (lo < hi) {                            // largest k with S[k] <= w
        uint32_t mid = (lo + hi + 1) >> 1;
        if (T.S[mid] <= w) lo = mid; else hi = mid - 1;
    }
    const uint32_t k = lo, nk = 32 - k;
    const uint32_t off = w - T.S[k];
    const uint32_t q = off / nk;
    round = (k ? T.sW[k - 1] : 0) + q;
    lane = select_bit(T.M[k], off - q * nk);
}

// Per-lane cursor over its own words.
struct LaneCursor {
    uint32_t r;        // next word to fetch
    uint32_t W;        // word count
    uint32_t k;        // current segment
    uint32_t base, stride, seg_end;

    MDZ_DEV void enter(const LaneTable& T, uint32_t t) {
        stride = 32 - k;
        uint32_t prevW = k ? T.sW[k - 1] : 0;
        base = T.S[k] - prevW * stride + __popc(T.M[k] & ((1u << t) - 1u));
        seg_end = T.sW[k];
    }
    MDZ_DEV void init(const LaneTable& T, uint32_t t, uint32_t words) {
        r = 0; W = words; k = 0;
        if (W) {
            while (T.sW[k] == 0) ++k;
            enter(T, t);
        }
    }
    MDZ_DEV uint32_t addr() const { return base + r * stride; }
    MDZ_DEV void advance(const LaneTable& T, uint32_t t) {
        ++r;
        if (r < W && r >= seg_end) {
            do { ++k; } while (T.sW[k] <= r);
            enter(T, t);
        }
    }
};

// ------------------------------------------------------------------------
// Stage raw chunk words [a, b) (both multiples of 4) into the LDS ring of
// `ring_words` (power of 2) with 16-byte loads; all 64 lanes participate.
MDZ_DEV void ring_stage(uint32_t* ring, uint32_t ring_words, const u32x4* __restrict__ src,
                        uint32_t a, uint32_t b) {
    const uint32_t l = lane_id();
    const uint32_t gmask = ring_words / 4 - 1;
    u32x4* ring4 = reinterpret_cast<u32x4*>(ring);
    uint32_t g = a / 4 + l, ge = b / 4;
    for (; g + 3 * kWave < ge; g += 4 * kWave) {
        u32x4 v0 = __builtin_nontemporal_load(src + g);
        u32x4 v1 = __builtin_nontemporal_load(src + g + kWave);
        u32x4 v2 = __builtin_nontemporal_load(src + g + 2 * kWave);
        u32x4 v3 = __builtin_nontemporal_load(src + g + 3 * kWave);
        ring4[g & gmask] = v0;
        ring4[(g + kWave) & gmask] = v1;
        ring4[(g + 2 * kWave) & gmask] = v2;
        ring4[(g + 3 * kWave) & gmask] = v3;
    }
    for (; g < ge; g += kWave) ring4[g & gmask] = __builtin_nontemporal_load(src + g);
}

// ------------------------------------------------------------------------
// Byte buffer: up to 8 (NB=8) or 16 (NB=16) stream bytes, little-endian,
// byte 0 = next unread stream byte.
template <int NB> struct ByteBuf;

template <> struct ByteBuf<8> {
    uint64_t v = 0;
    uint32_t nb = 0;
    MDZ_DEV void append(uint32_t w, uint32_t nbytes) { v |= (uint64_t)w << (8 * nb); nb += nbytes; }
    MDZ_DEV uint64_t lo() const { return v; }
    MDZ_DEV uint64_t hi() const { return 0; }
    MDZ_DEV void consume(uint32_t n) {
        v = n >= 8 ? 0 : v >> (8 * n);
        nb -= n;
    }
};

template <> struct ByteBuf<16> {
    unsigned __int128 v = 0;
    uint32_t nb = 0;
    MDZ_DEV void append(uint32_t w, uint32_t nbytes) {
        v |= (unsigned __int128)w << (8 * nb);
        nb += nbytes;
    }
    MDZ_DEV uint64_t lo() const { return (uint64_t)v; }
    MDZ_DEV uint64_t hi() const { return (uint64_t)(v >> 64); }
    MDZ_DEV void consume(uint32_t n) {
        v = n >= 16 ? 0 : v >> (8 * n);
        nb -= n;
    }
};

// ------------------------------------------------------------------------
// Line emitter: lane t owns lines t, t+32, t+64, ... of `line` bytes in its
// chunk's output.  Values are shifted into a 16-byte register window; every
// SW bytes are written with one store (SW = 16, 8, 4, 2 or 1, dividing line).
template <int EB, int SW> struct LineEmitter {
    uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    uint32_t na = 0;          // bytes pending in the window
    uint8_t* out;             // chunk output base
    uint64_t opos;            // chunk byte offset of the pending group
    uint64_t limit;           // bytes of this chunk that may be written
    uint32_t lrem;            // bytes left in the current line
    uint32_t line;

    // Starts at byte `lane_off` of lane t's own byte sequence (a multiple of SW).
    MDZ_DEV void init(uint8_t* chunk_out, uint64_t lim, uint32_t t, uint32_t line_bytes,
                      uint64_t lane_off) {
        out = chunk_out; limit = lim; line = line_bytes;
        const uint64_t ln = lane_off / line_bytes, in_line = lane_off - ln * line_bytes;
        opos = (ln * 32 + t) * line_bytes + in_line;
        lrem = line_bytes - (uint32_t)in_line;
        na = 0;
    }
    MDZ_DEV void push(uint64_t v) {
        if constexpr (EB == 4) {
            a0 = a1; a1 = a2; a2 = a3; a3 = (uint32_t)v;
        } else if constexpr (EB == 8) {
            a0 = a2; a1 = a3; a2 = (uint32_t)v; a3 = (uint32_t)(v >> 32);
        } else if constexpr (EB == 2) {
            a0 = alignbyte(a1, a0, 2); a1 = alignbyte(a2, a1, 2); a2 = alignbyte(a3, a2, 2);
            a3 = (a3 >> 16) | ((uint32_t)v << 16);
        } else {
            a0 = alignbyte(a1, a0, 1); a1 = alignbyte(a2, a1, 1); a2 = alignbyte(a3, a2, 1);
            a3 = (a3 >> 8) | ((uint32_t)v << 24);
        }
        na += EB;
        if (na == SW) flush();
    }
    MDZ_DEV void flush() {
        uint8_t* p = out + opos;
        if (opos + SW <= limit) {
            if constexpr (SW == 16) {
                *reinterpret_cast<uint4*>(p) = make_uint4(a0, a1, a2, a3);
            } else if constexpr (SW == 8) {
                *reinterpret_cast<uint2*>(p) = make_uint2(a2, a3);
            } else if constexpr (SW == 4) {
                *reinterpret_cast<uint32_t*>(p) = a3;
            } else if constexpr (SW == 2) {
                *reinterpret_cast<uint16_t*>(p) = (uint16_t)(a3 >> 16);
            } else {
                *p = (uint8_t)(a3 >> 24);
            }
        } else if (opos < limit) {
            // tail of a clipped output: bytewise
            uint32_t w[4] = {a0, a1, a2, a3};
            for (uint32_t i = 0; i < SW && opos + i < limit; ++i) {
                uint32_t bi = 16 - SW + i;
                p[i] = (uint8_t)(w[bi >> 2] >> (8 * (bi & 3)));
            }
        }
        na = 0;
        opos += SW;
        lrem -= SW;
        if (lrem == 0) { opos += (uint64_t)31 * line; lrem = line; }
    }
};

}  // namespace mdz

namespace mdz {

// ------------------------------------------------------------------------
// Whole-chunk staging for the lane-transposed codecs: every thread of the
// workgroup reads the chunk's interleaved payload with 16-byte loads (each
// byte once) and scatters each 4-byte word to its lane's contiguous LDS
// region at d32[lbase[lane]/4 + round].
MDZ_DEV void stage_chunk_lanes(const LaneTable& T, const uint32_t* lbase, uint32_t* d32,
                               const u32x4* __restrict__ src, uint32_t tid, uint32_t nthreads) {
    const uint32_t total = T.total, ngroups = (total + 3) / 4;
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
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t w = 4 * g + i;
                if (w < total) {
                    uint32_t lane, round;
                    lane_of_word(T, w, lane, round);
                    d32[(lbase[lane] >> 2) + round] = v[u][i];
                }
            }
        }
    }
}

// Wave-wide inclusive prefix sums.
template <int kCtrl, int kRowMask>
MDZ_DEV uint64_t dpp_src_u64(uint64_t v) {
    return ((uint64_t)dpp_src<kCtrl, kRowMask, 0u>((uint32_t)(v >> 32)) << 32) |
           dpp_src<kCtrl, kRowMask, 0u>((uint32_t)v);
}
MDZ_DEV uint64_t wave_incl_scan_u64(uint64_t v) {       // as wave_incl_scan_u32, in 64 bits
    v += dpp_src_u64<0x111, 0xf>(v);
    v += dpp_src_u64<0x112, 0xf>(v);
    v += dpp_src_u64<0x114, 0xf>(v);
    v += dpp_src_u64<0x118, 0xf>(v);
    v += dpp_src_u64<0x142, 0xa>(v);
    v += dpp_src_u64<0x143, 0xc>(v);
    return v;
}
// Inclusive prefix sum over the 64 lanes, in registers with DPP (call with
// every lane active): row_shr 1, 2, 4, 8 sum within each row of 16 lanes
// (sources outside the row read 0), then row_bcast:15 adds row 0's total to
// row 1 and row 2's to row 3, and row_bcast:31 adds lane 31 (rows 0 and 1)
// to rows 2 and 3.
MDZ_DEV uint32_t wave_incl_scan_u32(uint32_t v) {
    v += dpp_src<0x111, 0xf, 0u>(v);      // row_shr:1
    v += dpp_src<0x112, 0xf, 0u>(v);      // row_shr:2
    v += dpp_src<0x114, 0xf, 0u>(v);      // row_shr:4
    v += dpp_src<0x118, 0xf, 0u>(v);      // row_shr:8
    v += dpp_src<0x142, 0xa, 0u>(v);      // row_bcast:15
    v += dpp_src<0x143, 0xc, 0u>(v);      // row_bcast:31
    return v;
}
// v of lane 63 (wave-uniform).
MDZ_DEV uint64_t last_lane_u64(uint64_t v) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), kWave - 1) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, kWave - 1);
}
MDZ_DEV uint64_t shfl_u64(uint64_t v, int src) {
    const uint32_t lo = __shfl((uint32_t)v, src, kWave), hi = __shfl((uint32_t)(v >> 32), src, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// 8 stream bytes (little-endian) at byte p of a 4-byte-aligned LDS region.
MDZ_DEV uint64_t lds_peek8(const uint8_t* base, uint32_t p) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(base) + (p >> 2);
    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], k = p & 3;
    return ((uint64_t)alignbyte(w2, w1, k) << 32) | alignbyte(w1, w0, k);
}

// MSB-first field of w bits (1..64) starting `bit` bits after byte b0 of a
// 4-byte-aligned LDS region (reads bytes up to 12 past the field start).
MDZ_DEV uint64_t lds_bits_be(const uint8_t* base, uint32_t b0, uint64_t bit, uint32_t w) {
    const uint32_t b = b0 + (uint32_t)(bit >> 3), s = (uint32_t)(bit & 7), k = b & 3;
    const uint32_t* d = reinterpret_cast<const uint32_t*>(base) + (b >> 2);
    const uint32_t d0 = d[0], d1 = d[1], d2 = d[2];
    const uint32_t lo = alignbyte(d1, d0, k), mid = alignbyte(d2, d1, k);
    const uint64_t be = ((uint64_t)__builtin_bswap32(lo) << 32) | __builtin_bswap32(mid);
    const uint64_t v = s ? (be << s) | (((d2 >> (8 * k)) & 0xffu) >> (8 - s)) : be;
    return w >= 64 ? v : v >> (64 - w);
}

}  // namespace mdz
