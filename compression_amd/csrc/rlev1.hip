// RLE v1 decoder for gfx950.
//
// Replaces the reference's reader-warp / decoder-warp inflate kernel
// (include/rle_v1/decoder.h:426-470, decoder_warp :275-423).  Token grammar
// (decoder.h:296-411): int8 h < 0 -> -h LEB128 literals; h >= 0 -> int8 delta,
// LEB128 base, h+3 values base + i*delta (64-bit arithmetic, truncated to T).
//
// One workgroup (4 wave64s) per chunk:
//   1. stage: all 256 threads read the chunk's interleaved payload from HBM
//      once with 16-byte loads and scatter each 4-byte word to its lane's
//      contiguous region in LDS (the inverse of the compacted round-robin
//      interleave, via the lane-address table);
//   2. scan: one thread per lane stream walks its tokens -- literal groups
//      are skipped by counting LEB128 terminators 16 bytes at a time -- and
//      records the decoder state at every K-th value (8 segments per lane);
//   3. decode: 256 threads, one per (lane, segment), each decode exactly K
//      values from their checkpoint and write their lines with 16-byte stores.
// A chunk whose payload does not fit the LDS budget falls back to a
// single-wave decoder that streams the chunk through an LDS window.
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

enum : uint32_t { kHdr = 0, kDelta = 1, kBase = 2, kRun = 3, kLit = 4, kDone = 5 };
constexpr uint32_t kSegs = 16;           // segments per lane stream
constexpr uint32_t kThreads = 32 * kSegs;
constexpr uint32_t kBad = 0xffffffffu;

template <int EB> struct VarintSpec;
template <> struct VarintSpec<1> { static constexpr uint32_t kMax = 2; };
template <> struct VarintSpec<2> { static constexpr uint32_t kMax = 3; };
template <> struct VarintSpec<4> { static constexpr uint32_t kMax = 5; };
template <> struct VarintSpec<8> { static constexpr uint32_t kMax = 10; };

// Gather the 7-bit groups of the 8 bytes of x into one integer.
MDZ_DEV uint64_t gather7x8(uint64_t x) {
    return (x & 0x7full) | ((x >> 1) & (0x7full << 7)) | ((x >> 2) & (0x7full << 14)) |
           ((x >> 3) & (0x7full << 21)) | ((x >> 4) & (0x7full << 28)) |
           ((x >> 5) & (0x7full << 35)) | ((x >> 6) & (0x7full << 42)) |
           ((x >> 7) & (0x7full << 49));
}

// LEB128 decode from up to 16 peeked bytes (lo = bytes 0..7, hi = 8..15) of
// which `avail` are stream bytes.  Returns the length, or 0 with *err set:
// input overrun when the stream ends inside the value, malformed symbol when
// the value is longer than kMax bytes.  Mirrors the oracle's leb128().
template <int EB>
MDZ_DEV uint32_t varint_decode(uint64_t lo, uint64_t hi, uint32_t avail, uint64_t& val,
                               int32_t& err) {
    constexpr uint32_t kMax = VarintSpec<EB>::kMax;
    uint32_t len;
    if constexpr (kMax <= 8) {
        constexpr uint64_t tmask = kMax == 8 ? 0x8080808080808080ull
                                            : ((1ull << (8 * kMax)) - 1) & 0x8080808080808080ull;
        const uint64_t term = ~lo & tmask;
        len = term ? (__builtin_ctzll(term) >> 3) + 1 : 0;
    } else {
        const uint64_t tl = ~lo & 0x8080808080808080ull, th = ~hi & 0x8080ull;
        len = tl ? (__builtin_ctzll(tl) >> 3) + 1 : th ? 9 + (__builtin_ctzll(th) >> 3) : 0;
    }
    if (len == 0 || len > avail) {
        err = (len == 0 && avail >= kMax) ? MDZ_ST_BAD_SYMBOL : MDZ_ST_INPUT_OVERRUN;
        return 0;
    }
    uint64_t v = gather7x8(lo);
    if (len < 8) {
        v &= (1ull << (7 * len)) - 1;
    } else if constexpr (kMax > 8) {
        if (len > 8) {
            uint64_t h = (hi & 0x7f) | ((hi >> 1) & (0x7full << 7));
            h &= (1ull << (7 * (len - 8))) - 1;
            v |= h << 56;
        }
    }
    val = v;
    return len;
}

// 8 stream bytes starting at byte p of a 16-byte-aligned LDS lane region.
MDZ_DEV uint64_t peek8(const uint8_t* base, uint32_t p) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(base) + (p >> 2);
    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], k = p & 3;
    return ((uint64_t)alignbyte(w2, w1, k) << 32) | alignbyte(w1, w0, k);
}

template <int EB>
MDZ_DEV uint32_t read_varint(const uint8_t* base, uint32_t p, uint32_t end, uint64_t& val,
                             int32_t& err) {
    const uint64_t lo = peek8(base, p);
    uint64_t hi = 0;
    if constexpr (EB == 8) {
        if ((~lo & 0x8080808080808080ull) == 0) hi = peek8(base, p + 8);
    }
    return varint_decode<EB>(lo, hi, end > p ? end - p : 0, val, err);
}

// ---------------------------------------------------------------- tables --
struct SplitTables {
    uint32_t lbase[32];      // lane region byte offsets in the data area (16-aligned)
    uint32_t cl[32];
    uint32_t err;            // min over (lane*kSegs+segment) << 8 | status byte
    uint32_t fits;           // lane regions fit the LDS data area
    uint32_t extent_ok;      // lane lengths lie within the chunk's payload extent
    uint32_t nblk;           // 16-byte blocks in the data area in use
    union {
        LaneTable lt;                        // stage phase
        struct {                             // scan -> decode hand-off
            uint32_t p[kSegs - 1][32];
            uint32_t m[kSegs - 1][32];      // mode | rem << 8 | delta << 24
            uint32_t vlo[kSegs - 1][32];
            uint32_t vhi[kSegs - 1][32];
        } ck;
    };
};

MDZ_DEV uint32_t pack_err(uint32_t t, uint32_t j, int32_t e) {
    return ((t * kSegs + j) << 8) | ((uint32_t)e & 0xffu);
}

// LEB128 terminators (bytes with bit 7 clear) in a 16-byte block, as four
// dword masks, ignoring bytes before `from`.
MDZ_DEV void block_terms(u32x4 x, uint32_t from, uint32_t m[4]) {
    m[0] = ~x.x & 0x80808080u; m[1] = ~x.y & 0x80808080u;
    m[2] = ~x.z & 0x80808080u; m[3] = ~x.w & 0x80808080u;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int lo = (int)from - 4 * i;
        if (lo >= 4) m[i] = 0;
        else if (lo > 0) m[i] &= 0xffffffffu << (8 * lo);
    }
}

// Byte offset (within the block) just past the k-th set terminator (k >= 1).
MDZ_DEV uint32_t block_select(const uint32_t m[4], uint32_t k) {
    uint32_t di = 0, mm = m[0], c = __popc(m[0]);
    if (k > c) { k -= c; di = 1; mm = m[1]; c = __popc(m[1]);
        if (k > c) { k -= c; di = 2; mm = m[2]; c = __popc(m[2]);
            if (k > c) { k -= c; di = 3; mm = m[3]; } } }
    return di * 4 + (select_bit(mm, k - 1) >> 3) + 1;
}

// Position just past the k-th LEB128 terminator at or after q in a lane
// stream (`base`, `end` bytes).  `bcnt` holds the terminator count of every
// 16-byte block of the LDS data area; the lane's blocks start at `blk0`.
// Whole blocks are skipped 16 counts per LDS read.  kBad if the stream ends
// first.
MDZ_DEV uint32_t skip_varints(const uint8_t* base, const uint8_t* bcnt, uint32_t blk0, uint32_t q,
                              uint32_t k, uint32_t end) {
    if (k == 0) return q;
    if (q >= end) return kBad;
    const u32x4* blk = reinterpret_cast<const u32x4*>(base);
    const uint32_t last = (end - 1) >> 4;
    uint32_t b = q >> 4;
    uint32_t m[4];
    block_terms(blk[b], q & 15, m);
    const uint32_t c = __popc(m[0]) + __popc(m[1]) + __popc(m[2]) + __popc(m[3]);
    if (c < k) {
        k -= c;
        uint32_t gb = blk0 + b + 1;                 // global block index
        const uint32_t glast = blk0 + last;
        for (;;) {                                  // 16 block counts per LDS read
            if (gb > glast) return kBad;
            const uint32_t a = gb & ~15u;
            const u32x4 cn = *reinterpret_cast<const u32x4*>(bcnt + a);
            const uint32_t w[4] = {cn.x, cn.y, cn.z, cn.w};
            bool found = false;
            for (uint32_t i = gb - a; i < 16; ++i) {
                if (a + i > glast) return kBad;
                const uint32_t ci = (w[i >> 2] >> (8 * (i & 3))) & 0xffu;
                if (ci >= k) { gb = a + i; found = true; break; }
                k -= ci;
            }
            if (found) break;
            gb = a + 16;
        }
        b = gb - blk0;
        block_terms(blk[b], 0, m);
    }
    const uint32_t pos = b * 16 + block_select(m, k);
    return pos <= end ? pos : kBad;
}

// 32-bit LEB128 for T of up to 4 bytes: bytes 0..3 in lo, byte 4 in b4.
// Same results and status codes as varint_decode.
template <int EB>
MDZ_DEV uint32_t varint32(uint32_t lo, uint32_t b4, uint32_t avail, uint32_t& val, int32_t& err) {
    constexpr uint32_t kMax = VarintSpec<EB>::kMax;
    constexpr uint32_t tmask = kMax >= 4 ? 0x80808080u : kMax == 3 ? 0x808080u : 0x8080u;
    const uint32_t t = ~lo & tmask;
    uint32_t len = t ? (__builtin_ctz(t) >> 3) + 1 : (kMax >= 5 && !(b4 & 0x80) ? 5 : 0);
    if (len == 0 || len > avail) {
        err = (len == 0 && avail >= kMax) ? MDZ_ST_BAD_SYMBOL : MDZ_ST_INPUT_OVERRUN;
        return 0;
    }
    uint32_t v = (lo & 0x7fu) | ((lo >> 1) & 0x3f80u) | ((lo >> 2) & 0x1fc000u) |
                 ((lo >> 3) & 0xfe00000u);
    if (len < 4) v = __builtin_amdgcn_ubfe(v, 0, 7 * len);
    else if (len == 5) v |= b4 << 28;
    val = v;
    return len;
}

// ------------------------------------------------ windowed single wave ----
// Fallback for chunks whose payload exceeds the LDS budget: wave 0 stages
// the raw interleaved payload through an LDS ring; lanes 0..31 decode their
// streams from it, reading their words through the lane-address table; when
// every running lane waits on data beyond the window, the window slides to
// the lowest word still needed.  Returns the chunk status.
template <int EB, int SW>
MDZ_DEV int32_t rlev1_windowed(const u32x4* __restrict__ src, LaneTable& T, uint32_t cl,
                               uint32_t* ring, uint32_t ring_words, uint8_t* chunk_out,
                               uint64_t limit, uint32_t nvals, uint32_t line) {
    constexpr int NB = EB == 8 ? 16 : 8;
    constexpr uint32_t kMaxV = VarintSpec<EB>::kMax;
    constexpr uint32_t kRefillAt = NB - 4;
    const uint32_t l = lane_id(), t = l & 31;
    const bool decoder = l < 32;
    const uint32_t W = (cl + 3) / 4;
    const uint32_t chunk_words = (T.total + 3) & ~3u;

    LaneCursor cur;
    cur.init(T, t, decoder ? W : 0);
    ByteBuf<NB> bb;
    LineEmitter<EB, SW> em;
    em.init(chunk_out, limit, t, line, 0);
    uint32_t mode = decoder ? kHdr : kDone;
    uint32_t rem = 0, vo = 0;
    uint64_t rv = 0;
    int64_t rd = 0;
    int32_t err = MDZ_ST_OK;
    const uint32_t last_bytes = cl - 4 * (W ? W - 1 : 0);

    uint32_t hi = min(chunk_words, ring_words);
    ring_stage(ring, ring_words, src, 0, hi);
    wave_sync();
    const uint32_t rmask = ring_words - 1;
    for (;;) {
        while (mode != kDone) {
            while (bb.nb <= kRefillAt && cur.r < cur.W) {
                uint32_t a = cur.addr();
                if (a >= hi) break;
                bb.append(ring[a & rmask], cur.r + 1 == cur.W ? last_bytes : 4);
                cur.advance(T, t);
            }
            const uint32_t need = mode == kRun ? 0 : (mode == kLit || mode == kBase) ? kMaxV : 1;
            if (bb.nb < need && cur.r < cur.W) break;   // waits on the window
            if (mode == kRun) {
                em.push(rv);
                rv += (uint64_t)rd;
                ++vo;
                if (--rem == 0) mode = kHdr;
            } else if (mode == kLit || mode == kBase) {
                uint64_t v;
                uint32_t n = varint_decode<EB>(bb.lo(), bb.hi(), bb.nb, v, err);
                if (n == 0) { mode = kDone; break; }
                bb.consume(n);
                if (mode == kLit) {
                    em.push(v);
                    ++vo;
                    if (--rem == 0) mode = kHdr;
                } else {
                    rv = v;
                    mode = kRun;
                }
            } else if (mode == kHdr) {
                if (bb.nb == 0) { err = MDZ_ST_INPUT_OVERRUN; mode = kDone; break; }
                uint32_t h = (uint32_t)(bb.lo() & 0xff);
                bb.consume(1);
                if (h & 0x80) { rem = 256 - h; mode = kLit; }
                else { rem = h + 3; mode = kDelta; }
                if (rem > nvals - vo) { err = MDZ_ST_OUTPUT_OVERRUN; mode = kDone; break; }
            } else {   // kDelta
                if (bb.nb == 0) { err = MDZ_ST_INPUT_OVERRUN; mode = kDone; break; }
                rd = (int64_t)(int8_t)(bb.lo() & 0xff);
                bb.consume(1);
                mode = kBase;
            }
            if (vo == nvals) mode = kDone;
        }
        uint32_t needw = (mode != kDone && cur.r < cur.W) ? cur.addr() : 0xffffffffu;
        uint32_t lo = wave_min_u32(needw);
        if (lo == 0xffffffffu) break;
        lo &= ~3u;
        uint32_t nhi = min(chunk_words, lo + ring_words);
        wave_sync();
        if (nhi > hi) ring_stage(ring, ring_words, src, max(hi, lo), nhi);
        hi = max(hi, nhi);
        wave_sync();
    }
    uint64_t bad = __ballot(err != MDZ_ST_OK);
    if (!bad) return MDZ_ST_OK;
    return __shfl(err, (int)__builtin_ctzll(bad), kWave);
}

// ------------------------------------------------------------ the kernel --
// Scan: records the decoder state at values K, 2K, ... of lane t's stream.
template <int EB>
MDZ_DEV void scan_lane(SplitTables& tb, const uint8_t* base, const uint8_t* bcnt, uint32_t blk0,
                       uint32_t end, uint32_t t, uint32_t nvals, uint32_t K) {
    uint32_t p = 0, v = 0, jn = 1, target = K;
    int32_t e = MDZ_ST_OK;
    auto record = [&](uint32_t pos, uint32_t mode, uint32_t rem, uint64_t rv, int32_t d) {
        tb.ck.p[jn - 1][t] = pos;
        tb.ck.m[jn - 1][t] = mode | (rem << 8) | (((uint32_t)d & 0xffu) << 24);
        tb.ck.vlo[jn - 1][t] = (uint32_t)rv;
        tb.ck.vhi[jn - 1][t] = (uint32_t)(rv >> 32);
        ++jn;
        target += K;
    };
    while (jn < kSegs) {
        if (v == target) { record(p, kHdr, 0, 0, 0); continue; }
        if (p >= end) break;
        const uint64_t x = peek8(base, p);
        const uint32_t h = (uint32_t)(x & 0xff);
        if (h & 0x80) {
            const uint32_t n = 256 - h;
            if (n > nvals - v) break;
            uint32_t q = p + 1, done = 0;
            while (jn < kSegs && target < v + n) {
                const uint32_t k = target - v - done;
                q = skip_varints(base, bcnt, blk0, q, k, end);
                if (q == kBad) break;
                done += k;
                record(q, kLit, n - done, 0, 0);
            }
            if (q == kBad) break;
            q = skip_varints(base, bcnt, blk0, q, n - done, end);
            if (q == kBad) break;
            p = q;
            v += n;
        } else {
            const uint32_t cnt = h + 3;
            if (cnt > nvals - v || p + 1 >= end) break;
            const int32_t d = (int32_t)(int8_t)(x >> 8);
            uint64_t bv;
            const uint32_t len = read_varint<EB>(base, p + 2, end, bv, e);
            if (len == 0) break;
            const uint32_t pe = p + 2 + len;
            while (jn < kSegs && target < v + cnt) {
                const uint32_t i = target - v;
                record(pe, kRun, cnt - i, bv + (uint64_t)(int64_t)d * i, d);
            }
            p = pe;
            v += cnt;
        }
    }
    while (jn < kSegs) record(0, kDone, 0, 0, 0);    // past a damaged token
}

// Decode values [v0, v0+K) of one lane for T of up to 4 bytes: 32-bit
// arithmetic, the next token's header folded into the step of its first value.
template <int EB, int SW>
MDZ_DEV int32_t decode_seg32(const uint8_t* base, uint32_t end, uint32_t p, uint32_t mode,
                             uint32_t rem, uint32_t rv, uint32_t d, uint32_t v0, uint32_t K,
                             uint32_t nvals, LineEmitter<EB, SW>& em) {
    for (uint32_t i = 0; i < K; ++i) {
        if (rem == 0) {
            if (p >= end) return MDZ_ST_INPUT_OVERRUN;
            const uint64_t x = peek8(base, p);
            const uint32_t lo = (uint32_t)x, h = lo & 0xffu;
            rem = (h & 0x80) ? 256 - h : h + 3;
            if (rem > nvals - v0 - i) return MDZ_ST_OUTPUT_OVERRUN;
            if (h & 0x80) {
                mode = kLit;
                p += 1;
            } else {
                if (p + 1 >= end) return MDZ_ST_INPUT_OVERRUN;
                d = (uint32_t)(int32_t)(int8_t)(lo >> 8);
                int32_t e = MDZ_ST_OK;
                const uint32_t hi = (uint32_t)(x >> 32);
                const uint32_t len = varint32<EB>(alignbyte(hi, lo, 2), (hi >> 16) & 0xffu,
                                                  end - p - 2, rv, e);
                if (len == 0) return e;
                p += 2 + len;
                mode = kRun;
            }
        }
        uint32_t val;
        if (mode == kLit) {
            const uint64_t x = peek8(base, p);
            int32_t e = MDZ_ST_OK;
            const uint32_t len = varint32<EB>((uint32_t)x, (uint32_t)(x >> 32) & 0xffu,
                                              end - p, val, e);
            if (len == 0) return e;
            p += len;
        } else {
            val = rv;
            rv += d;
        }
        em.push(val);
        --rem;
    }
    return MDZ_ST_OK;
}

// Generic decode (any T, 64-bit arithmetic).
template <int EB, int SW>
MDZ_DEV int32_t decode_seg64(const uint8_t* base, uint32_t end, uint32_t p, uint32_t mode,
                             uint32_t rem, uint64_t rv, int64_t d, uint32_t v0, uint32_t K,
                             uint32_t nvals, LineEmitter<EB, SW>& em) {
    int32_t e = MDZ_ST_OK;
    for (uint32_t i = 0; i < K;) {
        if (mode == kHdr) {
            if (p >= end) return MDZ_ST_INPUT_OVERRUN;
            const uint64_t x = peek8(base, p);
            const uint32_t h = (uint32_t)(x & 0xff);
            if (h & 0x80) {
                rem = 256 - h;
                if (rem > nvals - v0 - i) return MDZ_ST_OUTPUT_OVERRUN;
                mode = kLit;
                p += 1;
            } else {
                rem = h + 3;
                if (rem > nvals - v0 - i) return MDZ_ST_OUTPUT_OVERRUN;
                if (p + 1 >= end) return MDZ_ST_INPUT_OVERRUN;
                d = (int64_t)(int8_t)(x >> 8);
                const uint32_t len = read_varint<EB>(base, p + 2, end, rv, e);
                if (len == 0) return e;
                p += 2 + len;
                mode = kRun;
            }
            continue;
        }
        uint64_t val;
        if (mode == kRun) {
            val = rv;
            rv += (uint64_t)d;
        } else {
            const uint32_t len = read_varint<EB>(base, p, end, val, e);
            if (len == 0) return e;
            p += len;
        }
        em.push(val);
        ++i;
        if (--rem == 0) mode = kHdr;
    }
    return MDZ_ST_OK;
}

template <int EB, int SW>
__global__ __launch_bounds__(kThreads) void rlev1_decode_kernel(
    const uint8_t* __restrict__ payload, const uint64_t* __restrict__ blk_off,
    const uint64_t* __restrict__ col_len, uint64_t n_chunks, uint64_t chunk_bytes,
    uint32_t line, uint8_t* __restrict__ dst, uint64_t dst_bytes, int32_t* __restrict__ status,
    uint32_t data_cap) {
    extern __shared__ __align__(16) uint32_t smem[];
    uint8_t* const data = reinterpret_cast<uint8_t*>(smem);
    uint8_t* const bcnt = data + data_cap + 16;                 // data_cap / 16 counts
    SplitTables& tb = *reinterpret_cast<SplitTables*>(bcnt + data_cap / 16);

    const uint64_t c = blockIdx.x;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const uint32_t nvals = (uint32_t)(chunk_bytes / 32 / EB);
    const uint32_t K = nvals / kSegs;
    const uint64_t out_base = c * chunk_bytes;
    const uint64_t limit = dst_bytes > out_base ? min(dst_bytes - out_base, chunk_bytes) : 0;
    const uint64_t c_off = blk_off[c], c_len = blk_off[c + 1] - c_off;
    const u32x4* src = reinterpret_cast<const u32x4*>(payload + c_off);

    // ---- lane table and lane regions (wave 0)
    if (wave == 0) {
        const uint32_t t = l & 31;
        const uint32_t cl = (uint32_t)col_len[c * 32 + t];
        lane_table_build(tb.lt, (cl + 3) / 4);
        const uint32_t r16 = (cl + 15) & ~15u;
        uint32_t incl = l < 32 ? r16 : 0;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            uint32_t y = __shfl_up(incl, o, kWave);
            if ((l & 31) >= (uint32_t)o) incl += y;
        }
        if (l < 32) { tb.lbase[t] = incl - r16; tb.cl[t] = cl; }
        if (l == 31) { tb.fits = incl <= data_cap; tb.nblk = incl / 16; }
        if (l == 0) {
            tb.err = 0xffffffffu;
            tb.extent_ok = (uint64_t)tb.lt.total * 4 <= c_len && !(c_off & 15);
        }
    }
    __syncthreads();

    if (!tb.extent_ok) {
        if (tid == 0 && status) status[c] = MDZ_ST_INPUT_OVERRUN;
        return;
    }
    if (!tb.fits) {
        if (wave == 0) {
            uint32_t ring_words = 256;
            while (ring_words * 2 <= data_cap / 4) ring_words *= 2;
            int32_t st = rlev1_windowed<EB, SW>(src, tb.lt, tb.cl[l & 31], smem, ring_words,
                                                dst + out_base, limit, nvals, line);
            if (l == 0 && status) status[c] = st;
        }
        return;
    }

    // ---- 1. stage + de-interleave into per-lane contiguous LDS regions
    {
        const uint32_t total = tb.lt.total, ngroups = (total + 3) / 4;
        uint32_t* d32 = smem;
        for (uint32_t g0 = tid; g0 < ngroups; g0 += 4 * kThreads) {
            u32x4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t g = g0 + u * kThreads;
                if (g < ngroups) v[u] = __builtin_nontemporal_load(src + g);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t g = g0 + u * kThreads;
                if (g >= ngroups) break;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t w = 4 * g + i;
                    if (w < total) {
                        uint32_t lane, round;
                        lane_of_word(tb.lt, w, lane, round);
                        d32[(tb.lbase[lane] >> 2) + round] = v[u][i];
                    }
                }
            }
        }
    }
    __syncthreads();
    // ---- per-16-byte-block LEB128 terminator counts (bytes past a lane's
    //      end count too; the scan bounds every position by the lane length)
    for (uint32_t b = tid; b < tb.nblk; b += kThreads) {
        const u32x4 x = reinterpret_cast<const u32x4*>(data)[b];
        bcnt[b] = (uint8_t)(__popc(~x.x & 0x80808080u) + __popc(~x.y & 0x80808080u) +
                            __popc(~x.z & 0x80808080u) + __popc(~x.w & 0x80808080u));
    }
    __syncthreads();

    // ---- 2. scan: one thread per lane stream, 4 per wave
    if (l < 32 / (kThreads / 64)) {
        const uint32_t t = l * (kThreads / 64) + wave;
        scan_lane<EB>(tb, data + tb.lbase[t], bcnt, tb.lbase[t] / 16, tb.cl[t], t, nvals, K);
    }
    __syncthreads();

    // ---- 3. decode: thread (lane t, segment j) writes values [jK, (j+1)K)
    {
        const uint32_t t = tid & 31, j = tid >> 5;
        const uint8_t* base = data + tb.lbase[t];
        const uint32_t end = tb.cl[t];
        uint32_t p = 0, mode = kHdr, rem = 0, vlo = 0, vhi = 0, dd = 0;
        if (j) {
            p = tb.ck.p[j - 1][t];
            const uint32_t m = tb.ck.m[j - 1][t];
            mode = m & 0xff;
            rem = (m >> 8) & 0xffff;
            dd = (uint32_t)(int32_t)(int8_t)(m >> 24);
            vlo = tb.ck.vlo[j - 1][t];
            vhi = tb.ck.vhi[j - 1][t];
        }
        int32_t e = MDZ_ST_OK;
        if (mode != kDone) {
            const uint32_t v0 = j * K;
            LineEmitter<EB, SW> em;
            em.init(dst + out_base, limit, t, line, (uint64_t)v0 * EB);
            if constexpr (EB <= 4) {
                e = decode_seg32<EB, SW>(base, end, p, mode, rem, vlo, dd, v0, K, nvals, em);
            } else {
                e = decode_seg64<EB, SW>(base, end, p, mode, rem,
                                         (uint64_t)vlo | ((uint64_t)vhi << 32),
                                         (int64_t)(int32_t)dd, v0, K, nvals, em);
            }
        }
        if (e != MDZ_ST_OK) atomicMin(&tb.err, pack_err(t, j, e));
    }
    __syncthreads();
    if (tid == 0 && status) {
        const uint32_t e = tb.err;
        status[c] = e == 0xffffffffu ? MDZ_ST_OK : (int32_t)(int8_t)(e & 0xff);
    }
}

template <int EB, int SW>
hipError_t launch(const mdz_lanes_t& in, void* dst, uint64_t dst_bytes, int32_t* status,
                  uint32_t data_cap, hipStream_t s) {
    const size_t smem = (size_t)data_cap + 16 + data_cap / 16 + sizeof(SplitTables);
    auto k = rlev1_decode_kernel<EB, SW>;
    hipError_t e = set_dynamic_lds((const void*)k, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((uint32_t)in.n_chunks), dim3(kThreads), smem, s,
                       (const uint8_t*)in.payload, in.blk_off, in.col_len, in.n_chunks,
                       in.chunk_bytes, in.lane_line, (uint8_t*)dst, dst_bytes, status, data_cap);
    return hipGetLastError();
}

template <int EB>
hipError_t launch_eb(const mdz_lanes_t& in, void* dst, uint64_t dst_bytes, int32_t* status,
                     uint32_t data_cap, uint32_t sw, hipStream_t s) {
    if (sw == 16) return launch<EB, 16>(in, dst, dst_bytes, status, data_cap, s);
    if (sw == 8) return launch<EB, 8>(in, dst, dst_bytes, status, data_cap, s);
    if constexpr (EB <= 4)
        if (sw == 4) return launch<EB, 4>(in, dst, dst_bytes, status, data_cap, s);
    if constexpr (EB <= 2)
        if (sw == 2) return launch<EB, 2>(in, dst, dst_bytes, status, data_cap, s);
    if constexpr (EB == 1)
        if (sw == 1) return launch<EB, 1>(in, dst, dst_bytes, status, data_cap, s);
    return hipErrorInvalidValue;
}

}  // namespace

int rlev1_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                     const mdz_lanes_tune_t* tune, hipStream_t s) {
    if (!in || !in->payload || !in->blk_off || !in->col_len || (!dst && dst_bytes))
        return MDZ_ERR_INVALID;
    const uint32_t eb = in->elem_bytes, line = in->lane_line;
    const uint64_t chunk = in->chunk_bytes;
    // each lane stream holds a multiple of 8 values (8 decode segments)
    if (!(eb == 1 || eb == 2 || eb == 4 || eb == 8) || line == 0 || line % eb || chunk % 32 ||
        (chunk / 32) % line || (chunk / 32 / eb) % kSegs || chunk / 32 / eb < kSegs ||
        chunk / 32 / eb > 0xffffffu)
        return MDZ_ERR_INVALID;
    if (in->n_chunks == 0) return MDZ_OK;
    if (in->n_chunks > 0x7fffffffull) return MDZ_ERR_UNSUPPORTED;
    if (((uintptr_t)in->payload & 15) != 0) return MDZ_ERR_INVALID;
    // LDS data area: the caller's bound on one chunk's payload (+ 12 bytes of
    // 16-byte alignment per lane region), else as much as a workgroup gets.
    const uint32_t lds_max = ((160 * 1024 - 16 - (uint32_t)sizeof(SplitTables) - 512) * 16 / 17);
    uint32_t cap = lds_max & ~255u;
    if (tune && tune->max_chunk_bytes) {
        uint64_t want = ((uint64_t)tune->max_chunk_bytes + 32 * 12 + 255) & ~255ull;
        if (want < cap) cap = (uint32_t)want;
    }
    if (tune && tune->ring_words) {          // forces the windowed path for testing
        if (tune->ring_words < 256 || (tune->ring_words & (tune->ring_words - 1)) ||
            tune->ring_words * 4ull > lds_max)
            return MDZ_ERR_INVALID;
        cap = tune->ring_words * 4;
    }
    // widest store that divides the line, the segment and the output alignment
    const uint64_t seg_bytes = chunk / 32 / kSegs;
    uint32_t sw = 16;
    while (sw > eb && ((line % sw) || ((uintptr_t)dst % sw) || (seg_bytes % sw))) sw >>= 1;
    hipError_t e;
    switch (eb) {
    case 1: e = launch_eb<1>(*in, dst, dst_bytes, status, cap, sw, s); break;
    case 2: e = launch_eb<2>(*in, dst, dst_bytes, status, cap, sw, s); break;
    case 4: e = launch_eb<4>(*in, dst, dst_bytes, status, cap, sw, s); break;
    default: e = launch_eb<8>(*in, dst, dst_bytes, status, cap, sw, s); break;
    }
    return record_hip(e, "rlev1_decode_kernel launch");
}

}  // namespace mdz
