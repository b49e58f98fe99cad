// RLE v2 device code shared by the per-chunk kernel and the batched
// persistent kernel.
#pragma once
// RLE v2 (ORC integer RLEv2, lane-transposed) decoder for gfx950.
//
// Replaces the reference's reader-warp / decoder-warp kernel
// (include/rlev2/decoder_new.h:618-657, decoder_warp :354-611) and host entry
// rlev2::decompress_gpu (include/rlev2/decoder_trans.h:615-681).
//
// One workgroup (8 wave64s) per chunk:
//   1. the chunk's interleaved payload is read with 16-byte loads and each
//      4-byte word scattered to its lane's contiguous LDS region (a per-
//      segment destination table and a stepped quotient place it);
//   2. header walk: 32 threads of wave 0, one per lane stream, walk the run
//      headers in parallel (one LDS read per header) and append one record
//      per run to an LDS pool, publishing the count as they go;
//   3. the other waves -- and wave 0 once its walk ends -- claim windows of
//      published records, check them one per lane with the oracle's checks
//      in its order (a lane's first failing run is its status), and decode
//      each run with the whole wave, one value per SIMD lane:
//        SHORT_REPEAT  lane i writes the repeated value;
//        DIRECT        lane i unpacks the i-th MSB-first bit field;
//        DELTA         fixed stride: base + i*delta; packed deltas: lane i
//                      unpacks delta i and a wave prefix sum makes values;
//        PATCHED_BASE  lane k reads patch entry k, a prefix sum of the gaps
//                      places it, and every lane ORs in the patches that land
//                      on its value (64-bit mask: decoder_new.h:533's 32-bit
//                      mask drops patch bits at positions >= 32).
//   Steps 2 and 3 overlap; they repeat if a chunk has more than kRunPool
//   runs.  Lane t's q-th value is element (q / RU) * 32 * RU + t * RU + q % RU
//   of the chunk, so consecutive SIMD lanes store consecutive elements.
//   Runs are checked and written concurrently, so for a lane that fails,
//   runs after its first error may still be written (its status is exact,
//   and every byte before the error point equals the oracle's).
#include <type_traits>

#include "lanes_dev.h"
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

constexpr uint32_t kThreads = 512;

__constant__ uint8_t kDecW[32] = {1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16,
                                  17, 18, 19, 20, 21, 22, 23, 24, 26, 28, 30, 32, 40, 48, 56, 64};

MDZ_DEV uint32_t closest_bits(uint32_t b) {     // utils.h __CLOSEST_FIXED_BIT_MAP
    if (b <= 24) return b ? b : 1;
    if (b <= 32) return (b + 1) & ~1u;
    return (b + 7) & ~7u;
}

MDZ_DEV uint64_t gather7(uint64_t x) {
    return (x & 0x7full) | ((x >> 1) & (0x7full << 7)) | ((x >> 2) & (0x7full << 14)) |
           ((x >> 3) & (0x7full << 21)) | ((x >> 4) & (0x7full << 28)) |
           ((x >> 5) & (0x7full << 35)) | ((x >> 6) & (0x7full << 42)) |
           ((x >> 7) & (0x7full << 49));
}

// Unsigned LEB128 of at most 10 bytes at byte p (uniform across the wave).
// Returns its length, or 0 with *err: input overrun if the lane stream ends
// inside it, malformed symbol if it is longer than 10 bytes.
MDZ_DEV uint32_t uvarint(const uint8_t* base, uint32_t p, uint32_t end, uint64_t& val, int32_t& err) {
    const uint64_t lo = lds_peek8(base, p), hi = lds_peek8(base, p + 8);
    const uint64_t tl = ~lo & 0x8080808080808080ull, th = ~hi & 0x8080ull;
    const uint32_t len = tl ? (__builtin_ctzll(tl) >> 3) + 1 : th ? 9 + (__builtin_ctzll(th) >> 3) : 0;
    const uint32_t avail = end > p ? end - p : 0;
    if (len == 0 || len > avail) {
        err = (len == 0 && avail >= 10) ? MDZ_ST_BAD_SYMBOL : MDZ_ST_INPUT_OVERRUN;
        return 0;
    }
    uint64_t v = gather7(lo);
    if (len < 8) {
        v &= (1ull << (7 * len)) - 1;
    } else if (len > 8) {
        uint64_t h = (hi & 0x7f) | ((hi >> 1) & (0x7full << 7));
        h &= (1ull << (7 * (len - 8))) - 1;
        v |= h << 56;
    }
    val = v;
    return len;
}

// Where lane t's q-th value goes: element (q/RU)*32*RU + t*RU + q%RU.
// READ_UNIT is a power of two in every reference configuration; that case
// is shifts and masks, the general one divides.
struct LanePlace {
    uint32_t t, ru, sh;      // sh = log2(ru) when ru is a power of two, else 32
    MDZ_DEV uint64_t elem(uint32_t q) const {
        if (sh < 32) return ((uint64_t)(q >> sh) * 32 + t) << sh | (q & (ru - 1));
        const uint32_t line = q / ru;
        return ((uint64_t)line * 32 + t) * ru + (q - line * ru);
    }
};

template <int EB>
MDZ_DEV void put_val(uint8_t* out, uint64_t limit, uint32_t q, const LanePlace& lp, uint64_t v) {
    const uint64_t off = lp.elem(q) * EB;
    if (off + EB > limit) {                 // clipped output: the bytes that fit
        for (uint64_t b = off; b < limit; ++b) out[b] = (uint8_t)(v >> (8 * (b - off)));
        return;
    }
    if constexpr (EB == 8) *reinterpret_cast<uint64_t*>(out + off) = v;
    else if constexpr (EB == 4) *reinterpret_cast<uint32_t*>(out + off) = (uint32_t)v;
    else if constexpr (EB == 2) *reinterpret_cast<uint16_t*>(out + off) = (uint16_t)v;
    else out[off] = (uint8_t)v;
}

constexpr uint32_t kWinDefault = 16;            // records a consumer wave claims at once
// Consumer windows a launch may use: they divide both record pools.
inline bool rlev2_window_ok(uint32_t w) { return w == 8 || w == 16 || w == 32 || w == 64; }
constexpr uint32_t kOpen = 0xffffffffu;         // final_ while the walk runs
constexpr uint32_t kSpinMax = 1u << 22;         // consumer wait bound (about 0.2 s)

// Bit width of a 5-bit width code (the reference's FixedBitSizes decode):
// 1..24, then 26, 28, 30, 32, then 40, 48, 56, 64.
MDZ_DEV uint32_t decode_width(uint32_t c) {
    // masks instead of compares: c >= 24 iff c + 8 >= 32, c >= 28 iff c + 4 >= 32
    const uint32_t ge24 = 0u - ((c + 8) >> 5), ge28 = 0u - ((c + 4) >> 5);
    return c + 1 + ((c - 23) & ge24) + ((6 * c - 162) & ge28);
}

// closest_bits() as mask arithmetic (the walk keeps it free of branches).
MDZ_DEV uint32_t closest_bits_sel(uint32_t b) {
    const uint32_t gt24 = 0u - (uint32_t)(b > 24), gt32 = 0u - (uint32_t)(b > 32);
    return max(b, 1u) + ((b & 1) & gt24 & ~gt32) + (((8 - (b & 7)) & 7) & gt32);
}

// Byte length (1..10) of the unsigned LEB128 at byte p, or 0 with *err as
// uvarint() reports it.
MDZ_DEV uint32_t uvarint_len(const uint8_t* base, uint32_t p, uint32_t end, int32_t& err) {
    const uint64_t lo = lds_peek8(base, p), hi = lds_peek8(base, p + 8);
    const uint64_t tl = ~lo & 0x8080808080808080ull, th = ~hi & 0x8080ull;
    const uint32_t len = tl ? (__builtin_ctzll(tl) >> 3) + 1 : th ? 9 + (__builtin_ctzll(th) >> 3) : 0;
    const uint32_t avail = end > p ? end - p : 0;
    if (len == 0 || len > avail) {
        err = (len == 0 && avail >= 10) ? MDZ_ST_BAD_SYMBOL : MDZ_ST_INPUT_OVERRUN;
        return 0;
    }
    return len;
}

// One run, as the header walk leaves it for the value pass.
struct Run {
    uint32_t p;      // lane byte offset: SHORT_REPEAT value bytes, DIRECT values,
                     // DELTA base varint, PATCHED_BASE base bytes
    uint32_t q;      // index of the run's first value in its lane
    uint32_t meta;   // kind | code << 2 | w << 7 | (len - 1) << 14 | lane << 23
    uint32_t aux;    // SHORT_REPEAT: nb; DELTA: n1 | n2 << 4;
                     // PATCHED_BASE: bw | pw << 4 | pgw << 11 | pll << 15 | cfb << 20
};

// Walks one run header at byte *p of a lane stream of `end` bytes whose next
// value is q: checks it as the oracle's mdz_ref_rlev2_lane_decode does (in
// the same order), fills r and advances *p.  Patch positions of a
// PATCHED_BASE run are checked later, by patch_list_check().
MDZ_DEV int32_t walk_run(const uint8_t* base, uint32_t end, uint32_t& p, uint32_t q, uint32_t nvals,
                         uint32_t lane, Run& r) {
    if (p >= end) return MDZ_ST_INPUT_OVERRUN;
    const uint64_t x = lds_peek8(base, p);
    const uint32_t first = (uint32_t)(x & 0xff), kind = first >> 6;
    r.q = q;
    if (kind == 0) {                                            // SHORT_REPEAT
        const uint32_t nb = ((first >> 3) & 7) + 1, cnt = (first & 7) + 3;
        if (p + 1 + nb > end) return MDZ_ST_INPUT_OVERRUN;
        if (cnt > nvals - q) return MDZ_ST_OUTPUT_OVERRUN;
        r.p = p + 1;
        r.meta = (cnt - 1) << 14 | lane << 23;
        r.aux = nb;
        p += 1 + nb;
        return MDZ_ST_OK;
    }
    if (p + 1 >= end) return MDZ_ST_INPUT_OVERRUN;
    const uint32_t code = (first >> 1) & 0x1f, w = decode_width(code);
    const uint32_t len = ((((first & 1) << 8)) | (uint32_t)((x >> 8) & 0xff)) + 1;
    r.meta = kind | code << 2 | w << 7 | (len - 1) << 14 | lane << 23;
    if (len > nvals - q) {
        // DIRECT, DELTA and PATCHED_BASE all check the value count first,
        // except that PATCHED_BASE checks its 4-byte header before it
        if (kind != 2 || p + 3 < end) return MDZ_ST_OUTPUT_OVERRUN;
    }
    if (kind == 1) {                                            // DIRECT
        const uint32_t bytes = (len * w + 7) / 8;
        if (p + 2 + bytes > end) return MDZ_ST_INPUT_OVERRUN;
        r.p = p + 2;
        r.aux = 0;
        p += 2 + bytes;
        return MDZ_ST_OK;
    }
    if (kind == 3) {                                            // DELTA
        if (len < 2) return MDZ_ST_BAD_SYMBOL;
        int32_t err = MDZ_ST_OK;
        const uint32_t n1 = uvarint_len(base, p + 2, end, err);
        if (!n1) return err;
        const uint32_t n2 = uvarint_len(base, p + 2 + n1, end, err);
        if (!n2) return err;
        const uint32_t pd = p + 2 + n1 + n2;
        r.p = p + 2;
        r.aux = n1 | n2 << 4;
        if (code == 0) {
            p = pd;
        } else {
            const uint32_t bytes = ((len - 2) * w + 7) / 8;
            if (pd + bytes > end) return MDZ_ST_INPUT_OVERRUN;
            p = pd + bytes;
        }
        return MDZ_ST_OK;
    }
    // PATCHED_BASE
    if (p + 3 >= end) return MDZ_ST_INPUT_OVERRUN;
    const uint32_t third = (uint32_t)((x >> 16) & 0xff), fourth = (uint32_t)((x >> 24) & 0xff);
    const uint32_t bw = ((third >> 5) & 7) + 1, pw = decode_width(third & 0x1f);
    const uint32_t pgw = ((fourth >> 5) & 7) + 1, pll = fourth & 0x1f;
    if (p + 4 + bw > end) return MDZ_ST_INPUT_OVERRUN;
    const uint32_t pb = p + 4 + bw, vbytes = (len * w + 7) / 8;
    if (pb + vbytes > end) return MDZ_ST_INPUT_OVERRUN;
    const uint32_t cfb = closest_bits(pw + pgw);
    r.p = p + 4;
    r.aux = bw | pw << 4 | pgw << 11 | pll << 15 | cfb << 20;
    p = pb + vbytes + (pll * cfb + 7) / 8;
    return MDZ_ST_OK;
}

// Header chain step of the walk: the byte length and value count of the run
// at byte p, from the 24 bytes at p alone (one LDS read).  Exact for every
// run that passes walk_run(); for a malformed run it only has to stop:
// *stop is set when the chain cannot go on past this run.
MDZ_DEV void chain_step(const uint8_t* base, uint32_t p, uint32_t end, uint32_t& size, uint32_t& cnt,
                        bool& stop) {
    const uint32_t* w32 = reinterpret_cast<const uint32_t*>(base) + (p >> 2);
    const uint32_t k = p & 3;
    const uint32_t d0 = w32[0], d1 = w32[1], d2 = w32[2], d3 = w32[3], d4 = w32[4], d5 = w32[5],
                   d6 = w32[6];
    // all seven words in one LDS round trip: keep the compiler from sinking
    // the DELTA-only loads into that branch (a second round trip per header)
    asm volatile("" ::"v"(d0), "v"(d1), "v"(d2), "v"(d3), "v"(d4), "v"(d5), "v"(d6));
    const uint32_t u0 = alignbyte(d1, d0, k), u1 = alignbyte(d2, d1, k), u2 = alignbyte(d3, d2, k),
                   u3 = alignbyte(d4, d3, k), u4 = alignbyte(d5, d4, k), u5 = alignbyte(d6, d5, k);
    const uint32_t b0 = u0 & 0xff, b1 = (u0 >> 8) & 0xff, b2 = (u0 >> 16) & 0xff, b3 = u0 >> 24;
    const uint32_t kind = b0 >> 6;
    const uint32_t code = (b0 >> 1) & 31, w = decode_width(code);
    const uint32_t len = (((b0 & 1) << 8) | b1) + 1;
    const uint32_t vbytes = (len * w + 7) / 8;
    uint32_t sz, n = len;
    bool bad = false;
    if (kind == 0) {                                         // SHORT_REPEAT
        sz = 2 + ((b0 >> 3) & 7);
        n = (b0 & 7) + 3;
    } else if (kind == 1) {                                  // DIRECT
        sz = 2 + vbytes;
    } else if (kind == 2) {                                  // PATCHED_BASE
        const uint32_t bw = ((b2 >> 5) & 7) + 1, pw = decode_width(b2 & 31);
        const uint32_t pgw = ((b3 >> 5) & 7) + 1, pll = b3 & 31;
        sz = 4 + bw + vbytes + (pll * closest_bits_sel(pw + pgw) + 7) / 8;
    } else {                                                 // DELTA: two LEB128 varints
        // bytes 2..9 and 10..17 as 64-bit words: the first varint's end is
        // the first clear bit 7 among them, the second's the first after it
        const uint64_t x0 = (uint64_t)alignbyte(u2, u1, 2) << 32 | alignbyte(u1, u0, 2);
        const uint64_t x1 = (uint64_t)alignbyte(u4, u3, 2) << 32 | alignbyte(u3, u2, 2);
        const uint32_t x2 = alignbyte(u5, u4, 2);
        const uint64_t t1 = ~x0 & 0x8080808080808080ull;
        uint32_t n1;
        if (t1) {
            n1 = ((uint32_t)__builtin_ctzll(t1) >> 3) + 1;
        } else {
            const uint32_t tt = ~(uint32_t)x1 & 0x8080u;
            n1 = tt ? 9 + ((uint32_t)__builtin_ctz(tt) >> 3) : 0;
        }
        // the 8 bytes after the first varint (bytes 2 + n1 ..)
        const uint32_t sh1 = 8 * (n1 & 7);
        const uint64_t lo = n1 < 8 ? x0 : x1, hi = n1 < 8 ? x1 : (uint64_t)x2;
        const uint64_t y = sh1 ? (lo >> sh1) | (hi << (64 - sh1)) : lo;
        const uint64_t t2 = ~y & 0x8080808080808080ull;
        uint32_t n2;
        if (t2) {
            n2 = ((uint32_t)__builtin_ctzll(t2) >> 3) + 1;
        } else {                                             // a second varint of 9 or 10 bytes
            auto term4 = [](uint32_t u) {
                const uint32_t t = (~u >> 7) & 0x01010101u;
                return (t | (t >> 7) | (t >> 14) | (t >> 21)) & 0xfu;
            };
            const uint32_t tm = term4(u0) | term4(u1) << 4 | term4(u2) << 8 | term4(u3) << 12 |
                                term4(u4) << 16 | term4(u5) << 20;
            const uint32_t tt = (tm >> (2 + n1)) & 0x3ff;
            n2 = tt ? (uint32_t)__builtin_ctz(tt) + 1 : 0;
        }
        bad = n1 == 0 || n2 == 0 || len < 2;
        sz = 2 + n1 + n2 + (code ? ((len - 2) * w + 7) / 8 : 0);
    }
    size = sz;
    cnt = n;
    stop = bad || p >= end || p + sz > end;
}

// Patch-list check of a PATCHED_BASE run (one thread): the first entry
// that cannot be read, or that lands past the run, decides the status
// (readability only falls as the entry index grows).  Returns MDZ_ST_OK
// when every entry is good.
MDZ_DEV int32_t patch_list_check(const uint8_t* base, uint32_t end, const Run& r) {
    const uint32_t w = (r.meta >> 7) & 127, len = ((r.meta >> 14) & 511) + 1;
    const uint32_t bw = r.aux & 15, pw = (r.aux >> 4) & 127, pll = (r.aux >> 15) & 31,
                   cfb = (r.aux >> 20) & 127;
    const uint32_t pp = r.p + bw + (len * w + 7) / 8;
    uint32_t pos = 0;
    for (uint32_t k = 0; k < pll; ++k) {
        if ((uint64_t)pp * 8 + (uint64_t)(k + 1) * cfb > (uint64_t)end * 8) return MDZ_ST_INPUT_OVERRUN;
        const uint64_t e = lds_bits_be(base, pp, (uint64_t)k * cfb, cfb);
        pos += pw < 64 ? (uint32_t)(e >> pw) : 0u;
        if (pos >= len) return MDZ_ST_BAD_SYMBOL;
    }
    return MDZ_ST_OK;
}

// Unsigned LEB128 of n (1..10) bytes at byte p (per lane).
MDZ_DEV uint64_t uvarint_value(const uint8_t* base, uint32_t p, uint32_t n) {
    uint64_t v = gather7(lds_peek8(base, p));
    if (n < 8) {
        v &= (1ull << (7 * n)) - 1;
    } else if (n > 8) {
        const uint64_t hi = lds_peek8(base, p + 8);
        const uint64_t h = ((hi & 0x7f) | ((hi >> 1) & (0x7full << 7))) & ((1ull << (7 * (n - 8))) - 1);
        v |= h << 56;
    }
    return v;
}

// MSB-first field of w bits (1..64) starting `bit` bits after byte b0 of the
// 4-byte-aligned LDS array (32-bit bit offset; reads up to 12 bytes past the
// field start).
MDZ_DEV uint64_t lds_bits_be32(const uint8_t* base, uint32_t b0, uint32_t bit, uint32_t w) {
    const uint32_t b = b0 + (bit >> 3), s = bit & 7, k = b & 3;
    const uint32_t* d = reinterpret_cast<const uint32_t*>(base) + (b >> 2);
    const uint32_t d0 = d[0], d1 = d[1], d2 = d[2];
    const uint32_t lo = alignbyte(d1, d0, k), mid = alignbyte(d2, d1, k);
    const uint64_t be = ((uint64_t)__builtin_bswap32(lo) << 32) | __builtin_bswap32(mid);
    const uint64_t v = s ? (be << s) | (((d2 >> (8 * k)) & 0xffu) >> (8 - s)) : be;
    return w >= 64 ? v : v >> (64 - w);
}

// Decodes run r (all 64 lanes of one wave; r is wave-uniform, r.p an
// absolute LDS byte offset).  a and b are the run's decoded parameters:
// SHORT_REPEAT value; DELTA base and delta (masked); PATCHED_BASE base.  A run whose values lie in one READ_UNIT line
// and inside the output is stored through a per-run pointer; otherwise each
// value is placed and clipped on its own.
template <int EB>
MDZ_DEV void run_values(const Run& rr, uint64_t a, uint64_t b, const uint8_t* data, uint8_t* out,
                        uint64_t limit, uint32_t ru, uint32_t sh) {
    constexpr uint64_t vmask = EB == 8 ? ~0ull : ((1ull << (8 * EB)) - 1);
    typedef typename std::conditional<EB == 8, uint64_t,
            typename std::conditional<EB == 4, uint32_t,
            typename std::conditional<EB == 2, uint16_t, uint8_t>::type>::type>::type elem_t;
    const uint32_t l = lane_id();
    const uint32_t meta = __builtin_amdgcn_readfirstlane(rr.meta);
    const uint32_t aux = __builtin_amdgcn_readfirstlane(rr.aux);
    const uint32_t p = __builtin_amdgcn_readfirstlane(rr.p), q = __builtin_amdgcn_readfirstlane(rr.q);
    const uint32_t kind = meta & 3, code = (meta >> 2) & 31, w = (meta >> 7) & 127;
    const uint32_t len = ((meta >> 14) & 511) + 1, t = meta >> 23;
    const LanePlace lp{t, ru, sh};
    elem_t* o = nullptr;
    if (sh < 32) {
        const uint32_t in_line = q & (ru - 1);
        const uint64_t e0 = ((uint64_t)((q >> sh) * 32 + t) << sh) | in_line;
        if (in_line + len <= ru && (e0 + len) * EB <= limit) o = reinterpret_cast<elem_t*>(out) + e0;
    }
    auto put = [&](uint32_t i, uint64_t v) {
        if (o) o[i] = (elem_t)v;
        else put_val<EB>(out, limit, q + i, lp, v);
    };
    if (kind == 0) {                                            // SHORT_REPEAT (value in a)
        if (l < len) put(l, a);
    } else if (kind == 1) {                                     // DIRECT
        for (uint32_t i0 = 0; i0 < len; i0 += kWave) {
            const uint32_t i = i0 + l;
            if (i < len) put(i, lds_bits_be32(data, p, i * w, w) & vmask);
        }
    } else if (kind == 3) {                                     // DELTA
        const uint32_t n1 = aux & 15, n2 = aux >> 4;
        const uint64_t v0 = a, dt = b;                          // base and delta, masked
        if (code == 0) {
            for (uint32_t i0 = 0; i0 < len; i0 += kWave) {
                const uint32_t i = i0 + l;
                if (i < len) put(i, (v0 + (uint64_t)i * dt) & vmask);
            }
        } else {
            const uint32_t nd = len - 2, pd = p + n1 + n2;
            const uint64_t v1 = (v0 + dt) & vmask;
            if (l < 2) put(l, l ? v1 : v0);
            // multiplier = (DATA_TYPE)delta > 0 ? 1 : -1 on the unsigned type
            const bool neg = dt == 0;
            uint64_t carry = v1;
            for (uint32_t k0 = 0; k0 < nd; k0 += kWave) {
                const uint32_t k = k0 + l;
                uint64_t d = k < nd ? lds_bits_be32(data, pd, k * w, w) : 0;
                if (neg) d = 0 - d;
                const uint64_t incl = wave_incl_scan_u64(d);
                if (k < nd) put(2 + k, (carry + incl) & vmask);
                carry += last_lane_u64(incl);
            }
        }
    } else {                                                    // PATCHED_BASE
        const uint32_t bw = aux & 15, pw = (aux >> 4) & 127, pll = (aux >> 15) & 31,
                       cfb = (aux >> 20) & 127;
        const uint64_t bv = a;                                  // patched base
        const uint32_t pb = p + bw, pp = pb + (len * w + 7) / 8;
        const uint64_t e = l < pll ? lds_bits_be32(data, pp, l * cfb, cfb) : 0;
        const uint32_t gap = (l < pll && pw < 64) ? (uint32_t)(e >> pw) : 0;
        const uint32_t pos = wave_incl_scan_u32(gap);
        const uint64_t pv = pw >= 64 ? e : e & ((1ull << pw) - 1);
        for (uint32_t i0 = 0; i0 < len; i0 += kWave) {
            const uint32_t i = i0 + l;
            uint64_t v = i < len ? lds_bits_be32(data, pb, i * w, w) : 0;
            for (uint32_t k = 0; k < pll; ++k) {
                const uint32_t pk = __builtin_amdgcn_readlane(pos, k);
                const uint32_t vlo = (uint32_t)__builtin_amdgcn_readlane((uint32_t)pv, k);
                const uint32_t vhi = (uint32_t)__builtin_amdgcn_readlane((uint32_t)(pv >> 32), k);
                const uint64_t vk = (uint64_t)vhi << 32 | vlo;
                if (pk == i) v |= w >= 64 ? 0 : vk << w;
            }
            if (i < len) put(i, (v + bv) & vmask);
        }
    }
}


// Run pool per workgroup: a whole chunk (32 lanes) or one part of NL lanes.
template <uint32_t NL> struct PartPool { static constexpr uint32_t value = NL == 32 ? 768 : 384; };

template <uint32_t POOL>
struct Rlev2Tables {
    uint32_t lbase[32];      // LDS byte offset of each lane stream of the part (16-byte
                             // aligned), kNoLane for lanes of other parts
    uint32_t cl[32];         // lane stream bytes
    uint32_t lerr[32];       // first failing run of each lane: q << 8 | status byte
    // The interleave in segments: segment k holds rounds sW[k-1] .. sW[k]-1,
    // in which the 32 - k lanes of mask M[k] are active; it starts at raw
    // word S[k].
    uint32_t sW[32];         // lane word counts, ascending
    uint32_t S[32];
    uint32_t M[32];
    uint32_t fits;
    uint32_t extent_ok;
    uint32_t pub;            // records the walk has published this epoch
    uint32_t claimed;        // records claimed by consumer waves (windows of Rlev2Args::win)
    uint32_t final_;         // records of this epoch once the walk has ended, else kOpen
    uint32_t more;           // a lane stream has runs left after this epoch
    uint32_t total;          // words of all lanes
    uint32_t stuck;          // a consumer gave up waiting (never expected)
    uint32_t pad[2];
    union {
        Run runs[POOL];
        struct {                 // staging only
            uint32_t dest[32][32];   // segment k, j-th active lane -> its LDS word base + the
                                     // segment's first round (kNoLane: another part's lane)
            uint8_t segof[kSegBlocks];  // segment of the first raw word of each 32-word block
        } st;
    };
};
static_assert(sizeof(Run) * PartPool<8>::value >= sizeof(uint32_t) * 32 * 32 + kSegBlocks,
              "staging tables fit the run pool");

// Launch arguments of one RLE v2 container decode.
struct Rlev2Args {
    const uint8_t* payload;
    const uint64_t* blk_off;
    const uint64_t* col_len;
    uint64_t n_chunks, chunk_bytes;
    uint32_t ru, data_cap;
    uint64_t payload_bytes;
    uint8_t* dst;
    uint64_t dst_bytes;
    int32_t* status;
    uint32_t win;            // records per consumer window: 8, 16, 32 or 64 (divides the pool)
};

// Decodes lanes [part*NL, part*NL + NL) of chunk c with one workgroup of NT
// threads; smem holds data_cap + 16 bytes of staged payload (the part's
// lanes) followed by the tables.  With NL = 32 the workgroup decodes the
// whole chunk and writes its status; with NL < 32 each part adds its first
// failing lane to status[c] as an atomic minimum of (lane << 8 | code), which
// status_finish_kernel turns into the code (status_fill_kernel clears it).
template <int EB, uint32_t NL, uint32_t NT>
MDZ_DEV void rlev2_part(const Rlev2Args& a, uint64_t c, uint32_t part, uint32_t* smem) {
    static_assert(NL >= 1 && NL <= 32 && 32 % NL == 0 && NT % kWave == 0 && NT >= 128, "part shape");
    constexpr uint32_t kRunPool = PartPool<NL>::value;
    typedef Rlev2Tables<kRunPool> Tables;
    constexpr uint64_t vmask = EB == 8 ? ~0ull : ((1ull << (8 * EB)) - 1);
    const uint8_t* __restrict__ payload = a.payload;
    const uint64_t* __restrict__ blk_off = a.blk_off;
    const uint64_t* __restrict__ col_len = a.col_len;
    const uint64_t n_chunks = a.n_chunks, chunk_bytes = a.chunk_bytes, payload_bytes = a.payload_bytes;
    const uint32_t ru = a.ru, data_cap = a.data_cap;
    const uint32_t win = a.win;
    uint8_t* __restrict__ dst = a.dst;
    const uint64_t dst_bytes = a.dst_bytes;
    int32_t* __restrict__ status = a.status;
    uint8_t* const data = reinterpret_cast<uint8_t*>(smem);
    Tables& tb = *reinterpret_cast<Tables*>(data + data_cap + 16);
    const uint32_t t0 = part * NL;                 // first lane of the part

    const uint32_t tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    MDZ_PHASE_START();
    const uint32_t nvals = (uint32_t)(chunk_bytes / EB / 32);
    const uint64_t out_base = c * chunk_bytes;
    const uint64_t limit = dst_bytes > out_base ? min(dst_bytes - out_base, chunk_bytes) : 0;
    const uint64_t c_off = blk_off[c];
    // chunk c ends at blk_off[c+1], except the last, whose end is the payload's
    const uint64_t c_end = c + 1 < n_chunks ? blk_off[c + 1] : (payload_bytes ? payload_bytes : ~0ull);

    if (wave == 0) lane_setup<NL>(tb, col_len, c, t0, data_cap, c_off, c_end);
    __syncthreads();
    if (!tb.extent_ok || !tb.fits) {
        // lanes longer than the chunk's payload, or a chunk larger than the
        // LDS budget (the host sizes the budget from the largest chunk)
        if (tid == 0 && status) {
            const int32_t e = tb.extent_ok ? MDZ_ST_OUTPUT_OVERRUN : MDZ_ST_INPUT_OVERRUN;
            if constexpr (NL == 32) status[c] = e;
            else atomicMin(&status[c], (int32_t)(t0 << 8 | ((uint32_t)e & 0xff)));
        }
        return;
    }
    if constexpr (NL == 32) {
        stage_tables(tb, tid, NT);
        __syncthreads();
        MDZ_PHASE(0);
        stage_lanes(tb, smem, reinterpret_cast<const u32x4*>(payload + c_off), tid, NT);
    } else {
        MDZ_PHASE(0);
        stage_part<NL, NT>(tb, smem, reinterpret_cast<const uint32_t*>(payload + c_off), t0, tid);
    }
    __syncthreads();
    MDZ_PHASE(1);

    const uint32_t sh = (ru & (ru - 1)) ? 32u : (uint32_t)__builtin_ctz(ru);
    uint8_t* const out = dst + out_base;
    // Steps 2-4 overlap: wave 0 walks the headers and publishes records as
    // it goes; every wave (wave 0 once its walk ends) claims windows of win
    // published records, checks them one per lane and decodes them.  Only
    // consumers wait, and only for records the walk will publish, so every
    // wait ends; a bound on the wait turns a broken invariant into a status.
    if (tid == 0) { tb.pub = 0; tb.claimed = 0; tb.final_ = kOpen; tb.stuck = 0; }
    __syncthreads();
    // header-chain state of lane stream t0 + tid (threads 0..NL-1 of wave 0)
    uint32_t wp = 0, wq = 0;
    bool done = tid >= NL;
    for (;;) {
        if (wave == 0) {
            // the walk is the chunk's critical path: raise this wave's issue
            // priority over the consumer waves while it runs
            __builtin_amdgcn_s_setprio(2);
            const uint32_t wl = t0 + l % NL;           // lanes past NL idle on a valid lane
            const uint8_t* const base = data + tb.lbase[wl];
            const uint32_t end = tb.cl[wl];
            uint32_t n = 0, pubd = 0;
            for (;;) {
                const bool has = !done && wq < nvals;
                uint32_t size = 0, cnt = 0;
                bool stop = true;
                if (has) chain_step(base, wp, end, size, cnt, stop);
                const uint64_t want = __ballot(has);
                if (!want) break;
                const uint32_t slot = n + __builtin_amdgcn_mbcnt_hi((uint32_t)(want >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)want, 0u));
                if (has && slot < kRunPool) {
                    // every run the chain reaches is recorded, a malformed one
                    // too: its check finds the status
                    *reinterpret_cast<uint4*>(&tb.runs[slot]) = make_uint4(wp, wq, wl << 23, 0u);
                    wp += size;
                    wq += cnt;
                    done = stop;
                }
                n = min(n + (uint32_t)__popcll(want), kRunPool);
                // publish when a consumer window has filled (each publish
                // waits for the records' LDS writes)
                if (n >= pubd + win) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    if (l == 0) __hip_atomic_store(&tb.pub, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    pubd = n;
                }
                if (n >= kRunPool) break;
            }
            __builtin_amdgcn_s_setprio(0);
            done = done || wq >= nvals;
            const uint64_t left = __ballot(!done);
            if (l == 0) {
                tb.more = left != 0;
                __hip_atomic_store(&tb.final_, n, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            MDZ_PHASE(2);
        }
        for (;;) {
            uint32_t w0 = 0;
            if (l == 0) w0 = atomicAdd(&tb.claimed, win);
            w0 = __builtin_amdgcn_readfirstlane(w0);
            uint32_t avail = 0;
            for (uint32_t it = 0;; ++it) {
                const uint32_t fin = __hip_atomic_load(&tb.final_, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (fin != kOpen) { avail = fin; break; }
                const uint32_t pv = __hip_atomic_load(&tb.pub, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (pv >= w0 + win) { avail = w0 + win; break; }
                if (it >= kSpinMax) { if (l == 0) tb.stuck = 1; break; }
                __builtin_amdgcn_s_sleep(2);
            }
            avail = __builtin_amdgcn_readfirstlane(avail);
            if (w0 >= avail) break;
            const uint32_t cnt = min(win, avail - w0);
            // check, one record per lane: the oracle's checks in its order
            // (patch lists included) complete the descriptor in registers,
            // and the run's parameters are decoded here, one run per lane
            Run r{};
            uint64_t pa = 0, pb = 0;
            bool ok = false;
            // lane errors as of now, lane t's in register lane t
            const uint32_t lerr_r = tb.lerr[l & 31];
            if (l < cnt) {
                const Run rec = tb.runs[w0 + l];
                const uint32_t t = rec.meta >> 23;
                const uint8_t* const base = data + tb.lbase[t];
                const uint32_t end = tb.cl[t];
                uint32_t pp = rec.p;
                int32_t e = walk_run(base, end, pp, rec.q, nvals, t, r);
                const uint32_t kind = r.meta & 3;
                if (e == MDZ_ST_OK && kind == 2) e = patch_list_check(base, end, r);
                if (e != MDZ_ST_OK) {
                    atomicMin(&tb.lerr[t], rec.q << 8 | ((uint32_t)e & 0xffu));
                } else {
                    ok = true;
                    if (kind == 0) {
                        pa = lds_bits_be32(base, r.p, 0, 8 * r.aux) & vmask;
                    } else if (kind == 3) {
                        const uint32_t n1 = r.aux & 15, n2 = r.aux >> 4;
                        const uint64_t zz = uvarint_value(base, r.p + n1, n2);
                        pa = uvarint_value(base, r.p, n1) & vmask;
                        pb = ((zz >> 1) ^ (0 - (zz & 1))) & vmask;
                    } else if (kind == 2) {
                        pa = lds_bits_be32(base, r.p, 0, 8 * (r.aux & 15));
                    }
                    r.p += tb.lbase[t];                 // absolute LDS offset for the value pass
                    // SHORT_REPEAT (3..10 values): written here by its lane
                    if (kind == 0 && rec.q < (tb.lerr[t] >> 8)) {
                        const LanePlace lp{t, ru, sh};
                        const uint32_t n = ((r.meta >> 14) & 511) + 1;
                        for (uint32_t i = 0; i < n; ++i) put_val<EB>(out, limit, rec.q + i, lp, pa);
                    }
                }
            }
            // the longer runs: one at a time with the whole wave
            const uint64_t okm = __ballot(ok && (r.meta & 3) != 0);
            for (uint64_t m = okm; m;) {
                const uint32_t j = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                Run rj;
                rj.p = __builtin_amdgcn_readlane(r.p, j);
                rj.q = __builtin_amdgcn_readlane(r.q, j);
                rj.meta = __builtin_amdgcn_readlane(r.meta, j);
                rj.aux = __builtin_amdgcn_readlane(r.aux, j);
                // a run at or after its lane's first known error is not written
                if (rj.q >= ((uint32_t)__builtin_amdgcn_readlane(lerr_r, rj.meta >> 23) >> 8)) continue;
                const uint64_t a = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(pa >> 32), j) << 32 |
                                   (uint32_t)__builtin_amdgcn_readlane((uint32_t)pa, j);
                const uint64_t b = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(pb >> 32), j) << 32 |
                                   (uint32_t)__builtin_amdgcn_readlane((uint32_t)pb, j);
                run_values<EB>(rj, a, b, data, out, limit, ru, sh);
            }
        }
        __syncthreads();
        MDZ_PHASE(3);
        if (!tb.more || tb.stuck) break;
        __syncthreads();                               // everyone has read more
        if (tid == 0) { tb.pub = 0; tb.claimed = 0; tb.final_ = kOpen; }
        __syncthreads();
    }
    if (tid == 0 && status) {
        uint32_t fl = 32, code = 0;                   // the part's first failing lane
        for (uint32_t t = t0; t < t0 + NL; ++t) {
            const uint32_t e = tb.lerr[t];
            if (e != 0xffffffffu) { fl = t; code = e & 0xff; break; }
        }
        if (tb.stuck) { fl = t0; code = (uint32_t)MDZ_ST_INTERNAL & 0xff; }
        if constexpr (NL == 32) {
            status[c] = fl < 32 ? (int32_t)(int8_t)code : MDZ_ST_OK;
        } else {
            if (fl < 32) atomicMin(&status[c], (int32_t)(fl << 8 | code));
        }
    }
}

}  // namespace
}  // namespace mdz
