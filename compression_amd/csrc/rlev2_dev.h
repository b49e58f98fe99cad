// RLE v2 device code shared by the per-chunk kernel and the batched
// persistent kernel.
#pragma once
// RLE v2 (ORC integer RLEv2, lane-transposed) decoder for gfx950.
//
// Replaces the reference's reader-warp / decoder-warp kernel
// (include/rlev2/decoder_new.h:618-657, decoder_warp :354-611) and host entry
// rlev2::decompress_gpu (include/rlev2/decoder_trans.h:615-681).
//
// One workgroup (4 wave64s) per chunk.  The chunk's interleaved payload is
// staged into LDS once and de-interleaved into 32 contiguous lane streams.
// Each wave then decodes whole lane streams cooperatively: the run headers
// are walked once per wave (uniform control flow), and each run's values are
// produced 64 at a time, one per SIMD lane:
//   SHORT_REPEAT  lane i writes the repeated value;
//   DIRECT        lane i unpacks the i-th MSB-first bit field;
//   DELTA         fixed stride: base + i*delta; packed deltas: lane i unpacks
//                 delta i and a wave prefix sum turns deltas into values;
//   PATCHED_BASE  lane k reads patch entry k, a prefix sum of the gaps places
//                 it, and every lane ORs in the patches that land on its value
//                 (64-bit mask: decoder_new.h:533's 32-bit mask drops patch
//                 bits at positions >= 32).
// Values are written straight to their final place: lane t's q-th value is
// element (q / RU) * 32 * RU + t * RU + q % RU of the chunk, so consecutive
// SIMD lanes store consecutive elements.
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

// Decodes lane stream t (LDS bytes [base, base+end)) into its nvals values.
// Called by all 64 lanes of one wave; every branch condition is uniform.
// Status codes follow the oracle's mdz_ref_rlev2_lane_decode.
template <int EB>
MDZ_DEV int32_t rlev2_lane_wave(const uint8_t* base, uint32_t end, uint32_t nvals, uint8_t* out,
                                uint64_t limit, const LanePlace lp) {
    constexpr uint64_t vmask = EB == 8 ? ~0ull : ((1ull << (8 * EB)) - 1);
    const uint32_t l = lane_id();
    uint32_t p = 0, q = 0;
    int32_t err = MDZ_ST_OK;
    while (q < nvals) {
        if (p >= end) return MDZ_ST_INPUT_OVERRUN;
        const uint64_t x = lds_peek8(base, p);
        const uint32_t first = (uint32_t)(x & 0xff), kind = first >> 6;
        if (kind == 0) {                                        // SHORT_REPEAT
            const uint32_t nb = ((first >> 3) & 7) + 1, cnt = (first & 7) + 3;
            if (p + 1 + nb > end) return MDZ_ST_INPUT_OVERRUN;
            if (cnt > nvals - q) return MDZ_ST_OUTPUT_OVERRUN;
            const uint64_t v = lds_bits_be(base, p + 1, 0, 8 * nb) & vmask;
            if (l < cnt) put_val<EB>(out, limit, q + l, lp, v);
            p += 1 + nb;
            q += cnt;
            continue;
        }
        if (p + 1 >= end) return MDZ_ST_INPUT_OVERRUN;
        const uint32_t code = (first >> 1) & 0x1f, w = kDecW[code];
        const uint32_t len = ((((first & 1) << 8)) | (uint32_t)((x >> 8) & 0xff)) + 1;
        if (kind == 1) {                                        // DIRECT
            if (len > nvals - q) return MDZ_ST_OUTPUT_OVERRUN;
            const uint32_t bytes = (len * w + 7) / 8;
            if (p + 2 + bytes > end) return MDZ_ST_INPUT_OVERRUN;
            for (uint32_t i0 = 0; i0 < len; i0 += kWave) {
                const uint32_t i = i0 + l;
                if (i < len)
                    put_val<EB>(out, limit, q + i, lp,
                                lds_bits_be(base, p + 2, (uint64_t)i * w, w) & vmask);
            }
            p += 2 + bytes;
            q += len;
        } else if (kind == 3) {                                 // DELTA
            if (len > nvals - q) return MDZ_ST_OUTPUT_OVERRUN;
            if (len < 2) return MDZ_ST_BAD_SYMBOL;
            uint64_t bv, zd;
            const uint32_t n1 = uvarint(base, p + 2, end, bv, err);
            if (!n1) return err;
            const uint32_t n2 = uvarint(base, p + 2 + n1, end, zd, err);
            if (!n2) return err;
            const uint32_t pd = p + 2 + n1 + n2;
            const uint64_t dt = ((zd >> 1) ^ (0 - (zd & 1))) & vmask;
            const uint64_t v0 = bv & vmask, v1 = (v0 + dt) & vmask;
            if (code == 0) {                                    // fixed stride
                for (uint32_t i0 = 0; i0 < len; i0 += kWave) {
                    const uint32_t i = i0 + l;
                    if (i < len) put_val<EB>(out, limit, q + i, lp, (v0 + (uint64_t)i * dt) & vmask);
                }
                p = pd;
            } else {
                const uint32_t nd = len - 2, bytes = (nd * w + 7) / 8;
                if (pd + bytes > end) return MDZ_ST_INPUT_OVERRUN;
                if (l < 2) put_val<EB>(out, limit, q + l, lp, l ? v1 : v0);
                // multiplier = (DATA_TYPE)delta > 0 ? 1 : -1 on the unsigned type
                const bool neg = dt == 0;
                uint64_t carry = v1;
                for (uint32_t k0 = 0; k0 < nd; k0 += kWave) {
                    const uint32_t k = k0 + l;
                    uint64_t d = k < nd ? lds_bits_be(base, pd, (uint64_t)k * w, w) : 0;
                    if (neg) d = 0 - d;
                    const uint64_t incl = wave_incl_scan_u64(d);
                    if (k < nd) put_val<EB>(out, limit, q + 2 + k, lp, (carry + incl) & vmask);
                    carry += shfl_u64(incl, kWave - 1);
                }
                p = pd + bytes;
            }
            q += len;
        } else {                                                // PATCHED_BASE
            if (p + 3 >= end) return MDZ_ST_INPUT_OVERRUN;
            const uint32_t third = (uint32_t)((x >> 16) & 0xff), fourth = (uint32_t)((x >> 24) & 0xff);
            const uint32_t bw = ((third >> 5) & 7) + 1, pw = kDecW[third & 0x1f];
            const uint32_t pgw = ((fourth >> 5) & 7) + 1, pll = fourth & 0x1f;
            if (len > nvals - q) return MDZ_ST_OUTPUT_OVERRUN;
            if (p + 4 + bw > end) return MDZ_ST_INPUT_OVERRUN;
            const uint64_t bv = lds_bits_be(base, p + 4, 0, 8 * bw);
            const uint32_t pb = p + 4 + bw, vbytes = (len * w + 7) / 8;
            if (pb + vbytes > end) return MDZ_ST_INPUT_OVERRUN;
            const uint32_t cfb = closest_bits(pw + pgw), pp = pb + vbytes;
            // patch entry k in lane k; the first entry (in order) that cannot
            // be read or lands past the run decides the status
            const uint64_t pmask = pw >= 64 ? ~0ull : ((1ull << pw) - 1);
            const bool readable = (uint64_t)pp * 8 + (uint64_t)(l + 1) * cfb <= (uint64_t)end * 8;
            const uint64_t e = (l < pll && readable) ? lds_bits_be(base, pp, (uint64_t)l * cfb, cfb) : 0;
            const uint32_t gap = (l < pll && pw < 64) ? (uint32_t)(e >> pw) : 0;
            const uint32_t pos = wave_incl_scan_u32(gap);
            const uint64_t bad = __ballot(l < pll && (!readable || pos >= len));
            if (bad) {
                const uint32_t k = (uint32_t)__builtin_ctzll(bad);
                return __shfl((int)readable, (int)k, kWave) ? MDZ_ST_BAD_SYMBOL : MDZ_ST_INPUT_OVERRUN;
            }
            const uint64_t pv = e & pmask;
            for (uint32_t i0 = 0; i0 < len; i0 += kWave) {
                const uint32_t i = i0 + l;
                uint64_t v = i < len ? lds_bits_be(base, pb, (uint64_t)i * w, w) : 0;
                for (uint32_t k = 0; k < pll; ++k) {
                    const uint32_t pk = __shfl(pos, (int)k, kWave);
                    const uint64_t vk = shfl_u64(pv, (int)k);
                    if (pk == i) v |= w >= 64 ? 0 : vk << w;
                }
                if (i < len) put_val<EB>(out, limit, q + i, lp, (v + bv) & vmask);
            }
            p = pp + (pll * cfb + 7) / 8;
            q += len;
        }
    }
    return MDZ_ST_OK;
}

struct Rlev2Tables {
    uint32_t lbase[32];
    uint32_t cl[32];
    uint32_t err;        // min over lane << 8 | status byte
    uint32_t fits;
    uint32_t extent_ok;
    uint32_t pad;
    LaneTable lt;
};

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
};

// Decodes chunk c with the whole workgroup (kThreads threads); smem holds
// data_cap + 16 bytes of staged payload followed by the tables.
template <int EB>
MDZ_DEV void rlev2_chunk(const Rlev2Args& a, uint64_t c, uint32_t* smem) {
    const uint8_t* __restrict__ payload = a.payload;
    const uint64_t* __restrict__ blk_off = a.blk_off;
    const uint64_t* __restrict__ col_len = a.col_len;
    const uint64_t n_chunks = a.n_chunks, chunk_bytes = a.chunk_bytes, payload_bytes = a.payload_bytes;
    const uint32_t ru = a.ru, data_cap = a.data_cap;
    uint8_t* __restrict__ dst = a.dst;
    const uint64_t dst_bytes = a.dst_bytes;
    int32_t* __restrict__ status = a.status;
    uint8_t* const data = reinterpret_cast<uint8_t*>(smem);
    Rlev2Tables& tb = *reinterpret_cast<Rlev2Tables*>(data + data_cap + 16);

    const uint32_t tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    MDZ_PHASE_START();
    const uint32_t nvals = (uint32_t)(chunk_bytes / EB / 32);
    const uint64_t out_base = c * chunk_bytes;
    const uint64_t limit = dst_bytes > out_base ? min(dst_bytes - out_base, chunk_bytes) : 0;
    const uint64_t c_off = blk_off[c];
    // chunk c ends at blk_off[c+1], except the last, whose end is the payload's
    const uint64_t c_end = c + 1 < n_chunks ? blk_off[c + 1] : (payload_bytes ? payload_bytes : ~0ull);

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
        if (l == 31) tb.fits = incl <= data_cap;
        if (l == 0) {
            tb.err = 0xffffffffu;
            tb.extent_ok = !(c_off & 15) && c_end >= c_off &&
                           (uint64_t)tb.lt.total * 4 <= c_end - c_off;
        }
    }
    __syncthreads();
    if (!tb.extent_ok || !tb.fits) {
        // lanes longer than the chunk's payload, or a chunk larger than the
        // LDS budget (the host sizes the budget from the largest chunk)
        if (tid == 0 && status) status[c] = tb.extent_ok ? MDZ_ST_OUTPUT_OVERRUN : MDZ_ST_INPUT_OVERRUN;
        return;
    }
    MDZ_PHASE(0);
    stage_chunk_lanes(tb.lt, tb.lbase, smem, reinterpret_cast<const u32x4*>(payload + c_off), tid,
                      kThreads);
    __syncthreads();
    MDZ_PHASE(1);

    const uint32_t sh = (ru & (ru - 1)) ? 32u : (uint32_t)__builtin_ctz(ru);
    for (uint32_t t = wave; t < 32; t += kThreads / kWave) {
        const int32_t e = rlev2_lane_wave<EB>(data + tb.lbase[t], tb.cl[t], nvals, dst + out_base,
                                              limit, LanePlace{t, ru, sh});
        if (e != MDZ_ST_OK && l == 0) atomicMin(&tb.err, (t << 8) | ((uint32_t)e & 0xffu));
    }
    MDZ_PHASE(2);
    __syncthreads();
    if (tid == 0 && status) {
        const uint32_t e = tb.err;
        status[c] = e == 0xffffffffu ? MDZ_ST_OK : (int32_t)(int8_t)(e & 0xff);
    }
    MDZ_PHASE(3);
}

}  // namespace
}  // namespace mdz
