// RLE v2 encode for gfx950 (lane-transposed container).
//
// Replaces rlev2::compress_gpu_transpose (reference include/rlev2/
// encoder_trans.h:428-537: block_encode :14-289 per lane, the transpose
// :291-349, the scans :466-492).  Payload, col_len and blk_off (blk_off[n]
// holding the uncompressed size, as the reference leaves it) equal that
// encoder's output byte for byte: run grouping of block_encode, the sub-mode
// choice of determineEncoding and the four writers (encoder.h:124-785).
//
// One wave per chunk; lane t < 32 encodes lane stream t, whose q-th value is
// chunk element (q / RU) * 32 * RU + t * RU + q % RU, read straight from the
// input.  A lane's pending values (up to 128) live in LDS ([i][lane], 32 KiB
// per wave); deltas and patched-base offsets are recomputed from them, and
// the patched-base writer builds its patch list in a second pass instead of
// storing it.  The scan and the
// interleave are the shared lane-encoder kernels (lanes_enc_dev.h).
#include "lanes_enc_dev.h"
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

constexpr uint32_t kMinRepeat = 3, kMaxLit = 128, kMaxShortRepeat = 10;

// Pending values; u32 storage when elements are 4 bytes or less (16 KiB per
// wave instead of 32), arithmetic always in 64 bits.
template <class LT>
struct V2EncLds {
    LT lit[kMaxLit][32];
};

// The 5-bit width code of a bit width, and the widths each writer rounds to.
MDZ_DEV uint32_t width_code(uint32_t bits) {
    if (bits <= 1) return 0;
    if (bits <= 24) return bits - 1;
    if (bits <= 26) return 24;
    if (bits <= 28) return 25;
    if (bits <= 30) return 26;
    if (bits <= 32) return 27;
    if (bits <= 40) return 28;
    if (bits <= 48) return 29;
    if (bits <= 56) return 30;
    return 31;
}
MDZ_DEV uint32_t code_width(uint32_t code) {
    return code < 24 ? code + 1 : code < 28 ? 26 + 2 * (code - 24) : 40 + 8 * (code - 28);
}
MDZ_DEV uint32_t fixed_width(uint32_t bits) {            // nearest encodable width >= bits
    return code_width(width_code(bits));
}
MDZ_DEV uint32_t aligned_width(uint32_t bits) {          // 1, 2, 4, 8, 16, 24, 32, ... 64
    if (bits <= 1) return 1;
    if (bits <= 2) return 2;
    if (bits <= 4) return 4;
    if (bits <= 8) return 8;
    if (bits <= 16) return 16;
    if (bits <= 24) return 24;
    if (bits <= 32) return 32;
    return fixed_width(bits);                       // 40, 48, 56, 64
}
MDZ_DEV uint32_t value_width(uint64_t v) {               // fixed width of the bit length
    return fixed_width(v ? 64 - __builtin_clzll(v) : 0);
}

template <class LT>
struct LaneEnc {
    V2EncLds<LT>& S;
    uint32_t t;
    uint8_t* out;
    uint64_t pos = 0;
    uint32_t n = 0, fix = 0, var = 0, delta_bits = 0;
    bool fixed_delta = false;

    MDZ_DEV LT& lit(uint32_t i) { return S.lit[i][t]; }
    // The reference keeps a delta array; every delta it writes out equals
    // lit(i + 1) - lit(i) (runs of equal values have delta 0), so it is
    // recomputed instead of stored.
    MDZ_DEV uint64_t dl(uint32_t i) { return (uint64_t)lit(i + 1) - (uint64_t)lit(i); }
    MDZ_DEV void put(uint32_t b) { out[pos++] = (uint8_t)b; }

    // Big-endian byte-aligned fields (widths >= 8) or MSB-first sub-byte fields.
    template <class Get>
    MDZ_DEV void write_aligned(Get get, uint32_t len, uint32_t bits) {
        if (bits < 8) {
            const uint32_t mask = (1u << bits) - 1, hops = 8 / bits, rem = len % hops, full = len - rem;
            for (uint32_t i = 0; i < full; i += hops) {
                uint32_t b = 0;
                for (uint32_t j = 0; j < hops; ++j) b |= ((uint32_t)get(i + j) & mask) << (8 - (j + 1) * bits);
                put(b & 0xFF);
            }
            if (rem) {
                int32_t sh = 8 - (int32_t)bits;
                uint32_t b = 0;
                for (uint32_t i = full; i < len; ++i) {
                    b |= ((uint32_t)get(i) & mask) << sh;
                    sh -= (int32_t)bits;
                }
                put(b & 0xFF);
            }
        } else {
            const int32_t bytes = (int32_t)bits / 8;
            for (uint32_t i = 0; i < len; ++i) {
                const uint64_t v = get(i);
                for (int32_t j = bytes - 1; j >= 0; --j) put((uint32_t)(v >> (j * 8)) & 0xFF);
            }
        }
    }
    // MSB-first bit packing across byte boundaries; a 64-bit field keeps all bits.
    struct BitWriter {
        LaneEnc& e;
        uint32_t left = 8, cur = 0;
        MDZ_DEV void field(uint64_t v, uint32_t bits) {
            uint32_t todo = bits;
            while (todo > left) {
                cur |= (uint32_t)(v >> (todo - left)) & 0xFF;
                todo -= left;
                v &= todo >= 64 ? ~0ull : ((1ull << todo) - 1);
                e.put(cur & 0xFF);
                cur = 0;
                left = 8;
            }
            left -= todo;
            cur |= (uint32_t)(v << left) & 0xFF;
            if (left == 0) {
                e.put(cur);
                cur = 0;
                left = 8;
            }
        }
        MDZ_DEV void flush() {
            if (left != 8) e.put(cur);
        }
    };
    MDZ_DEV void varint(uint64_t v) {
        for (;;) {
            if ((v & ~0x7Full) == 0) {
                put((uint32_t)v);
                return;
            }
            put(0x80 | (uint32_t)(v & 0x7F));
            v >>= 7;
        }
    }
    // Width at the p-th percentile of fixed widths of get(0..len) (histogram
    // over the 32 width codes, counted from the top).
    template <class Get>
    MDZ_DEV uint32_t percentile(Get get, uint32_t len, uint32_t p) {
        uint32_t hist[32];
        for (uint32_t c = 0; c < 32; ++c) hist[c] = 0;
        for (uint32_t i = 0; i < len; ++i) ++hist[width_code(value_width(get(i)))];
        p = 100 - p;
        int32_t plen = p == 0 ? 0 : (int32_t)((len * p - 1) / 100 + 1);
        for (int32_t c = 31; c >= 0; --c) {
            plen -= (int32_t)hist[c];
            if (plen < 0) return code_width((uint32_t)c);
        }
        return 0;
    }

    MDZ_DEV void short_repeat() {
        const uint64_t v = lit(0);
        const uint32_t vbytes = (value_width(v) - 1) / 8 + 1;
        put(((vbytes - 1) << 3) | (fix - kMinRepeat));
        for (int32_t i = (int32_t)vbytes - 1; i >= 0; --i) put((uint32_t)(v >> (i * 8)) & 0xFF);
        fix = 0;
        n = 0;
        var = 0;
    }
    MDZ_DEV void direct() {
        auto L = [&](uint32_t i) { return (uint64_t)lit(i); };
        const uint32_t bits = aligned_width(percentile(L, n, 100));
        const uint32_t len = var - 1;
        put(0x40 | (width_code(bits) << 1) | ((len & 0x100) >> 8));
        put(len & 0xFF);
        write_aligned(L, n, bits);
        var = 0;
        n = 0;
    }
    MDZ_DEV void delta() {
        uint32_t len, code = 0, bits = aligned_width(delta_bits);
        if (fixed_delta) {
            if (fix > kMinRepeat) {
                len = fix - 1;
                fix = 0;
            } else {
                len = var - 1;
                var = 0;
            }
        } else {
            if (bits == 1) bits = 2;
            code = width_code(bits);
            len = var - 1;
            var = 0;
        }
        put(0xC0 | (code << 1) | ((len & 0x100) >> 8));
        put(len & 0xFF);
        varint(lit(0));
        const int64_t d0 = n > 1 ? (int64_t)dl(0) : 0;
        varint(((uint64_t)d0 << 1) ^ (uint64_t)(d0 >> 63));          // zigzag
        if (!fixed_delta) write_aligned([&](uint32_t i) { return dl(i + 1); }, n - 2, bits);
        n = 0;
        var = 0;
    }
    // red(i) = lit(i) - base, recomputed where needed.
    MDZ_DEV void patched(uint64_t base, uint32_t b95, uint32_t b100) {
        auto red = [&](uint32_t i) { return (uint64_t)lit(i) - base; };
        uint64_t mask = (1ull << b95) - 1;
        uint32_t pw = fixed_width(b100 - b95);
        if (pw == 64) {
            pw = 56;
            b95 = 8;
            mask = (1ull << b95) - 1;
        }
        // pass 1: patch count and largest gap (gaps stay below 128)
        uint32_t np = 0, prev = 0, max_gap = 0;
        for (uint32_t i = 0; i < n; ++i)
            if (red(i) > mask) {
                max_gap = max(max_gap, i - prev);
                prev = i;
                ++np;
            }
        const uint32_t pgw = (max_gap == 0 && np != 0) ? 1 : value_width(max_gap);
        const uint32_t varlen = var - 1;
        const uint32_t vbytes = (value_width(base) + 1 - 1) / 8 + 1;
        put(0x80 | (width_code(b95) << 1) | ((varlen & 0x100) >> 8));
        put(varlen & 0xFF);
        put(((vbytes - 1) << 5) | width_code(pw));
        put(((pgw - 1) << 5) | np);
        for (int32_t i = (int32_t)vbytes - 1; i >= 0; --i) put(i >= 8 ? 0 : (uint32_t)(base >> (i * 8)) & 0xFF);
        {
            BitWriter bw{*this};
            const uint32_t w = fixed_width(b95);
            for (uint32_t i = 0; i < n; ++i) bw.field(red(i) & mask, w);
            bw.flush();
        }
        {   // pass 2: the patch list, gap << pw | high bits, in order
            BitWriter bw{*this};
            const uint32_t w = fixed_width(pgw + pw);
            prev = 0;
            for (uint32_t i = 0; i < n; ++i)
                if (red(i) > mask) {
                    bw.field((uint64_t)(i - prev) << pw | (red(i) >> b95), w);
                    prev = i;
                }
            bw.flush();
        }
        var = 0;
        n = 0;
    }
    MDZ_DEV void determine() {
        if (n <= kMinRepeat) {
            direct();
            return;
        }
        bool inc = true, dec = true;
        fixed_delta = true;
        uint64_t mn = lit(0), mx = lit(0), max_delta = 0;
        const uint64_t d_init = (uint64_t)lit(1) - (uint64_t)lit(0);
        for (uint32_t i = 1; i < n; ++i) {
            const uint64_t l1 = lit(i), l0 = lit(i - 1), cd = l1 - l0;
            mn = l1 < mn ? l1 : mn;
            mx = l1 > mx ? l1 : mx;
            inc = inc && l0 <= l1;
            dec = dec && l0 >= l1;
            fixed_delta = fixed_delta && cd == d_init;
            max_delta = cd > max_delta ? cd : max_delta;
        }
        // the reference's signed-subtract safety test with INT8_MIN fails
        // only when min == 0 and max <= 128
        if (!(mn > 0 || mx > mn + 128)) {
            direct();
            return;
        }
        if (mn == mx || fixed_delta) {
            delta();
            return;
        }
        if (d_init != 0) {
            delta_bits = value_width(max_delta);
            if (inc || dec) {
                delta();
                return;
            }
        }
        auto L = [&](uint32_t i) { return (uint64_t)lit(i); };
        const uint32_t b100 = percentile(L, n, 100), b90 = percentile(L, n, 90);
        if (b100 - b90 > 1) {
            auto R = [&](uint32_t i) { return (uint64_t)lit(i) - mn; };
            const uint32_t r95 = percentile(R, n, 95), r100 = percentile(R, n, 100);
            if (r100 != r95) {
                patched(mn, r95, r100);
                return;
            }
        }
        direct();
    }

    // block_encode's grouping over the lane's values (encoder_trans.h:82-196)
    template <class Val>
    MDZ_DEV uint64_t encode(Val val_at, uint64_t nv) {
        uint64_t prev_delta = 0;
        for (uint64_t k = 0; k < nv; ++k) {
            const uint64_t val = val_at(k);
            if (n == 0) {
                lit(n++) = val;
                fix = 1;
                var = 1;
                continue;
            }
            if (n == 1) {
                lit(n++) = val;
                prev_delta = val - (uint64_t)lit(0);
                if (val == (uint64_t)lit(0)) {
                    fix = 2;
                    var = 0;
                } else {
                    fix = 0;
                    var = 2;
                }
                continue;
            }
            const uint64_t cur_delta = val - (uint64_t)lit(n - 1);
            if (prev_delta == 0 && cur_delta == 0) {
                lit(n++) = val;
                if (var > 0) fix = 2;
                ++fix;
                if (fix >= kMinRepeat && var > 0) {
                    n -= kMinRepeat;
                    var -= kMinRepeat - 1;
                    determine();
                    prev_delta = 0;
                    for (uint32_t i = 0; i < kMinRepeat; ++i) lit(i) = val;
                    n = kMinRepeat;
                }
                if (fix == kMaxLit) {
                    if (var == 0) var = n;          // the run's literal count is its length
                    determine();
                }
                continue;
            }
            if (fix >= kMinRepeat) {
                if (fix <= kMaxShortRepeat) {
                    short_repeat();
                } else {
                    fixed_delta = true;
                    delta();
                }
            }
            if (fix > 0 && fix < kMinRepeat && val != (uint64_t)lit(n - 1)) {
                var = fix;
                fix = 0;
            }
            if (n == 0) {
                lit(n++) = val;
                fix = 1;
                var = 1;
            } else {
                prev_delta = val - (uint64_t)lit(n - 1);
                lit(n++) = (LT)val;
                ++var;
                if (var == kMaxLit) determine();
            }
        }
        if (n != 0) {
            if (var != 0) {
                determine();
            } else if (fix != 0) {
                if (fix < kMinRepeat) {
                    var = fix;
                    fix = 0;
                    determine();
                } else if (fix <= kMaxShortRepeat) {
                    short_repeat();
                } else {
                    fixed_delta = true;
                    delta();
                }
            }
        }
        return pos;
    }
};

template <class LT>
__global__ __launch_bounds__(kWave) void rlev2_encode_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes,
                                                             uint64_t chunk, uint32_t eb, uint32_t ru,
                                                             uint64_t lane_slot, uint8_t* __restrict__ work,
                                                             uint64_t* __restrict__ col_len,
                                                             uint64_t* __restrict__ chunk_len) {
    __shared__ V2EncLds<LT> S;
    const uint32_t l = lane_id();
    const uint64_t c = blockIdx.x;
    uint64_t len = 0;
    if (l < 32) {
        const uint64_t base = c * chunk, nv = chunk / eb / 32;
        auto val_at = [&](uint64_t q) -> uint64_t {
            const uint64_t el = (q / ru) * 32 * ru + (uint64_t)l * ru + q % ru;
            const uint64_t off = base + el * eb;
            if (off + eb > in_bytes) return 0;           // a partial last value reads as 0
            const uint8_t* p = in + off;
            if (((uintptr_t)p & (eb - 1)) == 0) {         // one load of the element's width
                if (eb == 8) return *reinterpret_cast<const uint64_t*>(p);
                if (eb == 4) return *reinterpret_cast<const uint32_t*>(p);
                if (eb == 2) return *reinterpret_cast<const uint16_t*>(p);
                return *p;
            }
            uint64_t v = 0;
            for (uint32_t b = 0; b < eb; ++b) v |= (uint64_t)p[b] << (8 * b);
            return v;
        };
        LaneEnc<LT> e{S, l, work + (c * 32 + l) * lane_slot};
        len = e.encode(val_at, nv);
        col_len[c * 32 + l] = len;
    }
    const uint32_t total = wave_sum_u32((uint32_t)((len + 3) / 4));
    if (l == 0) chunk_len[c] = ((uint64_t)total * 4 + 127) / 128 * 128;
}

uint64_t v2_lane_slot(uint64_t chunk, uint32_t eb) {
    return (chunk / eb / 32 * 10 + 64 + 15) & ~15ull;
}

}  // namespace

int rlev2_compress_bound(uint64_t in_bytes, uint64_t chunk, uint32_t eb, uint32_t ru, uint64_t* n_chunks,
                         uint64_t* work_bytes, uint64_t* payload_cap) {
    if (!(eb == 1 || eb == 2 || eb == 4 || eb == 8) || ru == 0 || chunk % (32ull * eb) ||
        (chunk / eb / 32) % ru || in_bytes == 0 || chunk > (1ull << 32))
        return MDZ_ERR_INVALID;
    const uint64_t n = (in_bytes + chunk - 1) / chunk;
    const uint64_t slot = v2_lane_slot(chunk, eb);
    if (n_chunks) *n_chunks = n;
    if (work_bytes) *work_bytes = n * 32 * slot + 8 * n;
    if (payload_cap) *payload_cap = n * ((32 * slot + 127) / 128 * 128);
    return MDZ_OK;
}

int rlev2_compress(const void* in, uint64_t in_bytes, uint64_t chunk, uint32_t eb, uint32_t ru, void* work,
                   void* payload, uint64_t* col_len, uint64_t* blk_off, uint64_t* payload_bytes, hipStream_t s) {
    uint64_t n = 0;
    int rc = rlev2_compress_bound(in_bytes, chunk, eb, ru, &n, nullptr, nullptr);
    if (rc) return rc;
    if (!in || !work || !payload || !col_len || !blk_off || n > (1ull << 24)) return MDZ_ERR_INVALID;
    const uint64_t slot = v2_lane_slot(chunk, eb);
    uint64_t* chunk_len = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(work) + n * 32 * slot);
    if (eb <= 4)
        hipLaunchKernelGGL(rlev2_encode_kernel<uint32_t>, dim3((uint32_t)n), dim3(kWave), 0, s,
                           static_cast<const uint8_t*>(in), in_bytes, chunk, eb, ru, slot,
                           static_cast<uint8_t*>(work), col_len, chunk_len);
    else
        hipLaunchKernelGGL(rlev2_encode_kernel<uint64_t>, dim3((uint32_t)n), dim3(kWave), 0, s,
                           static_cast<const uint8_t*>(in), in_bytes, chunk, eb, ru, slot,
                           static_cast<uint8_t*>(work), col_len, chunk_len);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return record_hip(e, "rlev2_encode_kernel launch");
    // blk_off[n] holds the uncompressed size; the true payload end goes to *payload_bytes
    hipLaunchKernelGGL(lanes_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, chunk_len, n, blk_off, payload_bytes,
                       in_bytes);
    if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "lanes_scan_kernel launch");
    hipLaunchKernelGGL(lanes_interleave_kernel, dim3((uint32_t)n), dim3(kIlvThreads), 0, s,
                       static_cast<const uint8_t*>(work), slot, col_len, blk_off, chunk_len, nullptr,
                       static_cast<uint8_t*>(payload));
    return record_hip(hipGetLastError(), "lanes_interleave_kernel launch");
}

}  // namespace mdz
