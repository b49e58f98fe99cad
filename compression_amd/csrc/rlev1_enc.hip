// RLE v1 encode for gfx950 (lane-transposed container).
//
// Replaces rle_v1::compress_gpu and its kernels (reference include/rle_v1/
// rlev1.h:24-97: setup_deflate sizes every lane stream, reduction_scan runs
// one thread over the chunk sizes, deflate encodes again and interleaves).
// The payload, col_len and blk_off are byte-identical to that encoder's
// (compression_warp, include/rle_v1/encoder.h:607-1018).
//
// Three launches, each encoding only once:
//   1. one wave per chunk; lane t < 32 encodes lane stream t, reading its
//      values straight from the chunk's lines t, t+32, ... (zero past the
//      input) and writing the stream to its own slot of the work buffer; the
//      wave then sums the rounded lane lengths into the chunk's length;
//   2. one workgroup scans the chunk lengths into blk_off;
//   3. one workgroup per chunk interleaves the lane streams in the
//      round-robin order of encoder.h:564-588 (lanes_enc_dev.h).
#include "lanes_enc_dev.h"
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

// eb little-endian bytes at `at` (bytes past the input read as 0): one load
// of the element's width when it is whole and aligned, else bytewise.
MDZ_DEV uint64_t load_elem(const uint8_t* in, uint64_t in_bytes, uint64_t at, uint32_t eb) {
    const uint8_t* p = in + at;
    if (at + eb <= in_bytes && ((uintptr_t)p & (eb - 1)) == 0) {
        if (eb == 8) return *reinterpret_cast<const uint64_t*>(p);
        if (eb == 4) return *reinterpret_cast<const uint32_t*>(p);
        if (eb == 2) return *reinterpret_cast<const uint16_t*>(p);
        return *p;
    }
    uint64_t v = 0;
    for (uint32_t b = 0; b < eb; ++b)
        if (at + b < in_bytes) v |= (uint64_t)in[at + b] << (8 * b);
    return v;
}

// Value i of lane t: lane byte i*eb sits in line t + 32*(i*eb / line).
MDZ_DEV uint64_t lane_value(const uint8_t* in, uint64_t in_bytes, uint64_t cbase, uint32_t line, uint32_t t,
                            uint64_t i, uint32_t eb) {
    const uint64_t o = i * eb, k = o / line, j = o - k * line;
    const uint64_t at = cbase + (t + 32 * k) * line + j;
    return load_elem(in, in_bytes, at, eb);
}

// `read - prev` and `prev + delta == read` in the element type's arithmetic
// (u8/u16 promote to int; u32 and u64 wrap), as encoder.h evaluates them.
MDZ_DEV int64_t elem_diff(uint64_t rd, uint64_t pv, uint32_t eb) {
    if (eb <= 2) return (int64_t)((int32_t)rd - (int32_t)pv);
    if (eb == 4) return (int64_t)(uint32_t)((uint32_t)rd - (uint32_t)pv);
    return (int64_t)(rd - pv);
}
MDZ_DEV bool elem_match(uint64_t pv, int32_t d, uint64_t rd, uint32_t eb) {
    if (eb <= 2) return (int32_t)pv + d == (int32_t)rd;
    if (eb == 4) return (uint32_t)((uint32_t)pv + (uint32_t)d) == (uint32_t)rd;
    return pv + (uint64_t)(int64_t)d == rd;
}

struct LaneOut {
    uint8_t* out;
    uint64_t pos;
    MDZ_DEV void byte(uint32_t b) { out[pos++] = (uint8_t)b; }
    MDZ_DEV void varint(uint64_t v) {               // unsigned LEB128
        do {
            uint32_t b = (uint32_t)(v & 0x7F);
            v >>= 7;
            byte(v ? b | 0x80 : b);
        } while (v);
    }
    MDZ_DEV void patch(uint64_t at, int32_t lit) { out[at] = (uint8_t)(int8_t)(-lit); }
};

// One lane stream: the greedy run / literal-group encoder.  State: the last
// two values (ring `ring`, `head`, `pend` of them not yet emitted), the run
// being grown (`cnt` values with step `delta`, `ok` while a step exists,
// first value `first`) and the open literal group (`lit` values, header at
// `lit_at`, written as a placeholder and patched to -lit when it closes).
MDZ_DEV uint64_t encode_lane(const uint8_t* in, uint64_t in_bytes, uint64_t cbase, uint32_t line, uint32_t t,
                             uint64_t n, uint32_t eb, uint8_t* dst) {
    LaneOut w{dst, 0};
    uint64_t ring[2] = {0, 0};
    uint32_t head = 0, pend = 0;
    int32_t cnt, lit = 0, delta = 0;
    bool ok;
    uint64_t first = 0, lit_at = 0, prev;

    uint64_t x = lane_value(in, in_bytes, cbase, line, t, 0, eb);
    cnt = 1;
    prev = x;
    ring[head] = x;
    head ^= 1;
    ++pend;
    x = lane_value(in, in_bytes, cbase, line, t, 1, eb);
    int64_t d = elem_diff(x, prev, eb);
    if (d > 127 || d < -128) {
        ok = false;
        lit_at = w.pos;
        w.varint(1);
        lit = 1;
        w.varint(ring[0]);
        --pend;
    } else {
        ok = true;
        delta = (int32_t)d;
        ++cnt;
    }
    prev = x;
    ring[head] = x;
    head ^= 1;
    ++pend;

    for (uint64_t i = 2; i < n; ++i) {
        if (lit == 127) {
            w.patch(lit_at, lit);
            lit = 0;
        }
        x = lane_value(in, in_bytes, cbase, line, t, i, eb);
        if (cnt == 1) {
            d = elem_diff(x, prev, eb);
            if (d > 127 || d < -128) {
                ok = false;
                if (lit == 0) {
                    lit_at = w.pos;
                    w.varint(1);
                }
                w.varint(ring[head ^ 1]);
                ++lit;
                --pend;
            } else {
                ok = true;
                delta = (int32_t)d;
                ++cnt;
            }
            ++pend;
        } else if (ok && elem_match(prev, delta, x, eb)) {
            ++cnt;
            if (cnt == 3) {
                first = ring[head];
                if (lit != 0) {
                    w.patch(lit_at, lit);
                    lit = 0;
                }
            } else if (cnt == 130) {                    // a run holds at most 129 values
                w.byte((uint32_t)(cnt - 4) & 0xFF);
                w.byte((uint32_t)delta & 0xFF);
                w.varint(first);
                cnt = 1;
                pend = 0;
                lit = 0;
            }
            pend = pend + 1 < 2 ? pend + 1 : 2;
        } else if (cnt >= 3) {                          // the run ends before x
            w.byte((uint32_t)(cnt - 3) & 0xFF);
            w.byte((uint32_t)delta & 0xFF);
            w.varint(first);
            cnt = 1;
            lit = 0;
            pend = 1;
        } else {                                        // a pair that is not a run
            if (lit == 0) {
                lit_at = w.pos;
                w.byte(3);
            }
            ++lit;
            w.varint(ring[head]);
            if (lit == 127) {
                w.patch(lit_at, lit);
                lit = 0;
            }
            d = elem_diff(x, prev, eb);
            if (d > 127 || d < -128) {
                if (lit == 0) {
                    lit_at = w.pos;
                    w.byte(3);
                }
                ok = false;
                pend = 0;
                w.varint(ring[head ^ 1]);
                ++lit;
                cnt = 1;
            } else {
                pend = 1;
                ok = true;
                delta = (int32_t)d;
                cnt = 2;
            }
            pend = pend + 1 < 2 ? pend + 1 : 2;
        }
        prev = x;
        ring[head] = x;
        head ^= 1;
    }

    if (cnt >= 3 && ok) {                               // the tail run
        w.byte((uint32_t)(cnt - 3) & 0xFF);
        w.byte((uint32_t)delta & 0xFF);
        w.varint(first);
    } else {                                            // the tail values as literals
        if (pend == 1) {
            if (lit == 127) {
                w.patch(lit_at, lit);
                lit = 0;
            }
            if (lit == 0) {
                lit_at = w.pos;
                w.varint(1);
            }
            w.varint(ring[head ^ 1]);
            ++lit;
        }
        if (pend == 2) {
            if (lit == 127) {
                w.patch(lit_at, lit);
                lit = 0;
            }
            if (lit == 0) {
                lit_at = w.pos;
                w.byte(1);
            }
            w.varint(ring[head]);
            ++lit;
            if (lit == 127) {
                w.patch(lit_at, lit);
                lit = 0;
            }
            if (lit == 0) {
                lit_at = w.pos;
                w.byte(1);
            }
            head ^= 1;
            w.varint(ring[head]);
            ++lit;
        }
        w.patch(lit_at, lit);                           // closes the group (always written)
    }
    return w.pos;
}

__global__ __launch_bounds__(kWave) void rlev1_encode_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes,
                                                             uint64_t chunk, uint32_t eb, uint32_t line,
                                                             uint64_t lane_slot, uint8_t* __restrict__ work,
                                                             uint64_t* __restrict__ col_len,
                                                             uint64_t* __restrict__ chunk_len) {
    const uint32_t l = lane_id();
    const uint64_t c = blockIdx.x;
    uint64_t len = 0;
    if (l < 32) {
        const uint64_t n = chunk / 32 / eb;
        len = encode_lane(in, in_bytes, c * chunk, line, l, n, eb, work + (c * 32 + l) * lane_slot);
        col_len[c * 32 + l] = len;
    }
    // the chunk: every lane rounded up to words, the sum to 128 bytes
    uint32_t words = (uint32_t)((len + 3) / 4);
    const uint32_t total = wave_sum_u32(words);
    if (l == 0) chunk_len[c] = ((uint64_t)total * 4 + 127) / 128 * 128;
}

uint64_t lane_slot_bytes(uint64_t chunk, uint32_t eb) {
    const uint64_t n = chunk / 32 / eb;
    return (n * 11 + 16 + 15) & ~15ull;          // LEB128 <= 10 bytes per value, plus headers
}

}  // namespace

int rlev1_compress_bound(uint64_t in_bytes, uint64_t chunk, uint32_t eb, uint32_t line, uint64_t* n_chunks,
                         uint64_t* work_bytes, uint64_t* payload_cap) {
    if (in_bytes == 0 || chunk % 32 || line == 0 || (chunk / 32) % line || line % eb ||
        (eb != 1 && eb != 2 && eb != 4 && eb != 8) || (chunk / 32) / eb < 2 || chunk > (1ull << 32))
        return MDZ_ERR_INVALID;
    const uint64_t n = (in_bytes + chunk - 1) / chunk;
    const uint64_t slot = lane_slot_bytes(chunk, eb);
    if (n_chunks) *n_chunks = n;
    if (work_bytes) *work_bytes = n * 32 * slot + 8 * n;
    if (payload_cap) *payload_cap = n * ((32 * slot + 127) / 128 * 128);
    return MDZ_OK;
}

int rlev1_compress(const void* in, uint64_t in_bytes, uint64_t chunk, uint32_t eb, uint32_t line, void* work,
                   void* payload, uint64_t* col_len, uint64_t* blk_off, hipStream_t s) {
    uint64_t n = 0;
    int rc = rlev1_compress_bound(in_bytes, chunk, eb, line, &n, nullptr, nullptr);
    if (rc) return rc;
    if (!in || !work || !payload || !col_len || !blk_off || n > (1ull << 24)) return MDZ_ERR_INVALID;
    const uint64_t slot = lane_slot_bytes(chunk, eb);
    uint64_t* chunk_len = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(work) + n * 32 * slot);
    hipLaunchKernelGGL(rlev1_encode_kernel, dim3((uint32_t)n), dim3(kWave), 0, s, static_cast<const uint8_t*>(in),
                       in_bytes, chunk, eb, line, slot, static_cast<uint8_t*>(work), col_len, chunk_len);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return record_hip(e, "rlev1_encode_kernel launch");
    hipLaunchKernelGGL(lanes_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, chunk_len, n, blk_off, nullptr, ~0ull);
    if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "lanes_scan_kernel launch");
    hipLaunchKernelGGL(lanes_interleave_kernel, dim3((uint32_t)n), dim3(kIlvThreads), 0, s,
                       static_cast<const uint8_t*>(work), slot, col_len, blk_off, chunk_len, nullptr,
                       static_cast<uint8_t*>(payload));
    return record_hip(hipGetLastError(), "lanes_interleave_kernel launch");
}

}  // namespace mdz
