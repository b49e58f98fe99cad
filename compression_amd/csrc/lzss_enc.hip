// LZSS encode for gfx950 (lane-transposed container).
//
// Replaces lzss::compress_gpu (reference include/lzss/lzss_gpu2.h:1952-2123:
// compress_init_func / compress_func_new per lane :332-494,511-707 with the
// window scan find_match :70-109, the lane sort :1063-1085 and the
// interleave :497-509).  Payload, col_len, col_map and blk_off equal that
// encoder's output byte for byte.
//
// One wave per chunk; lane t < 32 encodes lane stream t (chunk bytes
// 128 k + 4 t + j, j < 4) greedily: at each position the longest match of
// 3..34 bytes within the lane's last 2048 bytes, ties to the oldest, else a
// literal; tokens in groups of 8 behind a flag byte.  Matches are found
// through hash chains over 3-byte prefixes: chain heads (u16) in LDS, chain
// links in the work buffer; every candidate's length is measured, so the
// choice equals the oldest-first window scan's.  The wave then ranks the 32
// lanes by length (descending, ties to the higher lane id) into col_map, and
// the shared kernels scan the chunk lengths and interleave the sorted lanes.
#include "lanes_enc_dev.h"
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

constexpr uint32_t kHist = 2048, kMinLen = 3, kMaxLen = 34, kLenBits = 5;
constexpr uint32_t kHashBits = 8, kNoPos = 0xFFFF;

struct LzEncLds {
    uint16_t head[32][1u << kHashBits];
    uint64_t len[32];
};

struct LaneSrc {
    const uint8_t* chunk;
    uint32_t t;
    uint32_t nd;                                 // lane dwords (lane bytes / 4)
    MDZ_DEV uint32_t at(uint32_t i) const { return chunk[128u * (i >> 2) + 4u * t + (i & 3u)]; }
    // lane bytes i..i+3 (zero past the lane's end): two aligned dwords of the
    // lane's rows, shifted together
    MDZ_DEV uint32_t dword(uint32_t k) const {
        return k < nd ? *reinterpret_cast<const uint32_t*>(chunk + 128u * k + 4u * t) : 0u;
    }
    MDZ_DEV uint32_t word_at(uint32_t i) const {
        const uint32_t k = i >> 2, sh = i & 3u, lo = dword(k);
        return sh ? alignbyte(dword(k + 1), lo, sh) : lo;
    }
    MDZ_DEV uint32_t hash3(uint32_t i) const {          // lane bytes i, i+1, i+2
        const uint32_t v = word_at(i) & 0xFFFFFFu;
        return (v * 2654435761u) >> (32 - kHashBits);
    }
};

MDZ_DEV uint64_t encode_lzss_lane(const LaneSrc& s, uint32_t n, uint16_t* head, uint16_t* prev, uint8_t* out) {
    for (uint32_t h = 0; h < (1u << kHashBits); ++h) head[h] = (uint16_t)kNoPos;
    auto insert = [&](uint32_t p) {
        if (p + kMinLen > n) return;
        const uint32_t h = s.hash3(p);
        prev[p] = head[h];
        head[h] = (uint16_t)p;
    };
    uint64_t o = 1, flag_at = 0;
    uint32_t ntok = 0, flags = 0;
    out[0] = 0;
    for (uint32_t cur = 0; cur < n;) {
        // the longest match of the window, ties to the oldest
        uint32_t best = 0, best_q = 0;
        if (n - cur >= kMinLen && cur >= 2) {
            const uint32_t count = cur < kHist ? cur : kHist;
            const uint32_t lo = cur - count + 1;
            const uint32_t look = n - cur;
            for (uint32_t q = head[s.hash3(cur)]; q != kNoPos && q >= lo; q = prev[q]) {
                uint32_t cap = min(cur - q, look);
                cap = cap < kMaxLen ? cap : kMaxLen;
                uint32_t f = 0;                  // common prefix, four bytes at a time
                while (f < cap) {
                    const uint32_t x = s.word_at(q + f) ^ s.word_at(cur + f);
                    if (x) {
                        f += (uint32_t)__builtin_ctz(x) >> 3;
                        break;
                    }
                    f += 4;
                }
                f = f < cap ? f : cap;
                if (f >= best) {
                    best = f;
                    best_q = q;
                }
            }
        }
        if (best < kMinLen) {
            flags |= 1u << ntok;
            out[o++] = (uint8_t)s.at(cur);
            insert(cur);
            cur += 1;
        } else {
            const uint32_t v = ((cur - best_q) << kLenBits) | (best - kMinLen);
            out[o] = (uint8_t)v;
            out[o + 1] = (uint8_t)(v >> 8);
            o += 2;
            for (uint32_t i = 0; i < best; ++i) insert(cur + i);
            cur += best;
        }
        if (++ntok == 8) {                      // the next group's flag byte is reserved at once
            out[flag_at] = (uint8_t)flags;
            flags = 0;
            ntok = 0;
            flag_at = o++;
            out[flag_at] = 0;
        }
    }
    if (ntok) out[flag_at] = (uint8_t)flags;
    return o;
}

__global__ __launch_bounds__(kWave) void lzss_encode_kernel(const uint8_t* __restrict__ in, uint64_t chunk,
                                                            uint64_t lane_slot, uint8_t* __restrict__ work,
                                                            uint16_t* __restrict__ links,
                                                            uint64_t* __restrict__ col_len,
                                                            uint8_t* __restrict__ col_map,
                                                            uint64_t* __restrict__ chunk_len) {
    __shared__ LzEncLds S;
    const uint32_t l = lane_id();
    const uint64_t c = blockIdx.x;
    const uint32_t lb = (uint32_t)(chunk / 32);
    if (l < 32) {
        const LaneSrc src{in + c * chunk, l, lb / 4};
        S.len[l] = encode_lzss_lane(src, lb, S.head[l], links + (c * 32 + l) * lb,
                                    work + (c * 32 + l) * lane_slot);
    }
    wave_sync();
    // sorted slot of lane t: longer lanes first, equal lengths by higher lane id
    uint32_t words = 0;
    if (l < 32) {
        const uint64_t mine = S.len[l];
        uint32_t rank = 0;
        for (uint32_t u = 0; u < 32; ++u) {
            const uint64_t other = S.len[u];
            rank += (other > mine || (other == mine && u > l)) ? 1u : 0u;
        }
        col_len[c * 32 + rank] = mine;
        col_map[c * 32 + rank] = (uint8_t)l;
        words = (uint32_t)((mine + 3) / 4);
    }
    const uint32_t total = wave_sum_u32(words);
    if (l == 0) chunk_len[c] = ((uint64_t)total * 4 + 127) / 128 * 128;
}

uint64_t lz_lane_slot(uint64_t chunk) {
    const uint64_t lb = chunk / 32;
    return (lb + lb / 8 + 16 + 15) & ~15ull;          // all literals: one flag byte per 8
}

}  // namespace

int lzss_compress_bound(uint64_t in_bytes, uint64_t chunk, uint64_t* n_chunks, uint64_t* work_bytes,
                        uint64_t* payload_cap) {
    // whole chunks only (a trailing partial chunk is dropped, as the reference
    // encoder drops it); lane positions fit the u16 chains
    if (chunk == 0 || chunk % 128 || chunk > (1ull << 20) || in_bytes < chunk) return MDZ_ERR_INVALID;
    const uint64_t n = in_bytes / chunk;
    const uint64_t slot = lz_lane_slot(chunk);
    if (n_chunks) *n_chunks = n;
    if (work_bytes) *work_bytes = n * 32 * slot + n * chunk * 2 + 8 * n;
    if (payload_cap) *payload_cap = n * ((32 * slot + 127) / 128 * 128);
    return MDZ_OK;
}

int lzss_compress(const void* in, uint64_t in_bytes, uint64_t chunk, void* work, void* payload, uint64_t* col_len,
                  uint8_t* col_map, uint64_t* blk_off, hipStream_t s) {
    uint64_t n = 0;
    int rc = lzss_compress_bound(in_bytes, chunk, &n, nullptr, nullptr);
    if (rc) return rc;
    if (!in || !work || !payload || !col_len || !col_map || !blk_off || n > (1ull << 24)) return MDZ_ERR_INVALID;
    if ((uintptr_t)in & 3) return MDZ_ERR_INVALID;   // lanes are read a dword at a time
    const uint64_t slot = lz_lane_slot(chunk);
    uint8_t* w = static_cast<uint8_t*>(work);
    uint16_t* links = reinterpret_cast<uint16_t*>(w + n * 32 * slot);
    uint64_t* chunk_len = reinterpret_cast<uint64_t*>(w + n * 32 * slot + n * chunk * 2);
    hipLaunchKernelGGL(lzss_encode_kernel, dim3((uint32_t)n), dim3(kWave), 0, s, static_cast<const uint8_t*>(in),
                       chunk, slot, w, links, col_len, col_map, chunk_len);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return record_hip(e, "lzss_encode_kernel launch");
    hipLaunchKernelGGL(lanes_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, chunk_len, n, blk_off, nullptr, ~0ull);
    if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "lanes_scan_kernel launch");
    hipLaunchKernelGGL(lanes_interleave_kernel, dim3((uint32_t)n), dim3(kIlvThreads), 0, s,
                       static_cast<const uint8_t*>(w), slot, col_len, blk_off, chunk_len, col_map,
                       static_cast<uint8_t*>(payload));
    return record_hip(hipGetLastError(), "lanes_interleave_kernel launch");
}

}  // namespace mdz
