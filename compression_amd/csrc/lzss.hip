// LZSS lane-transposed decoder for gfx950, and the lane-transposed byte RLE
// that shares its container and flag-group grammar.
//
// Replaces lzss::kernel_decompress / decompress_func (reference
// include/lzss/lzss_gpu2.h:1898-1907,114-325), which runs one thread per lane
// stream with a 2 KiB history array in scratch memory and 4-byte strided
// stores, and brle_trans::decompress_func (include/brle/brle_trans.h:72-202),
// one thread per lane as well.  A brle_trans 2-byte token is a run (count
// byte, value byte) instead of a match; the kernel is instantiated for each
// meaning (kRle), and a run-length stream's map holds output bytes directly
// (one byte per entry: half the LDS of the 16-bit source map, so more
// workgroups fit a CU), so nothing is resolved.
//
// Container: chunk c holds 32 lane streams; sorted position s has
// col_len[32c+s] bytes and decodes original lane t = col_map[32c+s]; lengths
// are non-increasing in s, so row k of the payload is word k of every sorted
// lane with more than k words, in sorted order.  Lane t's output byte i lands
// at chunk byte 128*(i/4) + 4t + i%4.  Token grammar: flag byte, then 8
// tokens, bit b (LSB first) 1 = literal byte, 0 = match v = 2 bytes LE,
// length (v & 31) + 3, distance v >> 5.
//
// One workgroup decodes G original lanes of one chunk (32/G parts per chunk;
// the parts of a chunk are given block ids that map to the same XCD so the
// payload rows they all read come from one L2):
//   1. wave 0 validates the chunk's metadata and builds suffix sums of the
//      word counts; R[r] = sum_s min(W_s, r) gives row r's first word;
//   2. the G lanes' words are gathered from the payload rows into LDS
//      (each load instruction touches G-word runs of 64/G rows);
//   3. one thread per lane walks the flag-group chain (group length is
//      17 - popcount(flag)), recording every group start -- the only
//      serial step;
//   4. one wave per lane: threads own contiguous groups, sum each group's
//      output length, prefix-sum them across the wave, then write a u16
//      source map per output byte: 0x8000|byte for literals, the source
//      position for match bytes (byte-serial semantics, so overlapping
//      matches repeat their source); for run tokens, a byte map of values;
//   5. each map entry is resolved by following source positions down to a
//      literal (positions strictly decrease) and replaced by its value, which
//      shortens later walks;
//   6. the G lanes' bytes are written back as G*4-byte row segments.
//
// Status: chunk status is that of its lowest original lane with an error
// (the first failing token in that lane's stream order), as the oracle
// reports it.  Parts combine by atomicMin on (lane << 8 | code) after an init
// fill; a finishing pass turns the key back into the status code.
#include <type_traits>

#include "mdz_device.h"
#include "mdz_internal.h"
#include "status_dev.h"

#include <climits>

namespace mdz {
namespace {

constexpr uint32_t kNoErr = 0xffffffffu;

struct LzssGeom {
    uint32_t lane_out;     // output bytes per lane (chunk / 32)
    uint32_t cap_stream;   // longest valid lane stream: all literals + flags + reserved byte
    uint32_t cap_in;       // LDS bytes per lane input (>= cap_stream + 16, multiple of 16)
    uint32_t gmax;         // group-table entries per lane (>= ceil(cap_stream / 9))
    uint32_t rows;         // R-table entries (>= words of the longest lane)
    uint32_t r_bytes;      // R-table bytes, multiple of 16
};

struct LzssHdr {
    uint32_t L[32];        // stream bytes per sorted position
    uint32_t W[32];        // words per sorted position
    uint32_t SW[33];       // SW[s] = sum of W[s..31]
    uint32_t ng[8];        // groups per local lane
    uint32_t bad;
    uint8_t inv[32];       // original lane -> sorted position
    uint32_t pad[2];
};
static_assert(sizeof(LzssHdr) % 16 == 0, "LDS carve alignment");

MDZ_DEV void report(int32_t* status, uint64_t c, uint32_t lane, int code) {
    if (status) atomicMin(&status[c], (int32_t)((lane << 8) | ((uint32_t)code & 0xffu)));
}

// Stores v, v + inc, v + 2*inc, ... into m[i, e) as u16 entries, two per
// 32-bit store (m is 4-byte aligned; entries stay below 0x10000).
MDZ_DEV void fill_u16(uint16_t* m, uint32_t i, uint32_t e, uint32_t v, uint32_t inc) {
    if ((i & 1u) && i < e) {
        m[i++] = (uint16_t)v;
        v += inc;
    }
    uint32_t pk = (v & 0xffffu) | ((v + inc) << 16);
    const uint32_t step = 2 * inc * 0x10001u;
    for (; i + 2 <= e; i += 2, pk += step) *reinterpret_cast<uint32_t*>(m + i) = pk;
    if (i < e) m[i] = (uint16_t)pk;
}

// Byte-map fill: m[i..e) = v (aligned dword stores in the middle).
MDZ_DEV void fill_u8(uint8_t* m, uint32_t i, uint32_t e, uint32_t v) {
    for (; (i & 3u) && i < e; ++i) m[i] = (uint8_t)v;
    const uint32_t pk = (v & 0xffu) * 0x01010101u;
    for (; i + 4 <= e; i += 4) *reinterpret_cast<uint32_t*>(m + i) = pk;
    for (; i < e; ++i) m[i] = (uint8_t)v;
}

// A run-length stream's map holds output bytes (values only, one byte
// each); an LZSS stream's holds 16-bit sources (a literal's byte with bit 15
// set, or an earlier output position) until resolution.
template <bool kRle> using MapT = typename std::conditional<kRle, uint8_t, uint16_t>::type;

// Output bytes of a 2-byte token from its first byte: an LZSS match,
// (b & 31) + 3; a byte-RLE run (brle_trans), the count byte itself.
template <bool kRle>
MDZ_DEV uint32_t tok_len(uint32_t b) {
    return kRle ? b : (b & 31u) + 3u;
}

template <int G, bool kRle>
__global__ __launch_bounds__(64 * G) void lzss_decode_kernel(
    const uint8_t* __restrict__ payload, uint64_t payload_bytes, const uint64_t* __restrict__ blk_off,
    const uint64_t* __restrict__ col_len, const uint8_t* __restrict__ col_map, uint64_t n_chunks,
    uint8_t* __restrict__ dst, uint64_t dst_bytes, int32_t* __restrict__ status, LzssGeom geo) {
    constexpr uint32_t P = 32 / G, NT = 64 * G;
    extern __shared__ __align__(16) uint8_t smem[];
    LzssHdr& h = *reinterpret_cast<LzssHdr*>(smem);
    uint16_t* const R = reinterpret_cast<uint16_t*>(smem + sizeof(LzssHdr));
    uint8_t* const inb = smem + sizeof(LzssHdr) + geo.r_bytes;
    MapT<kRle>* const map = reinterpret_cast<MapT<kRle>*>(inb + G * geo.cap_in);
    uint16_t* const gs = reinterpret_cast<uint16_t*>(map + G * geo.lane_out);
    uint16_t* const go = gs + G * geo.gmax;

    // XCD-aware placement: block ids congruent mod 8 share an XCD; the P parts
    // of chunk c get ids 8*(P*(c/8) + part) + c%8.
    const uint32_t bid = blockIdx.x, xcd = bid & 7u, rr = bid >> 3;
    const uint32_t part = rr % P;
    const uint64_t c = (uint64_t)(rr / P) * 8 + xcd;
    if (c >= n_chunks) return;
    const uint32_t tid = threadIdx.x, ln = tid & 63u, wv = tid >> 6;
    const uint32_t lane_out = geo.lane_out;
    const uint64_t beg = blk_off[c];

    // ---- 1. metadata (wave 0; lanes 32..63 mirror 0..31)
    if (wv == 0) {
        const uint32_t s = ln & 31u;
        const uint64_t Ls = col_len[32 * c + s];
        const uint32_t ms = col_map[32 * c + s];
        const uint64_t Lp = s ? col_len[32 * c + s - 1] : ~0ull;
        bool bad = Ls > geo.cap_stream || ms > 31u || Ls > Lp;
        uint32_t seen = ms < 32u ? 1u << ms : 0u;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) seen |= __shfl_xor(seen, o, 64);
        const uint32_t Ws = bad ? 0u : (uint32_t)((Ls + 3) / 4);
        const uint32_t incl = wave_incl_scan_u32(ln < 32 ? Ws : 0u);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 31);
        const uint64_t end = blk_off[c + 1];
        bad = bad || seen != 0xffffffffu || end < beg || end > payload_bytes || (beg & 3u) ||
              4ull * total > end - beg;
        const bool any_bad = __ballot(bad) != 0;
        if (ln < 32) {
            h.L[s] = (uint32_t)(bad ? 0 : Ls);
            h.W[s] = Ws;
            h.SW[s] = total - (incl - Ws);
            if (ms < 32u) h.inv[ms] = (uint8_t)s;
        }
        if (ln == 0) { h.SW[32] = 0; h.bad = any_bad; }
    }
    __syncthreads();
    if (h.bad) {
        if (tid == 0 && part == 0) report(status, c, 0, MDZ_ST_BAD_HEADER);
        return;
    }

    // ---- 2. row table R[r] = SW[n_r] + r * n_r, n_r = #{s : W_s > r}
    uint32_t mw = 0;
#pragma unroll
    for (int j = 0; j < G; ++j) mw = max(mw, h.W[h.inv[part * G + j]]);
    for (uint32_t r = tid; r < mw; r += NT) {
        uint32_t lo = 0, hi = 32;                  // first s with W_s <= r (W descending)
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (h.W[mid] > r) lo = mid + 1; else hi = mid;
        }
        R[r] = (uint16_t)(h.SW[lo] + r * lo);
    }
    __syncthreads();

    // ---- 3. gather the G lanes' words: thread -> (row, local lane)
    {
        const uint32_t j = tid & (G - 1);
        const uint32_t sj = h.inv[part * G + j], Wj = h.W[sj];
        const uint32_t* src = reinterpret_cast<const uint32_t*>(payload + beg);
        uint32_t* d32 = reinterpret_cast<uint32_t*>(inb + j * geo.cap_in);
        // eight rows per thread in flight: row indices, then the loads, then
        // the LDS stores (one store per iteration would wait out a global
        // load before the next index read)
        constexpr uint32_t kStep = NT / G;
        for (uint32_t r0 = tid / G; r0 < mw; r0 += 8 * kStep) {
            uint32_t at[8], w[8];
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t r = r0 + k * kStep;
                at[k] = r < Wj ? R[r] + sj : 0xffffffffu;
            }
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) w[k] = at[k] != 0xffffffffu ? src[at[k]] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t r = r0 + k * kStep;
                if (r < Wj) d32[r] = w[k];
            }
        }
    }
    __syncthreads();

    // ---- 4. group-chain walk, one thread per local lane
    if (wv == 0 && ln < G) {
        const uint32_t sj = h.inv[part * G + ln], L = h.L[sj];
        const uint8_t* d = inb + ln * geo.cap_in;
        uint16_t* g = gs + ln * geo.gmax;
        uint32_t p = 0, n = 0;
        while (p < L) {                            // L <= cap_stream bounds n by gmax
            g[n++] = (uint16_t)p;
            p += 17u - __popc(d[p]);
        }
        h.ng[ln] = n;
    }
    __syncthreads();

    // ---- 5. one wave per lane: lengths, offsets, source map, resolution
    {
        const uint32_t j = wv, t = part * G + j;
        const uint32_t L = h.L[h.inv[t]];
        const uint8_t* d = inb + j * geo.cap_in;
        MapT<kRle>* mp = map + j * lane_out;
        const uint16_t* gj = gs + j * geo.gmax;
        uint16_t* oj = go + j * geo.gmax;
        const uint32_t ng = h.ng[j];
        const uint32_t k = (ng + 63) >> 6, g0 = min(ng, ln * k), g1 = min(ng, g0 + k);

        // pass A: output bytes of each group's tokens that lie inside the stream
        uint32_t sum = 0;
        for (uint32_t g = g0; g < g1; ++g) {
            const uint32_t p = gj[g], f = d[p];
            uint32_t ol;
            if (p + 17u - __popc(f) <= L) {
                ol = __popc(f);
                const uint32_t nf = ~f & 0xffu;
                for (uint32_t m = nf; m; m &= m - 1) {
                    const uint32_t b = __builtin_ctz(m);
                    ol += tok_len<kRle>(d[p + 1 + b + __popc(nf & ((1u << b) - 1))]);
                }
            } else {                               // stream ends inside this group
                ol = 0;
                uint32_t q = p + 1;
                for (uint32_t b = 0; b < 8; ++b) {
                    const uint32_t lit = (f >> b) & 1u, need = 2u - lit;
                    if (q + need > L) break;
                    ol += lit ? 1u : tok_len<kRle>(d[q]);
                    q += need;
                }
            }
            oj[g] = (uint16_t)min(ol, 0xffffu);
            sum += ol;
        }
        const uint32_t incl = wave_incl_scan_u32(sum);
        const uint32_t T = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        uint32_t run = incl - sum;
        for (uint32_t g = g0; g < g1; ++g) {
            const uint32_t ol = oj[g];
            oj[g] = (uint16_t)min(run, 0xffffu);
            run += ol;
        }
        wave_sync();

        // pass B: source map; first error of this thread's tokens in stream order
        uint32_t err = kNoErr;
        for (uint32_t g = g0; g < g1; ++g) {
            uint32_t o = oj[g];
            if (o >= lane_out) break;
            const uint32_t p = gj[g], f = d[p];
            uint32_t q = p + 1;
            bool stop = false;
            for (uint32_t b = 0; b < 8 && o < lane_out; ++b) {
                const uint32_t tok = g * 8 + b;
                if ((f >> b) & 1u) {
                    if (q + 1 > L) {
                        err = min(err, (tok << 8) | ((uint32_t)MDZ_ST_INPUT_OVERRUN & 0xffu));
                        stop = true;
                        break;
                    }
                    if constexpr (kRle) mp[o++] = d[q++];
                    else mp[o++] = (uint16_t)(0x8000u | d[q++]);
                } else {
                    if (q + 2 > L) {
                        err = min(err, (tok << 8) | ((uint32_t)MDZ_ST_INPUT_OVERRUN & 0xffu));
                        stop = true;
                        break;
                    }
                    const uint32_t v = (uint32_t)d[q] | ((uint32_t)d[q + 1] << 8);
                    q += 2;
                    const uint32_t len = tok_len<kRle>(v & 0xffu), dist = v >> 5;
                    const uint32_t e = min(o + len, lane_out);
                    if constexpr (kRle) {          // run: `len` copies of the value byte
                        fill_u8(mp, o, e, v >> 8);
                        if (o + len > lane_out)
                            err = min(err, (tok << 8) | ((uint32_t)MDZ_ST_OUTPUT_OVERRUN & 0xffu));
                    } else if (dist == 0 || dist > o) {
                        err = min(err, (tok << 8) | ((uint32_t)MDZ_ST_BAD_DISTANCE & 0xffu));
                        fill_u16(mp, o, e, 0x8000u, 0);
                    } else {
                        fill_u16(mp, o, e, o - dist, 1);   // entry i holds i - dist
                        if (o + len > lane_out)
                            err = min(err, (tok << 8) | ((uint32_t)MDZ_ST_OUTPUT_OVERRUN & 0xffu));
                    }
                    o += len;
                }
            }
            if (stop) break;
        }
        if (T < lane_out) {                        // stream ended early: fill the rest
            for (uint32_t i = T + ln; i < lane_out; i += 64) mp[i] = kRle ? 0u : 0x8000u;
            err = min(err, ((ng * 8) << 8) | ((uint32_t)MDZ_ST_INPUT_OVERRUN & 0xffu));
        }
        wave_sync();

        // resolve: follow sources down to a literal, store the value back
        // (a run-length stream's map holds values only)
        // (both chains of a pair in flight; a resolved chain re-reads mp[0]
        // instead of branching)
        for (uint32_t i = 2 * ln; !kRle && i < lane_out; i += 128) {
            uint32_t* w = reinterpret_cast<uint32_t*>(reinterpret_cast<uint16_t*>(mp) + i);
            uint32_t e0 = *w & 0xffffu, e1 = *w >> 16;
            if ((e0 & e1 & 0x8000u) == 0) {
                do {
                    const uint32_t v0 = mp[(e0 & 0x8000u) ? 0u : e0], v1 = mp[(e1 & 0x8000u) ? 0u : e1];
                    e0 = (e0 & 0x8000u) ? e0 : v0;
                    e1 = (e1 & 0x8000u) ? e1 : v1;
                } while ((e0 & e1 & 0x8000u) == 0);
                *w = e0 | (e1 << 16);
            }
        }
        const uint32_t lane_err = wave_min_u32(err);
        if (ln == 0 && lane_err != kNoErr) report(status, c, t, (int)(int8_t)(lane_err & 0xffu));
    }
    __syncthreads();

    // ---- 6. write G*4-byte row segments
    {
        const uint64_t out0 = c * (uint64_t)lane_out * 32;
        const uint32_t nrows = lane_out / 4;
        for (uint32_t idx = tid; idx < nrows * G; idx += NT) {
            const uint32_t j = idx & (G - 1), r = idx / G;
            uint32_t word;
            if constexpr (kRle) {
                word = *reinterpret_cast<const uint32_t*>(map + j * lane_out + 4 * r);
            } else {
                const uint64_t e = *reinterpret_cast<const uint64_t*>(map + j * lane_out + 4 * r);
                word = (uint32_t)(e & 0xff) | (uint32_t)((e >> 8) & 0xff00) | (uint32_t)((e >> 16) & 0xff0000) |
                       (uint32_t)((e >> 24) & 0xff000000u);
            }
            const uint64_t at = out0 + 128ull * r + 4ull * (part * G + j);
            if (at + 4 <= dst_bytes) {
                *reinterpret_cast<uint32_t*>(dst + at) = word;
            } else {
                for (uint32_t b = 0; b < 4; ++b)
                    if (at + b < dst_bytes) dst[at + b] = (uint8_t)(word >> (8 * b));
            }
        }
    }
}

template <int G, bool kRle>
int launch(const mdz_lanes_t* in, const LzssGeom& geo, void* dst, uint64_t dst_bytes,
           int32_t* status, hipStream_t s) {
    const size_t smem = sizeof(LzssHdr) + geo.r_bytes +
                        (size_t)G * (geo.cap_in + sizeof(MapT<kRle>) * geo.lane_out + 4ull * geo.gmax);
    hipError_t e = set_dynamic_lds((const void*)lzss_decode_kernel<G, kRle>, smem);
    if (e != hipSuccess) return record_hip(e, "lzss_decode_kernel LDS");
    const uint64_t blocks = (in->n_chunks + 7) / 8 * 8 * (32 / G);
    hipLaunchKernelGGL((lzss_decode_kernel<G, kRle>), dim3((uint32_t)blocks), dim3(64 * G), smem, s,
                       (const uint8_t*)in->payload, in->payload_bytes, in->blk_off, in->col_len,
                       in->col_map, in->n_chunks, (uint8_t*)dst, dst_bytes, status, geo);
    return record_hip(hipGetLastError(), "lzss_decode_kernel launch");
}

template <bool kRle>
int lanes_flag_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                          hipStream_t s) {
    if (!in || !in->payload || !in->blk_off || !in->col_len || !in->col_map || (!dst && dst_bytes))
        return MDZ_ERR_INVALID;
    const uint64_t chunk = in->chunk_bytes;
    if (chunk == 0 || chunk % 128) return MDZ_ERR_INVALID;
    if (chunk > 128 * 1024) return MDZ_ERR_UNSUPPORTED;       // lane output > 4 KiB
    const uint64_t n = in->n_chunks;
    if (n == 0) return MDZ_OK;
    if (n > (1ull << 27)) return MDZ_ERR_UNSUPPORTED;
    LzssGeom geo;
    geo.lane_out = (uint32_t)(chunk / 32);
    geo.cap_stream = geo.lane_out + (geo.lane_out + 7) / 8 + 1;
    geo.cap_in = (geo.cap_stream + 15) / 16 * 16 + 16;
    geo.gmax = (geo.cap_stream / 9 + 3) & ~1u;
    geo.rows = geo.cap_in / 4;
    geo.r_bytes = (2 * geo.rows + 15) / 16 * 16;
    if (status) {
        hipLaunchKernelGGL(status_fill_kernel, dim3((uint32_t)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096)),
                           dim3(256), 0, s, status, n);
        int rc = record_hip(hipGetLastError(), "lzss status init");
        if (rc) return rc;
    }
    int rc = geo.lane_out <= 1024 ? launch<8, kRle>(in, geo, dst, dst_bytes, status, s)
           : geo.lane_out <= 2048 ? launch<4, kRle>(in, geo, dst, dst_bytes, status, s)
                                  : launch<2, kRle>(in, geo, dst, dst_bytes, status, s);
    if (rc || !status) return rc;
    hipLaunchKernelGGL(status_finish_kernel, dim3((uint32_t)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096)),
                       dim3(256), 0, s, status, n);
    return record_hip(hipGetLastError(), "lzss status finish");
}
}  // namespace

int lzss_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                    hipStream_t s) {
    return lanes_flag_decompress<false>(in, dst, dst_bytes, status, s);
}

int brle_trans_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                          hipStream_t s) {
    return lanes_flag_decompress<true>(in, dst, dst_bytes, status, s);
}

}  // namespace mdz
