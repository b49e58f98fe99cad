// Plain (per-chunk) LZSS decoder for gfx950.
//
// Replaces lzss::kernel_decompress(in, out, in_n_bytes, out_n_bytes,
// out_chunk_size, n_chunks, lens) and its plain decompress_func (reference
// include/lzss/lzss_gpu2.h:1908-1913,718-792), which runs one thread per chunk
// through the whole token stream.  Chunk c's stream is data[lens[c-1],
// lens[c]) (lens = inclusive prefix sums of the stream lengths, written by
// compress_func :1450-1566 and the scan that follows it).  Token grammar:
// flag byte, then 8 tokens, bit b (LSB first) 1 = literal byte, 0 = match
// v = 2 bytes LE, length (v & 31) + 3, distance v >> 5 (1..2047); a group
// spans 17 - popcount(flag) bytes.
//
// One 512-thread workgroup per chunk.  The stream is handled in pieces of
// kPiece bytes staged in LDS; for each piece:
//   1. exit tables: for every 64-byte segment and each of the 17 offsets at
//      which the segment's first group can start (groups span 9..17 bytes),
//      the offset past the segment end where that chain of groups leaves it
//      -- walks of at most 8 steps, all independent;
//   2. wave 0 composes the tables across the segments from the known entry
//      of the first one (one v_readlane per segment), giving every segment
//      its true first group and the next piece its entry;
//   3. one thread per segment sums its groups' output bytes, a block scan
//      gives each group its output offset, and the group list is written;
//   4. the output is built in windows of kWin bytes preceded by kHist bytes of
//      history: a u16 source map per byte (kLit|value for literals and
//      resolved bytes, else the map index of the byte it copies) is filled
//      with 8 lanes per group, one per token; a match that overlaps itself
//      (distance < length) points into its first period, so runs do not form
//      byte-by-byte chains.  Every entry is then resolved by following sources
//      down to a value (sources strictly precede their byte) and the value is
//      stored back, which shortens later walks;
//   5. a finished window is written out as 16-byte stores and its last kHist
//      bytes become the next window's history.
// Status: the first failing token in stream order -- the lowest output
// offset among bad distances and matches that run past the chunk end, or the
// output offset at which the stream runs out -- as a byte-serial decoder
// that stops when the chunk's output is complete reports it.
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

constexpr uint32_t kNT = 512;                   // threads per workgroup (8 waves)
constexpr uint32_t kSeg = 64;                   // stream bytes per segment
constexpr uint32_t kPiece = 16384;              // stream bytes whose groups one pass handles
constexpr uint32_t kNSeg = kPiece / kSeg;
constexpr uint32_t kStage = kPiece + 64;        // a group reaches at most 16 bytes past its start
constexpr uint32_t kGCap = (kPiece - 1) / 9 + 2;  // groups that can start in one piece, + sentinel
constexpr uint32_t kHist = 2048;                // distances are 11 bits
constexpr uint32_t kWin = 16384;                // output bytes per window
constexpr uint32_t kLit = 0x8000u;              // map entry holds a value
constexpr uint32_t kNoErr = 0xffffffffu;
constexpr uint64_t kMaxChunk = 1ull << 28;      // error keys are (offset << 2 | kind) in 32 bits

static_assert(kHist + kWin <= kLit, "map indices must stay below the value flag");
static_assert(kPiece / 2 * 34 < (1u << 20), "piece output must fit the 20-bit scan field");

struct PlainLds {
    uint16_t map[kHist + kWin];  // [0, kHist): history; [kHist, ...): the window
    uint32_t gout[kGCap];        // output offset (chunk-relative) of each group of the piece
    uint16_t gpos[kGCap];        // stream offset (piece-relative) of each group
    uint32_t wtot[kNT / 64];
    uint8_t stage[kStage];
    uint8_t ex[kNSeg * 17];
    uint8_t entry[kNSeg];
    uint32_t err;                // min over (output offset << 2 | kind)
    uint32_t next_entry;         // first group's offset in the next piece
};
static_assert(sizeof(PlainLds) <= 80 * 1024, "two workgroups per CU");

// Output bytes of the tokens of the group at stage[p] (flag f) that lie
// wholly inside the first `avail` staged bytes of the stream.
MDZ_DEV uint32_t group_out(const uint8_t* st, uint32_t p, uint32_t f, uint32_t avail) {
    uint32_t ol = 0;
    if (p + 17u - __popc(f) <= avail) {
        ol = __popc(f);
        const uint32_t nf = ~f & 0xffu;
        for (uint32_t m = nf; m; m &= m - 1) {
            const uint32_t b = __builtin_ctz(m);
            ol += (st[p + 1 + b + __popc(nf & ((1u << b) - 1))] & 31u) + 3u;
        }
    } else {                                  // the stream ends inside this group
        uint32_t q = p + 1;
        for (uint32_t b = 0; b < 8; ++b) {
            const uint32_t lit = (f >> b) & 1u, need = 2u - lit;
            if (q + need > avail) break;
            ol += lit ? 1u : (st[q] & 31u) + 3u;
            q += need;
        }
    }
    return ol;
}

__global__ __launch_bounds__(kNT) void lzss_plain_kernel(
    const uint8_t* __restrict__ data, uint64_t data_bytes, const uint64_t* __restrict__ lens,
    uint64_t n_chunks, uint64_t chunk, uint8_t* __restrict__ dst, uint64_t dst_bytes,
    int32_t* __restrict__ status) {
    extern __shared__ __align__(16) uint8_t smem[];
    PlainLds& S = *reinterpret_cast<PlainLds*>(smem);
    const uint64_t c = blockIdx.x;
    if (c >= n_chunks) return;
    const uint32_t tid = threadIdx.x, ln = tid & 63u, wv = tid >> 6;
    const uint64_t beg = c ? lens[c - 1] : 0, end = lens[c];
    const uint64_t o0 = c * chunk;
    const uint32_t n = o0 >= dst_bytes ? 0u : (uint32_t)(dst_bytes - o0 < chunk ? dst_bytes - o0 : chunk);
    if (end < beg || end > data_bytes) {
        if (tid == 0 && status) status[c] = MDZ_ST_BAD_HEADER;
        return;
    }
    const uint64_t L = end - beg;
    const uint8_t* const src = data + beg;
    uint8_t* const out = dst + o0;

    for (uint32_t i = tid; i < kHist; i += kNT) S.map[i] = (uint16_t)kLit;
    if (tid == 0) S.err = kNoErr;

    uint32_t produced = 0;   // output offset at which the current piece's groups start
    uint32_t entry0 = 0;     // offset of the current piece's first group
    uint32_t win = 0;        // output offset of the window the map holds
    uint32_t my_err = kNoErr;
    for (uint64_t pb = 0; pb < L && produced < n; pb += kPiece) {
        const uint32_t plen = (uint32_t)(L - pb < kPiece ? L - pb : kPiece);
        const uint32_t avail = (uint32_t)(L - pb < kStage ? L - pb : kStage);
        const uint32_t nseg = (plen + kSeg - 1) / kSeg;

        // ---- stage stream bytes [pb, pb + avail) as words; loads are aligned
        // words whose first byte lies inside the stream
        {
            const uintptr_t ga = reinterpret_cast<uintptr_t>(src + pb);
            const uint32_t mis = (uint32_t)(ga & 3u);
            const uint32_t* aw = reinterpret_cast<const uint32_t*>(ga - mis);
            const uint32_t nwords = (avail + mis + 3) / 4;
            uint32_t* st32 = reinterpret_cast<uint32_t*>(S.stage);
            for (uint32_t k = tid; k < kStage / 4; k += kNT) {
                const uint32_t lo = k < nwords ? aw[k] : 0u;
                const uint32_t hi = k + 1 < nwords ? aw[k + 1] : 0u;
                uint32_t w = alignbyte(hi, lo, mis);
                const uint32_t b0 = 4 * k;
                if (b0 + 4 > avail) w = b0 >= avail ? 0u : w & ((1u << (8 * (avail - b0))) - 1u);
                st32[k] = w;
            }
        }
        __syncthreads();

        // ---- 1. exit tables
        for (uint32_t w = tid; w < nseg * 17; w += kNT) {
            const uint32_t s = w / 17, e = w - s * 17;
            const uint32_t lim = (s + 1) * kSeg;
            uint32_t p = s * kSeg + e;
            while (p < lim) p += 17u - __popc(S.stage[p]);
            S.ex[w] = (uint8_t)(p - lim);
        }
        __syncthreads();

        // ---- 2. compose across segments (wave 0; the chain value is uniform)
        if (wv == 0) {
            uint32_t cur = __builtin_amdgcn_readfirstlane(entry0);
            for (uint32_t s = 0; s < nseg; s += 8) {
                uint32_t row[8];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    row[j] = (s + j < nseg && ln < 17) ? S.ex[(s + j) * 17 + ln] : 0u;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (s + j < nseg) {
                        if (ln == 0) S.entry[s + j] = (uint8_t)cur;
                        cur = (uint32_t)__builtin_amdgcn_readlane((int)row[j], (int)cur);
                    }
                }
            }
            if (ln == 0) S.next_entry = cur;
        }
        __syncthreads();
        const uint32_t next_entry = S.next_entry;

        // ---- 3. groups: per-segment counts and output bytes, scan, list
        uint32_t v = 0;
        if (tid < nseg) {
            const uint32_t lim = min((tid + 1) * kSeg, plen);
            uint32_t cnt = 0, sum = 0;
            for (uint32_t p = tid * kSeg + S.entry[tid]; p < lim;) {
                const uint32_t f = S.stage[p];
                sum += group_out(S.stage, p, f, avail);
                ++cnt;
                p += 17u - __popc(f);
            }
            v = (cnt << 20) | sum;
        }
        const uint32_t incl = wave_incl_scan_u32(v);
        if (ln == 63) S.wtot[wv] = incl;
        __syncthreads();
        uint32_t base = 0, total = 0;
#pragma unroll
        for (uint32_t k = 0; k < kNT / 64; ++k) {
            const uint32_t t = S.wtot[k];
            base += k < wv ? t : 0u;
            total += t;
        }
        if (tid < nseg) {
            const uint32_t ex = base + incl - v;
            const uint32_t lim = min((tid + 1) * kSeg, plen);
            uint32_t gi = ex >> 20, o = produced + (ex & 0xfffffu);
            for (uint32_t p = tid * kSeg + S.entry[tid]; p < lim;) {
                const uint32_t f = S.stage[p];
                S.gpos[gi] = (uint16_t)p;
                S.gout[gi] = o;
                o += group_out(S.stage, p, f, avail);
                ++gi;
                p += 17u - __popc(f);
            }
        }
        const uint32_t ng = total >> 20, pout = total & 0xfffffu;
        if (tid == 0) S.gout[ng] = produced + pout;
        __syncthreads();

        // ---- 4./5. windows
        const uint32_t pend = min(produced + pout, n);
        for (uint32_t x0 = produced; x0 < pend;) {
            const uint32_t wend = min(win + kWin, pend);
            // groups [glo, ghi) overlap [x0, wend): glo = last with gout <= x0
            uint32_t lo = 0, hi = ng;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (S.gout[mid] <= x0) lo = mid; else hi = mid;
            }
            const uint32_t glo = lo;
            lo = glo; hi = ng;                     // ghi = first with gout >= wend
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (S.gout[mid] < wend) lo = mid + 1; else hi = mid;
            }
            const uint32_t ghi = lo;
            const uint32_t mb = kHist - win;       // map index of output offset x: x + mb
            for (uint32_t g = glo + (tid >> 3); g < ghi; g += kNT / 8) {
                const uint32_t k = tid & 7u;
                const uint32_t gp = S.gpos[g], go = S.gout[g];
                const uint32_t f = S.stage[gp];
                const uint32_t lit = (f >> k) & 1u;
                const uint32_t q = gp + 1 + k + __popc(~f & ((1u << k) - 1u) & 0xffu);
                const uint32_t b0 = S.stage[q], b1 = S.stage[q + 1];
                const uint32_t tl = q + 2 - lit <= avail ? (lit ? 1u : (b0 & 31u) + 3u) : 0u;
                uint32_t inc = tl;
#pragma unroll
                for (uint32_t o = 1; o < 8; o <<= 1) {
                    const uint32_t y = __shfl_up(inc, o, 8);
                    inc += k >= o ? y : 0u;
                }
                const uint32_t off = go + inc - tl;
                const uint32_t xa = max(off, x0), xb = min(off + tl, wend);
                if (lit) {
                    if (xa < xb) S.map[off + mb] = (uint16_t)(kLit | b0);
                } else if (tl) {
                    const uint32_t d = (b0 | (b1 << 8)) >> 5;
                    const bool bad = d == 0 || d > off;
                    if (off >= x0 && off < wend) {
                        if (bad) my_err = min(my_err, off << 2);
                        else if (off + tl > n) my_err = min(my_err, (off << 2) | 1u);
                    }
                    if (bad) {
                        for (uint32_t x = xa; x < xb; ++x) S.map[x + mb] = (uint16_t)kLit;
                    } else {
                        uint32_t r = xa - off;
                        if (r >= d) r %= d;
                        const uint32_t sb = off - d + mb;
                        for (uint32_t x = xa; x < xb; ++x) {
                            S.map[x + mb] = (uint16_t)(sb + r);
                            r = r + 1 == d ? 0u : r + 1;
                        }
                    }
                }
            }
            __syncthreads();

            // resolve [x0, wend)
            for (uint32_t x = x0 + tid; x < wend; x += kNT) {
                uint32_t e = S.map[x + mb];
                if (!(e & kLit)) {
                    do { e = S.map[e]; } while (!(e & kLit));
                    S.map[x + mb] = (uint16_t)e;
                }
            }
            __syncthreads();

            const bool full = wend == win + kWin;
            if (full || wend == n) {
                const uint32_t cnt = wend - win;
                uint8_t* ob = out + win;
                const bool al = (reinterpret_cast<uintptr_t>(ob) & 15u) == 0;
                for (uint32_t u = tid; u * 16 < cnt; u += kNT) {
                    const uint32_t b = u * 16;
                    const u32x4 m0 = *reinterpret_cast<const u32x4*>(&S.map[kHist + b]);
                    const u32x4 m1 = *reinterpret_cast<const u32x4*>(&S.map[kHist + b + 8]);
                    u32x4 wv4;
                    wv4.x = __builtin_amdgcn_perm(m0.y, m0.x, 0x06040200u);
                    wv4.y = __builtin_amdgcn_perm(m0.w, m0.z, 0x06040200u);
                    wv4.z = __builtin_amdgcn_perm(m1.y, m1.x, 0x06040200u);
                    wv4.w = __builtin_amdgcn_perm(m1.w, m1.z, 0x06040200u);
                    if (al && b + 16 <= cnt) {
                        *reinterpret_cast<u32x4*>(ob + b) = wv4;
                    } else {
                        const uint32_t m = min(16u, cnt - b);
                        for (uint32_t j = 0; j < m; ++j) ob[b + j] = (uint8_t)(wv4[j >> 2] >> (8 * (j & 3)));
                    }
                }
                if (full) {                         // last kHist bytes become the history
                    for (uint32_t i = tid; i < kHist / 2; i += kNT)
                        reinterpret_cast<uint32_t*>(S.map)[i] =
                            reinterpret_cast<const uint32_t*>(S.map + kWin)[i];
                    win += kWin;
                }
                __syncthreads();
            }
            x0 = wend;
        }
        produced += pout;
        entry0 = next_entry;
    }

    if (my_err != kNoErr) atomicMin(&S.err, my_err);
    __syncthreads();
    if (tid == 0 && status) {
        uint32_t key = S.err;
        if (produced < n) key = min(key, (produced << 2) | 2u);
        const int32_t codes[3] = {MDZ_ST_BAD_DISTANCE, MDZ_ST_OUTPUT_OVERRUN, MDZ_ST_INPUT_OVERRUN};
        status[c] = key == kNoErr ? MDZ_ST_OK : codes[key & 3u];
    }
}

}  // namespace

int lzss_plain_decompress(const void* data, uint64_t data_bytes, const uint64_t* lens, uint64_t n_chunks,
                          uint64_t chunk, void* dst, uint64_t dst_bytes, int32_t* status, hipStream_t s) {
    if (n_chunks == 0) return MDZ_OK;
    if (!data || !lens || (!dst && dst_bytes) || chunk == 0)
        return set_error(MDZ_ERR_INVALID, "lzss plain: null buffer or zero chunk size");
    if (chunk > kMaxChunk)
        return set_error(MDZ_ERR_UNSUPPORTED, "lzss plain: chunk of %llu bytes above %llu",
                         (unsigned long long)chunk, (unsigned long long)kMaxChunk);
    if (n_chunks > 0x7fffffffull)
        return set_error(MDZ_ERR_UNSUPPORTED, "lzss plain: %llu chunks", (unsigned long long)n_chunks);
    const size_t smem = sizeof(PlainLds);
    hipError_t e = set_dynamic_lds((const void*)lzss_plain_kernel, smem);
    if (e != hipSuccess) return record_hip(e, "lzss_plain_kernel LDS");
    hipLaunchKernelGGL(lzss_plain_kernel, dim3((uint32_t)n_chunks), dim3(kNT), smem, s,
                       (const uint8_t*)data, data_bytes, lens, n_chunks, chunk, (uint8_t*)dst, dst_bytes,
                       status);
    return record_hip(hipGetLastError(), "lzss_plain_kernel launch");
}

}  // namespace mdz
