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
constexpr uint32_t kPiece = 8192;               // stream bytes whose groups one pass handles
constexpr uint32_t kNSeg = kPiece / kSeg;
constexpr uint32_t kStage = kPiece + 64;        // a group reaches at most 16 bytes past its start
constexpr uint32_t kGCap = (kPiece - 1) / 9 + 1;  // groups that can start in one piece
constexpr uint32_t kHist = 2048;                // distances are 11 bits
constexpr uint32_t kWin = 8192;                 // output bytes per window
constexpr uint32_t kLit = 0x8000u;              // map entry holds a value
constexpr uint32_t kWTab = kPiece / 2 * 34 / kWin + 2;  // window boundaries inside one piece's output
constexpr uint32_t kNoErr = 0xffffffffu;
constexpr uint64_t kMaxChunk = 1ull << 28;      // error keys are (offset << 2 | kind) in 32 bits

static_assert(kHist + kWin <= kLit, "map indices must stay below the value flag");
static_assert(kPiece / 2 * 34 < (1u << 20), "piece output must fit the 20-bit scan field");

struct PlainLds {
    uint16_t map[kHist + kWin];  // [0, kHist): history; [kHist, ...): the window
    uint32_t gout[kGCap];        // output offset (chunk-relative) of each group of the piece
    uint32_t gpf[kGCap];         // stream offset (piece-relative) | flag byte << 16, per group
    uint32_t wtot[kNT / 64];
    uint16_t wglo[kWTab];        // group covering each window boundary inside the piece's output
    uint8_t stage[kStage];
    uint8_t ex[kNSeg * 17];
    uint8_t entry[kNSeg];
    uint32_t err;                // min over (output offset << 2 | kind)
    uint32_t next_entry;         // first group's offset in the next piece
};
static_assert(sizeof(PlainLds) <= 160 * 1024 / 4, "four workgroups per CU");

// Output bytes of the tokens of the group at stage[p] (flag f) that lie
// wholly inside the first `avail` staged bytes of the stream.  Token t starts
// at p + 1 + t + (matches before it); all eight first bytes are read.
MDZ_DEV uint32_t group_out(const uint8_t* st, uint32_t p, uint32_t f, uint32_t avail) {
    uint32_t ol = 0, q = p + 1;
#pragma unroll
    for (uint32_t t = 0; t < 8; ++t) {
        const uint32_t lt = (f >> t) & 1u, c0 = st[q];
        ol += q + 2 - lt <= avail ? (lt ? 1u : (c0 & 31u) + 3u) : 0u;
        q += 2 - lt;
    }
    return ol;
}

__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(8))) void lzss_plain_kernel(
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
    MDZ_PHASE_START();
    for (uint64_t pb = 0; pb < L && produced < n; pb += kPiece) {
        const uint32_t plen = (uint32_t)(L - pb < kPiece ? L - pb : kPiece);
        const uint32_t avail = (uint32_t)(L - pb < kStage ? L - pb : kStage);
        const uint32_t nseg = (plen + kSeg - 1) / kSeg;

        // ---- stage stream bytes [pb, pb + avail) as words (bytes past avail
        // in the last word zeroed; later words are never used); loads are
        // aligned words whose first byte lies inside the stream, four words
        // per thread in flight
        {
            const uint32_t mis = (uint32_t)(reinterpret_cast<uintptr_t>(src + pb) & 3u);
            const uint32_t* aw = reinterpret_cast<const uint32_t*>(src + pb - mis);
            const uint32_t nwords = (avail + mis + 3) / 4, nw = (avail + 3) / 4;
            uint32_t* st32 = reinterpret_cast<uint32_t*>(S.stage);
            for (uint32_t k0 = tid; k0 < nw; k0 += 4 * kNT) {
                uint32_t lo[4], hi[4];
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    const uint32_t k = k0 + j * kNT;
                    lo[j] = k < nwords ? aw[k] : 0u;
                    hi[j] = k + 1 < nwords ? aw[k + 1] : 0u;
                }
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    const uint32_t k = k0 + j * kNT, b0 = 4 * k;
                    if (k < nw) {
                        const uint32_t w = alignbyte(hi[j], lo[j], mis);
                        st32[k] = b0 + 4 > avail ? w & ((1u << (8 * (avail - b0))) - 1u) : w;
                    }
                }
            }
        }
        __syncthreads();
        MDZ_PHASE(0);

        // ---- 1. exit tables
        for (uint32_t w = tid; w < nseg * 17; w += kNT) {
            const uint32_t s = w / 17, e = w - s * 17;
            const uint32_t lim = (s + 1) * kSeg;
            uint32_t p = s * kSeg + e;
            while (p < lim) p += 17u - __popc(S.stage[p]);
            S.ex[w] = (uint8_t)(p - lim);
        }
        __syncthreads();
        MDZ_PHASE(1);

        // ---- 2. compose across segments (wave 0; the chain value is uniform).
        // Rows are loaded one block ahead; lane j of the wave keeps the entry
        // of segment (s0 + j), written out once per 64 segments.
        if (wv == 0) {
            uint32_t cur = __builtin_amdgcn_readfirstlane(entry0);
            uint32_t ra[8], rb[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j) ra[j] = (j < nseg && ln < 17) ? S.ex[j * 17 + ln] : 0u;
            uint32_t mine = 0;
            for (uint32_t s = 0; s < nseg; s += 8) {
#pragma unroll
                for (uint32_t j = 0; j < 8; ++j)
                    rb[j] = (s + 8 + j < nseg && ln < 17) ? S.ex[(s + 8 + j) * 17 + ln] : 0u;
#pragma unroll
                for (uint32_t j = 0; j < 8; ++j) {
                    mine = ln == ((s + j) & 63u) ? cur : mine;
                    const uint32_t nx = (uint32_t)__builtin_amdgcn_readlane((int)ra[j], (int)cur);
                    cur = s + j < nseg ? nx : cur;
                }
                if (((s + 8) & 63u) == 0 || s + 8 >= nseg) {
                    const uint32_t s0 = s & ~63u;
                    if (s0 + ln < nseg && ln <= s + 7 - s0) S.entry[s0 + ln] = (uint8_t)mine;
                }
#pragma unroll
                for (uint32_t j = 0; j < 8; ++j) ra[j] = rb[j];
            }
            if (ln == 0) S.next_entry = cur;
        }
        __syncthreads();
        MDZ_PHASE(2);
        const uint32_t next_entry = S.next_entry;

        // ---- 3. groups.  a: one thread per segment lists its group starts
        // (count, scan, write); b: one thread per 4 consecutive groups sums
        // their tokens' output bytes, a block scan gives output offsets, and
        // each group covering a window boundary records itself there.
        uint32_t cnt = 0;
        if (tid < nseg) {
            const uint32_t lim = min((tid + 1) * kSeg, plen);
            for (uint32_t p = tid * kSeg + S.entry[tid]; p < lim; p += 17u - __popc(S.stage[p])) ++cnt;
        }
        uint32_t incl = wave_incl_scan_u32(cnt);
        if (ln == 63) S.wtot[wv] = incl;
        __syncthreads();
        uint32_t base = 0, ng = 0;
#pragma unroll
        for (uint32_t k = 0; k < kNT / 64; ++k) {
            const uint32_t t = S.wtot[k];
            base += k < wv ? t : 0u;
            ng += t;
        }
        if (tid < nseg) {
            const uint32_t lim = min((tid + 1) * kSeg, plen);
            uint32_t gi = base + incl - cnt;
            for (uint32_t p = tid * kSeg + S.entry[tid]; p < lim; ++gi) {
                const uint32_t f = S.stage[p];
                S.gpf[gi] = p | (f << 16);
                p += 17u - __popc(f);
            }
        }
        __syncthreads();
        uint32_t gl[4], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t g = 4 * tid + j;
            gl[j] = 0;
            if (g < ng) {
                const uint32_t pf = S.gpf[g];
                gl[j] = group_out(S.stage, pf & 0xffffu, pf >> 16, avail);
            }
            sum += gl[j];
        }
        incl = wave_incl_scan_u32(sum);
        if (ln == 63) S.wtot[wv] = incl;
        __syncthreads();
        uint32_t pout = 0;
        base = 0;
#pragma unroll
        for (uint32_t k = 0; k < kNT / 64; ++k) {
            const uint32_t t = S.wtot[k];
            base += k < wv ? t : 0u;
            pout += t;
        }
        const uint32_t kb0 = produced / kWin + 1;   // first window boundary above `produced`
        {
            uint32_t o = produced + base + incl - sum;
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t g = 4 * tid + j;
                if (g < ng) {
                    S.gout[g] = o;
                    if (gl[j])
                        for (uint32_t b = max((o + kWin - 1) / kWin, kb0); b * kWin < o + gl[j]; ++b)
                            S.wglo[b - kb0] = (uint16_t)g;
                }
                o += gl[j];
            }
        }
        __syncthreads();
        MDZ_PHASE(3);

        // ---- 4./5. windows
        const uint32_t pend = min(produced + pout, n);
        for (uint32_t x0 = produced; x0 < pend;) {
            const uint32_t wend = min(win + kWin, pend);
            // groups [glo, ghi) overlap [x0, wend): x0 is the piece start or a
            // window boundary, wend the piece end or the next boundary
            const uint32_t glo = x0 == produced ? 0u : S.wglo[x0 / kWin - kb0];
            const uint32_t ghi = wend == pend ? ng : S.wglo[wend / kWin - kb0] + 1u;
            const uint32_t mb = kHist - win;       // map index of output offset x: x + mb
            for (uint32_t g = glo + (tid >> 3); g < ghi; g += kNT / 8) {
                const uint32_t k = tid & 7u;
                const uint32_t pf = S.gpf[g];
                uint32_t off = S.gout[g];
                const uint32_t f = pf >> 16;
                uint32_t q = (pf & 0xffffu) + 1, tl = 0, lit = 0, b0 = 0, b1 = 0;
#pragma unroll
                for (uint32_t t = 0; t < 8; ++t) {     // tokens 0..k: offset and bytes
                    const uint32_t lt = (f >> t) & 1u;
                    const uint32_t c0 = S.stage[q], c1 = S.stage[q + 1];
                    const uint32_t len = q + 2 - lt <= avail ? (lt ? 1u : (c0 & 31u) + 3u) : 0u;
                    off += t < k ? len : 0u;
                    tl = t == k ? len : tl;
                    lit = t == k ? lt : lit;
                    b0 = t == k ? c0 : b0;
                    b1 = t == k ? c1 : b1;
                    q += 2 - lt;
                }
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
                    } else if (d >= tl) {              // no self-overlap: entry i holds i - d
                        uint32_t i = xa + mb;
                        const uint32_t ie = xb + mb;
                        if ((i & 1u) && i < ie) { S.map[i] = (uint16_t)(i - d); ++i; }
                        uint32_t pk = (i - d) | ((i + 1 - d) << 16);
                        for (; i + 2 <= ie; i += 2, pk += 0x20002u)
                            *reinterpret_cast<uint32_t*>(&S.map[i]) = pk;
                        if (i < ie) S.map[i] = (uint16_t)(i - d);
                    } else {                           // first period of the match
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
            MDZ_PHASE(4);

            // resolve [x0, wend) two entries per thread and step, both chains in
            // flight; a resolved chain re-reads map[0] (a value) instead of
            // branching.  The pair's low entry may precede x0 (already a
            // value); its high entry past wend is left as it is.
            for (uint32_t i = ((x0 + mb) & ~1u) + 2 * tid; i < wend + mb; i += 2 * kNT) {
                const uint32_t w = *reinterpret_cast<const uint32_t*>(&S.map[i]);
                const bool hi_in = i + 1 < wend + mb;
                uint32_t e0 = w & 0xffffu, e1 = hi_in ? w >> 16 : kLit;
                if ((e0 & e1 & kLit) == 0) {
                    do {
                        const uint32_t a0 = (e0 & kLit) ? 0u : e0, a1 = (e1 & kLit) ? 0u : e1;
                        const uint32_t v0 = S.map[a0], v1 = S.map[a1];
                        e0 = (e0 & kLit) ? e0 : v0;
                        e1 = (e1 & kLit) ? e1 : v1;
                    } while ((e0 & e1 & kLit) == 0);
                    if (hi_in) *reinterpret_cast<uint32_t*>(&S.map[i]) = e0 | (e1 << 16);
                    else S.map[i] = (uint16_t)e0;
                }
            }
            __syncthreads();
            MDZ_PHASE(5);

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
                MDZ_PHASE(6);
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
