// Byte RLE device code shared by the per-chunk kernel and the batched
// persistent kernel.
#pragma once
// Byte RLE (ORC byte RLE) decoder for gfx950.
//
// Replaces brle::kernel_decompress / brle::decode (reference
// include/brle/brle.h:26-109, one thread per 256 KiB chunk) and
// brle::decompress_gpu (:305-342).  Control byte c: c <= 127 -> a run of
// c+3 copies of the next byte; c >= 128 -> c-127 literal bytes follow.
//
// One workgroup of 256 threads per chunk:
//   1. the chunk's compressed bytes are staged into LDS with 16-byte loads;
//   2. token discovery in parallel: the bytes are cut into 256 windows;
//      thread i walks the token chain speculatively from kOverlap bytes
//      before its window and notes the first token start in its window and
//      the first one past it.  A wrong starting point usually falls into
//      step with the true chain, so adjacent threads agree where their
//      chains meet; a window that disagrees re-walks from its predecessor's
//      exit, in parallel rounds until every window agrees;
//   3. a block-wide prefix sum of the windows' output lengths gives every
//      thread its output offset;
//   4. output is cut at 16-byte boundaries of the destination: each thread
//      writes the aligned blocks from the boundary at or after its window's
//      output start to the one at or after its end, assembling each block in
//      registers from token pieces (runs splat, literals read from LDS at
//      the block offset), so all stores but the chunk's two end blocks are
//      aligned 16-byte stores.
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

constexpr uint32_t kBrleThreads = 256;  // workgroup of the per-chunk kernel
constexpr uint32_t kOverlap = 96;       // bytes walked before a window to fall into step
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kBrlePad = 64;       // staged bytes past the input that expansion may read

template <uint32_t NT>
struct BrleTables {
    uint32_t first[NT];         // first true token start in window i
    uint32_t exit_[NT];         // first token start at or past the next window
    uint32_t olen[NT];          // output bytes of window i's tokens
    uint32_t ooff[NT];          // output offset of window i
    uint32_t wtot[NT / 64];     // per-wave output totals
    uint32_t trunc;             // output bytes of a last token that overruns the input
    uint32_t total;
    uint32_t any[2];            // repair-round flags
};


// Token length in input bytes and output bytes for the token at p.
MDZ_DEV void token_at(const uint8_t* d, uint32_t p, uint32_t& in_len, uint32_t& out_len) {
    const uint32_t c = d[p];
    if (c > 127) { in_len = c - 126; out_len = c - 127; }
    else { in_len = 2; out_len = c + 3; }
}

// Walks the chain from `from` over window [ws, we): returns the first token
// start >= ws in *first, the first start >= we as the result, and the output
// bytes of tokens starting in [ws, we) in *olen.  Stops at the end L.
MDZ_DEV uint32_t walk(const uint8_t* d, uint32_t L, uint32_t from, uint32_t ws, uint32_t we,
                      uint32_t& first, uint32_t& olen) {
    uint32_t p = from, o = 0;
    first = kNone;
    while (p < we && p < L) {
        uint32_t il, ol;
        token_at(d, p, il, ol);
        if (p >= ws) {
            if (first == kNone) first = p;
            o += ol;
        }
        p += il;
    }
    if (first == kNone) first = p;     // no token starts inside the window
    olen = o;
    return p;
}

// One 16-byte output block assembled in registers.  Pieces of tokens are
// merged in output order: merge(f, w) keeps the block's bytes below offset f
// and takes w's bytes from f up (w carries each byte at its block position),
// so no shifting of the block is needed; bytes above the last piece are
// garbage until a later piece (or the block's end) covers them.
struct Block16 {
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;

    MDZ_DEV static uint32_t below(int32_t x) {       // bytes [0, x) of a word
        return x >= 4 ? ~0u : x <= 0 ? 0u : (1u << (8 * x)) - 1u;
    }
    MDZ_DEV void merge(uint32_t f, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3) {
        const int32_t x = (int32_t)f;
        const uint32_t m0 = below(x), m1 = below(x - 4), m2 = below(x - 8), m3 = below(x - 12);
        w0 = (w0 & m0) | (v0 & ~m0);
        w1 = (w1 & m1) | (v1 & ~m1);
        w2 = (w2 & m2) | (v2 & ~m2);
        w3 = (w3 & m3) | (v3 & ~m3);
    }
    MDZ_DEV uint32_t byte(uint32_t j) const {
        const uint32_t w = j < 4 ? w0 : j < 8 ? w1 : j < 12 ? w2 : w3;
        return (w >> (8 * (j & 3))) & 0xffu;
    }
};

// 16 bytes of the staged input starting at signed byte index i of the
// 4-byte aligned LDS array w32 (i >= -16: the bytes before the input are
// the tables, read and then masked out by the caller).
MDZ_DEV void lds_bytes16(const uint32_t* w32, int32_t i, uint32_t& v0, uint32_t& v1, uint32_t& v2,
                         uint32_t& v3) {
    const int32_t wi = i >> 2;                  // floor
    const uint32_t sh = (uint32_t)i & 3u;
    const uint32_t u0 = w32[wi], u1 = w32[wi + 1], u2 = w32[wi + 2], u3 = w32[wi + 3],
                   u4 = w32[wi + 4];
    v0 = alignbyte(u1, u0, sh);
    v1 = alignbyte(u2, u1, sh);
    v2 = alignbyte(u3, u2, sh);
    v3 = alignbyte(u4, u3, sh);
}

// Launch arguments of one byte-RLE container decode.
struct BrleArgs {
    const uint8_t* data;
    const uint64_t* ptr;
    uint64_t n_chunks, chunk_bytes;
    uint8_t* dst;
    uint64_t dst_bytes;
    int32_t* status;
    uint32_t data_cap;
};

// Decodes chunk c with a workgroup of NT threads (one window each); smem
// holds the tables followed by data_cap bytes of staged input.
template <uint32_t NT>
MDZ_DEV void brle_chunk(const BrleArgs& a, uint64_t c, uint32_t* smem) {
    constexpr uint32_t kThreads = NT;
    const uint8_t* __restrict__ data = a.data;
    const uint64_t* __restrict__ ptr = a.ptr;
    const uint64_t chunk_bytes = a.chunk_bytes, dst_bytes = a.dst_bytes;
    uint8_t* __restrict__ dst = a.dst;
    int32_t* __restrict__ status = a.status;
    const uint32_t data_cap = a.data_cap;
    BrleTables<NT>& tb = *reinterpret_cast<BrleTables<NT>*>(smem);
    uint8_t* const lds = reinterpret_cast<uint8_t*>(smem) + sizeof(BrleTables<NT>);
    static_assert(sizeof(BrleTables<NT>) % 16 == 0 && sizeof(BrleTables<NT>) >= 16,
                  "staged input must stay 16-byte aligned with table bytes before it");

    const uint32_t tid = threadIdx.x;
    MDZ_PHASE_START();
    const uint64_t beg = ptr[c], fin = ptr[c + 1];
    const uint64_t out_base = c * chunk_bytes;
    const uint64_t cap = dst_bytes > out_base ? min(dst_bytes - out_base, chunk_bytes) : 0;
    if (fin < beg || fin - beg + kBrlePad > data_cap) {
        if (tid == 0 && status) status[c] = MDZ_ST_BAD_HEADER;
        return;
    }
    const uint32_t L = (uint32_t)(fin - beg);

    // ---- 1. stage: aligned 16-byte blocks covering [beg, fin)
    const uint64_t a0 = beg & ~15ull;
    const uint32_t skew = (uint32_t)(beg - a0);
    const uint32_t nblk = (skew + L + 15) / 16;
    const u32x4* src = reinterpret_cast<const u32x4*>(data + a0);
    for (uint32_t b = tid; b < nblk; b += kThreads) {
        if (a0 + 16ull * (b + 1) <= fin) {
            reinterpret_cast<u32x4*>(lds)[b] = __builtin_nontemporal_load(src + b);
        } else {                                // last block: no read past the data
            for (uint32_t k = 0; k < 16; ++k) {
                const uint64_t at = a0 + 16ull * b + k;
                lds[16 * b + k] = at < fin ? data[at] : 0;
            }
        }
    }
    if (tid == 0) { tb.trunc = 0; tb.any[0] = 0; }
    __syncthreads();
    const uint8_t* d = lds + skew;
    MDZ_PHASE(0);

    // ---- 2. speculative token discovery
    const uint32_t W = (L + kThreads - 1) / kThreads;
    const uint32_t ws = min(L, tid * W), we = min(L, ws + W);
    {
        const uint32_t from = tid == 0 ? 0 : (ws > kOverlap ? ws - kOverlap : 0);
        uint32_t first, olen;
        const uint32_t ex = walk(d, L, from, ws, we, first, olen);
        tb.first[tid] = first;
        tb.exit_[tid] = ex;
        tb.olen[tid] = olen;
    }
    // repair rounds: a window whose first token start is not the previous
    // window's exit re-walks from that exit.  Window 0 starts at the true
    // entry, so after round r windows 0..r are right; a wrong chain usually
    // joins the true one inside its window, so few rounds are needed.
    // The "any window wrong" flag alternates between two words: round r
    // sets any[r&1] and clears any[(r+1)&1] before the barrier, so each flag
    // is zero when its round starts and nobody reads it while it is cleared.
    __syncthreads();
    MDZ_PHASE(1);
    for (uint32_t round = 0;; ++round) {
        const bool wrong = tid > 0 && tb.first[tid] != tb.exit_[tid - 1];
        const uint32_t e = wrong ? tb.exit_[tid - 1] : 0;
        if (wrong) tb.any[round & 1] = 1;
        if (tid == 0) tb.any[(round + 1) & 1] = 0;
        __syncthreads();
        if (!tb.any[round & 1]) break;
        if (wrong) {
            uint32_t first, olen;
            tb.exit_[tid] = walk(d, L, e, ws, we, first, olen);
            tb.first[tid] = first;
            tb.olen[tid] = olen;
        }
        __syncthreads();
    }
    // The chain is now exact.  Its last token overran the input iff the last
    // exit lies past L; the window holding that token (first < we, exit > L)
    // finds it and notes its output length, which is not written (the
    // reference decoder stops before such a token).
    const uint32_t my_first = tb.first[tid];
    if (my_first < we && tb.exit_[tid] > L) {
        uint32_t p = my_first, il = 0, ol = 0;
        for (;;) {
            token_at(d, p, il, ol);
            if (p + il > L) break;
            p += il;
        }
        tb.trunc = ol;
    }
    MDZ_PHASE(2);

    // ---- 3. output offsets (block-wide exclusive scan over NT/64 waves)
    {
        const uint32_t l = tid & 63, wv = tid >> 6;
        const uint32_t v = tb.olen[tid];
        const uint32_t incl = wave_incl_scan_u32(v);
        if (l == 63) tb.wtot[wv] = incl;
        __syncthreads();
        uint32_t base = 0, tot = 0;
#pragma unroll
        for (uint32_t k = 0; k < NT / 64; ++k) {
            base += k < wv ? tb.wtot[k] : 0;
            tot += tb.wtot[k];
        }
        tb.ooff[tid] = base + incl - v;
        if (tid == 0) tb.total = tot;
    }
    __syncthreads();

    // Status, as the reference decoder would stop: a token that crosses the
    // output end (OUTPUT_OVERRUN; the bytes that fit are written) comes
    // before a truncated last token in chain order (INPUT_OVERRUN).
    const uint32_t trunc = tb.trunc;
    const uint64_t t_ok = (uint64_t)tb.total - trunc;   // output of the complete tokens
    const uint64_t lim = min(t_ok, cap);
    if (tid == 0 && status)
        status[c] = t_ok > cap ? MDZ_ST_OUTPUT_OVERRUN : trunc ? MDZ_ST_INPUT_OVERRUN : MDZ_ST_OK;
    MDZ_PHASE(3);

    // ---- 4. expand.  Output positions are cut at 16-byte boundaries of the
    // destination address: thread t writes [up(ooff[t]), up(ooff[t+1])),
    // so every store is an aligned 16-byte store except at the two ends of
    // the chunk.  A thread starts at its window's first token, skips the
    // bytes below its first boundary (the previous thread writes them) and
    // runs on into the next window's tokens up to its last boundary.
    {
        uint8_t* const out = dst + out_base;
        const uint32_t mis = (uint32_t)((uintptr_t)out & 15);
        auto up = [&](uint64_t q) -> uint64_t { return ((q + mis + 15) & ~15ull) - mis; };
        const uint64_t o_me = tb.ooff[tid];
        const uint64_t q0 = tid == 0 ? 0 : min(up(o_me), lim);
        const uint64_t q1 = tid == kThreads - 1 ? lim : min(up((uint64_t)tb.ooff[tid + 1]), lim);
        const uint32_t* const w32 = reinterpret_cast<const uint32_t*>(lds);
        uint32_t p = my_first;                   // next token to open
        uint32_t rem = 0;                        // bytes left in the open token
        uint32_t lit = 0, val = 0;               // literal flag / run byte splat
        int32_t sidx = 0;                        // LDS index of the next literal byte
        uint64_t q = o_me;
        auto open = [&]() {
            const uint32_t cc = d[p], nb = d[p + 1];
            lit = cc > 127;
            rem = lit ? cc - 127 : cc + 3;
            val = nb * 0x01010101u;
            sidx = (int32_t)(skew + p + 1);
            p += lit ? rem + 1 : 2;
        };
        // skip to the first boundary (fewer than 16 bytes)
        while (q < q0) {
            if (rem == 0) open();
            const uint32_t n = (uint32_t)min((uint64_t)rem, q0 - q);
            rem -= n; sidx += (int32_t)n; q += n;
        }
        Block16 blk;
        uint32_t f = (uint32_t)((q + mis) & 15), f0 = f;
        while (q < q1) {
            if (rem == 0) open();
            const uint32_t n = (uint32_t)min(min((uint64_t)rem, (uint64_t)(16 - f)), q1 - q);
            uint32_t v0 = val, v1 = val, v2 = val, v3 = val;
            if (lit) lds_bytes16(w32, sidx - (int32_t)f, v0, v1, v2, v3);
            blk.merge(f, v0, v1, v2, v3);
            f += n; q += n; rem -= n; sidx += (int32_t)n;
            if (f == 16 || q == q1) {
                uint8_t* const bp = out + (q - f);   // block start (aligned)
                if (f0 == 0 && f == 16) {
                    *reinterpret_cast<u32x4*>(bp) = u32x4{blk.w0, blk.w1, blk.w2, blk.w3};
                } else {
                    for (uint32_t j = f0; j < f; ++j) bp[j] = (uint8_t)blk.byte(j);
                }
                f = 0; f0 = 0;
            }
        }
    }
    MDZ_PHASE(4);
}

}  // namespace
}  // namespace mdz
