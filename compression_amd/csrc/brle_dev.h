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
//   4. each thread expands its window's tokens through a 16-byte register
//      window and writes 16-byte aligned stores (bytes at the two ragged
//      ends of its output range are stored singly).
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

constexpr uint32_t kBrleThreads = 256;  // workgroup of the per-chunk kernel
constexpr uint32_t kOverlap = 96;       // bytes walked before a window to fall into step
constexpr uint32_t kNone = 0xffffffffu;

template <uint32_t NT>
struct BrleTables {
    uint32_t first[NT];         // first true token start in window i
    uint32_t exit_[NT];         // first token start at or past the next window
    uint32_t olen[NT];          // output bytes of window i's tokens
    uint32_t ooff[NT];          // output offset of window i
    uint32_t wtot[NT / 64];     // per-wave output totals
    uint32_t err;               // min over (input position << 8 | status byte)
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

// 16-byte register window for aligned output stores.  Bytes before the
// first 16-byte boundary of a thread's output range, and after the last one,
// are stored singly; everything between goes out in aligned 16-byte stores.
struct ByteSink {
    uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0, na = 0;
    uint8_t* out;
    uint64_t pos;      // next output byte (offset from out)
    uint64_t cap;      // bytes that may be written

    MDZ_DEV bool head() const { return na == 0 && (((uintptr_t)out + pos) & 15) != 0; }
    MDZ_DEV void put1(uint32_t b) {
        if (head()) {
            if (pos < cap) out[pos] = (uint8_t)b;
            ++pos;
            return;
        }
        a0 = alignbyte(a1, a0, 1); a1 = alignbyte(a2, a1, 1); a2 = alignbyte(a3, a2, 1);
        a3 = (a3 >> 8) | (b << 24);
        ++pos;
        if (++na == 16) flush16();
    }
    MDZ_DEV void put4(uint32_t w) {                // window active, na % 4 == 0
        a0 = a1; a1 = a2; a2 = a3; a3 = w;
        pos += 4;
        na += 4;
        if (na == 16) flush16();
    }
    MDZ_DEV void flush16() {
        const uint64_t at = pos - 16;
        if (at + 16 <= cap) {
            *reinterpret_cast<u32x4*>(out + at) = u32x4{a0, a1, a2, a3};
        } else {
            const uint32_t w[4] = {a0, a1, a2, a3};
            for (uint32_t i = 0; i < 16 && at + i < cap; ++i)
                out[at + i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
        }
        na = 0;
    }
    MDZ_DEV void finish() {                        // ragged tail
        const uint32_t w[4] = {a0, a1, a2, a3};
        const uint64_t at = pos - na;
        for (uint32_t i = 0; i < na; ++i) {
            const uint32_t bi = 16 - na + i;
            if (at + i < cap) out[at + i] = (uint8_t)(w[bi >> 2] >> (8 * (bi & 3)));
        }
        na = 0;
    }
    MDZ_DEV void fill(uint32_t b, uint32_t n) {
        const uint32_t w = b * 0x01010101u;
        while (n) {
            if (!head() && (na & 3) == 0 && n >= 4) { put4(w); n -= 4; }
            else { put1(b); --n; }
        }
    }
};

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

    const uint32_t tid = threadIdx.x;
    const uint64_t beg = ptr[c], fin = ptr[c + 1];
    const uint64_t out_base = c * chunk_bytes;
    const uint64_t cap = dst_bytes > out_base ? min(dst_bytes - out_base, chunk_bytes) : 0;
    if (fin < beg || fin - beg + 32 > data_cap) {
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
    if (tid == 0) { tb.err = kNone; tb.any[0] = 0; }
    __syncthreads();
    const uint8_t* d = lds + skew;

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

    // ---- 4. expand this window's tokens
    {
        ByteSink sink;
        sink.out = dst + out_base;
        sink.cap = cap;
        sink.pos = tb.ooff[tid];
        uint32_t p = tb.first[tid];
        uint32_t e = MDZ_ST_OK, epos = 0;
        while (p < we && p < L) {
            const uint32_t cc = d[p];
            uint32_t il, ol;
            token_at(d, p, il, ol);
            if (p + il > L) { e = MDZ_ST_INPUT_OVERRUN; epos = p; break; }
            // a token that crosses the output end still writes the bytes
            // that fit (the sink drops the rest), then stops the window
            const bool over = sink.pos + ol > cap;
            if (cc > 127) {
                for (uint32_t k = 0; k < ol; ++k) sink.put1(d[p + 1 + k]);
            } else {
                sink.fill(d[p + 1], ol);
            }
            if (over) { e = MDZ_ST_OUTPUT_OVERRUN; epos = p; break; }
            p += il;
        }
        sink.finish();
        if (e != MDZ_ST_OK) atomicMin(&tb.err, (epos << 8) | ((uint32_t)e & 0xffu));
    }
    __syncthreads();
    if (tid == 0 && status) {
        const uint32_t e = tb.err;
        status[c] = e == kNone ? MDZ_ST_OK : (int32_t)(int8_t)(e & 0xff);
    }
}

}  // namespace
}  // namespace mdz
