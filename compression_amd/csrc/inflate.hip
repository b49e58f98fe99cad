// Deflate / zlib inflate for gfx950: two wavefronts per stream.
//
// Replaces the reference's deflate::decompress_gpu inflate kernels (reference
// include/zlib/zlib.h:832-962; reader, decoder and writer warps of one lane
// each doing useful work) and cuDF's gpuinflate (include/cudf_deflate.cuh:
// 1229-1345,1399-1407) behind the same batched descriptor/status interface.
//
// Per stream, one 128-thread workgroup: a producer wave and a consumer wave
// joined by a queue of three items in LDS (a round's symbol list, or a
// stored block), with a published and a finished count, so either wave can
// run up to two items ahead of the other (a barrier only at the start and
// end of each stream).  The rounds' lists share one ring of 6,144 symbols:
// a round keeps up to 4,096, and the producer lists a round once the
// consumer has freed room for it.
//   * producer: parses block headers, builds each block's tables in LDS
//     (literal/length: 10-bit root plus subtables; distance: 8-bit root plus
//     subtables; codes longer than a subtable holds take a canonical
//     count/symbol decode, which also rejects codes an incomplete set does
//     not define), and decodes symbols in rounds.  A round stages 64 spans
//     of 512 input bits in LDS; lane t decodes its span as if a symbol
//     started at its first bit and marks each start it meets.  Huffman codes
//     resynchronise, so the true symbol sequence, followed past the end of
//     lane t's span, soon meets a start a later lane marked (at most 3 spans
//     ahead); the lanes it passes through form a chain, each lists its piece
//     of the sequence, and the pieces are placed by popcounts in the symbol
//     ring.  An end-of-block, an error, a lost hand-off or a full list ends
//     the round there and the next round starts from it;
//   * consumer: turns each published round's list into output in a 34.5 KiB
//     LDS ring holding the last 32 KiB: batches of up to 64 symbols (about
//     2 KiB of output), literals stored by their lanes, then matches copied
//     in order, one byte per lane (byte j of a match reads start - dist +
//     (j mod dist), so overlapping matches need no serial step); stored
//     blocks are copied 32 bytes per lane; finished output is stored to HBM
//     with dword stores at dword-aligned addresses.
// Errors follow zlib's checks and order (invalid block type, stored length,
// code-set completeness, repeat codes, missing end-of-block, invalid codes,
// distance too far back) with cuDF's status codes.  parse_hdr selects the
// framing: raw Deflate, zlib (2-byte header checked) or gzip (cuDF's header
// rules, applied by a prepass kernel; the 8-byte footer is excluded from the
// Deflate data); with MDZ_HDR_VERIFY a kernel after the decode checks each
// trailer (Adler-32, or CRC-32 and ISIZE) against the output written.
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "deflate_hdr_dev.h"
#include "mdz_device.h"
#include "mdz_internal.h"

// Device memory of mdz_inflate_batch_sched_noalloc: word 0 is its work-queue
// cursor.  Every device the library runs on holds one copy until the process
// ends; calls that overlap on one device share it (and race).  Host code uses
// the per-device address from hipGetSymbolAddress, never this name.
static __device__ __align__(256) uint32_t scratch[64];

// Phase timing of the two roles (profiling builds only): lane 0 of each wave
// adds the cycles since its previous mark to mdz_phase_cycles[k].
#ifdef MDZ_PHASE_PROF
#define INF_PHASE_START() uint64_t inf_phase_t_ = clock64()
#define INF_PHASE(k)                                                                 \
    do {                                                                             \
        const uint64_t inf_now_ = clock64();                                         \
        if (lane_id() == 0) atomicAdd(&mdz_phase_cycles[k], inf_now_ - inf_phase_t_); \
        inf_phase_t_ = inf_now_;                                                     \
    } while (0)
#else
#define INF_PHASE_START() ((void)0)
#define INF_PHASE(k) ((void)0)
#endif

namespace mdz {
namespace {

constexpr uint32_t kBatchOut = 2048;                    // a batch closes once this much output is queued
constexpr uint32_t kRing = 32768 + kBatchOut + 512;     // history + one batch (< kBatchOut + 258 bytes)
constexpr uint32_t kLitRoot = 10, kLitCap = 1024 + 512; // two-level tables: root bits, root + subtables
constexpr uint32_t kDistRoot = 8, kDistCap = 256 + 384;
constexpr uint32_t kIndirect = 0x400;                   // root entry pointing at a subtable
constexpr uint32_t kSpanLog = 9, kSpan = 1u << kSpanLog; // input bits each lane takes per round
constexpr uint32_t kReach = 3;                          // spans a hand-off may run ahead
constexpr uint32_t kEvents = 16;                        // span-decode events a lane keeps
constexpr uint32_t kStageWords = 64 * kSpan / 32 + 8;   // one round's input, plus a symbol's overhang
// Staged words and hand-off bits are stored with one padding word after every
// 16: lane t's span starts at word 16t, so unpadded the 64 lanes' accesses
// would fall on a few LDS banks; padded, word w sits at w + w / 16.
MDZ_DEV uint32_t pad16(uint32_t w) { return w + (w >> 4); }
constexpr uint32_t kStagePadded = kStageWords + kStageWords / 16 + 1;
constexpr uint32_t kSymCap = 4096;                      // true-path symbols kept per round
constexpr uint32_t kSymRing = 6144;                     // symbol ring the rounds' lists share
constexpr uint32_t kSlots = 3;                          // queue depth (items) between the two waves
MDZ_DEV uint32_t sym_wrap(uint32_t x) { return x >= kSymRing ? x - kSymRing : x; }   // x < 2 * kSymRing
// Polls (each after an s_sleep of about 64 cycles) before a wave waiting on
// the other gives up on the stream with MDZ_ST_INTERNAL: far above any
// item's decode or output time, so only a broken queue reaches it.  The
// launch passes the limit; a test lowers it (MDZ_INFLATE_WAIT_POLLS) to
// drive the give-up path from the host.
constexpr uint32_t kWaitPolls = 1u << 24;

// Table entry: [3:0] code length (0: no entry, take the canonical path),
// [7:4] extra bits, [9:8] kind (0 literal, 1 end of block, 2 length/distance,
// 3 invalid), [31:16] value (literal byte, base length or base distance).  A
// root entry with kIndirect set points at a subtable instead: [14:11] its
// index bits, [31:16] its first entry; subtable entries hold the full code
// length.
enum : uint32_t { kLit = 0, kEob = 1, kRef = 2, kBad = 3 };

__constant__ uint16_t kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                                      35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2,
                                      3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
__constant__ uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193,
                                       257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145,
                                       8193, 12289, 16385, 24577};
__constant__ uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6,
                                       7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
__constant__ uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

MDZ_DEV uint32_t lit_entry_bits(uint32_t sym) {          // kind / extra / value for a litlen symbol
    if (sym < 256) return (kLit << 8) | (sym << 16);
    if (sym == 256) return kEob << 8;
    if (sym <= 285) return (kRef << 8) | ((uint32_t)kLenExtra[sym - 257] << 4) |
                           ((uint32_t)kLenBase[sym - 257] << 16);
    return kBad << 8;
}
MDZ_DEV uint32_t dist_entry_bits(uint32_t sym) {
    if (sym < 30) return (kRef << 8) | ((uint32_t)kDistExtra[sym] << 4) | ((uint32_t)kDistBase[sym] << 16);
    return kBad << 8;
}

// One canonical code: counts per length, symbols by (length, value), LUT.
struct Code {
    uint16_t cnt[16];
    uint16_t sym[288];
};

// Block-header scratch; it shares its space with the round's symbol list.
struct HdrScratch {
    uint8_t lens[320];                         // literal/length then distance code lengths
    uint32_t cl_lut[128];                      // code-length code: 7-bit table, complete
    Code cl;
};

// One entry of the queue between the producer wave (block headers and
// symbol decode) and the consumer wave (LZ77 output and stores).
enum : uint32_t { kItemRound = 0, kItemStored = 1, kItemNone = 2 };
struct Item {
    uint32_t kind;        // a round's symbols, a stored block, or nothing
    uint32_t nsym;        // round: symbols listed in the item's buffer
    int32_t status;       // after the item's output: the stream's status (0: none)
    uint32_t last;        // the stream ends with this item
    uint64_t src;         // stored: stream byte offset of the data
    uint64_t take;        // stored: bytes present (take < LEN ends in an input overrun)
    uint32_t sym0;        // round: symbol ring index of its first symbol
};

struct Ctrl {
    uint32_t qi;          // queue position of the stream in hand
    uint32_t halt;        // a wave has stopped the stream (consumer finished or failed, or a wait ran out)
    uint32_t n_pub;       // items the producer has published
    uint32_t n_fin;       // items the consumer has finished with
    uint32_t sym_fin;     // symbols of finished items (a running count, mod 2^32)
    uint32_t n_stored, n_fixed, n_dyn;
    uint64_t consumed;    // Deflate input bits through the last item
    uint64_t desc[4];     // a live queue's descriptor, read once by thread 0
};

struct InflateLds {
    uint32_t lit[kLitCap];                     // literal/length table, 10-bit root
    uint32_t dist[kDistCap];                   // distance table, 8-bit root
    Code lit_c, dist_c;                        // canonical tables of the same codes
    uint32_t in[kStagePadded];                 // the round's input bits (padded: pad16)
    union {
        HdrScratch h;
        uint32_t bm[64 * (kSpan / 32) * 17 / 16];   // per lane: symbol starts met in its span (padded)
    } u;
    uint32_t sym[kSymRing];                    // rounds' symbols, in stream order, as a ring
    Item item[kSlots];
    Ctrl ctrl;
    uint8_t ring[kRing];                       // the consumer's output window
    uint32_t sink[64];                         // per-lane target of stores past a match's end
};
static_assert(sizeof(InflateLds) <= 80 * 1024, "two workgroups per CU");

MDZ_DEV uint32_t bitrev(uint32_t v, uint32_t n) { return __builtin_bitreverse32(v) >> (32 - n); }

// Builds code `c` and its LUT from lens[0..n) (whole wave).  kind: 0 litlen,
// 1 distance, 2 code lengths.  Returns 0, or -1 for an over-subscribed set or
// an incomplete one that zlib rejects (incomplete is allowed only for
// literal/length and distance sets whose longest code has length 1).
MDZ_DEV int build_code(Code& c, uint32_t* lut, uint32_t lut_bits, const uint8_t* lens, uint32_t n,
                       uint32_t kind) {
    const uint32_t l = lane_id();
    uint32_t cnt = 0;                          // lane k < 16 counts length k
    for (uint32_t s0 = 0; s0 < n; s0 += 64) {
        const uint32_t s = s0 + l;
        const uint32_t len = s < n ? lens[s] : 0;
#pragma unroll
        for (uint32_t k = 1; k < 16; ++k) {
            const uint32_t b = __popcll(__ballot(len == k));
            if (l == k) cnt += b;
        }
    }
    if (l < 16) c.cnt[l] = (uint16_t)(l == 0 ? 0 : cnt);
    wave_sync();
    // completeness (zlib inftrees.c rules), max length, symbol offsets, first codes
    int left = 1, maxlen = 0;
    uint32_t offs[16], code = 0, first[16];
    offs[1] = 0;
#pragma unroll
    for (uint32_t k = 1; k < 16; ++k) {
        const uint32_t ck = c.cnt[k];
        left = (left << 1) - (int)ck;
        if (left < 0) return -1;
        if (ck) maxlen = (int)k;
        code = (code + (k > 1 ? c.cnt[k - 1] : 0)) << 1;
        first[k] = code;
        if (k < 15) offs[k + 1] = offs[k] + ck;
    }
    if (left > 0 && (kind == 2 || maxlen != 1) && maxlen != 0) return -1;
    // every LUT entry starts as "canonical path"; short codes overwrite theirs.
    // An empty code-length code decodes, as zlib's table does, every pattern
    // as length 0 with 1 bit.
    const uint32_t lsize = 1u << lut_bits;
    const uint32_t init = (kind == 2 && maxlen == 0) ? 1u : 0u;
    for (uint32_t i = l; i < lsize; i += 64) lut[i] = init;
    wave_sync();
    uint32_t run[16];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) run[k] = 0;
    for (uint32_t s0 = 0; s0 < n; s0 += 64) {
        const uint32_t s = s0 + l;
        const uint32_t len = s < n ? lens[s] : 0;
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t k = 1; k < 16; ++k) {
            const uint64_t b = __ballot(len == k);
            if (len == k) rank = run[k] + __popcll(b & ((1ull << l) - 1));
            run[k] += __popcll(b);
        }
        if (len) {
            uint32_t off = 0, fc = 0;
#pragma unroll
            for (uint32_t k = 1; k < 16; ++k)
                if (k == len) { off = offs[k]; fc = first[k]; }
            c.sym[off + rank] = (uint16_t)s;
            if (len <= lut_bits) {
                const uint32_t e = len | (kind == 0 ? lit_entry_bits(s) : kind == 1 ? dist_entry_bits(s)
                                                                                    : (s << 16));
                for (uint32_t j = bitrev(fc + rank, len); j < lsize; j += 1u << len) lut[j] = e;
            }
        }
    }
    wave_sync();
    return 0;
}

MDZ_DEV bool is_marker(uint32_t v, uint32_t rb) { return v > rb && v < 16; }

// Builds the two-level table T (root bits rb, cap entries) and the canonical
// tables c of the literal/length (kind 0) or distance (kind 1) code with
// lengths lens[0..n), whole wave.  Codes of at most rb bits fill root
// entries; each root prefix of longer codes gets a subtable of 2^(m - rb)
// entries, m its longest code.  If the subtables do not fit in cap, those
// root entries stay empty and long codes decode through c.  Returns 0, or -1
// for a set zlib rejects (over-subscribed, or incomplete unless its longest
// code has length 1).  Empty entries of incomplete sets mark patterns no code
// defines.
MDZ_DEV int build_table2(Code& c, uint32_t* T, uint32_t rb, uint32_t cap, const uint8_t* lens,
                         uint32_t n, uint32_t kind) {
    const uint32_t l = lane_id();
    uint32_t cnt = 0;                          // lane k < 16 counts length k
    for (uint32_t s0 = 0; s0 < n; s0 += 64) {
        const uint32_t s = s0 + l;
        const uint32_t len = s < n ? lens[s] : 0;
#pragma unroll
        for (uint32_t k = 1; k < 16; ++k) {
            const uint32_t b = __popcll(__ballot(len == k));
            if (l == k) cnt += b;
        }
    }
    if (l < 16) c.cnt[l] = (uint16_t)(l == 0 ? 0 : cnt);
    wave_sync();
    int left = 1, maxlen = 0;
    uint32_t offs[17], code = 0, first[16];
    offs[0] = offs[1] = 0;
#pragma unroll
    for (uint32_t k = 1; k < 16; ++k) {
        const uint32_t ck = c.cnt[k];
        left = (left << 1) - (int)ck;
        if (left < 0) return -1;
        if (ck) maxlen = (int)k;
        code = (code + (k > 1 ? c.cnt[k - 1] : 0)) << 1;
        first[k] = code;
        offs[k + 1] = offs[k] + ck;
    }
    if (left > 0 && maxlen != 1 && maxlen != 0) return -1;
    const uint32_t R = 1u << rb;
    for (uint32_t i = l; i < cap; i += 64) T[i] = 0;
    wave_sync();
    // sorted symbols, root entries of short codes, longest code per long prefix
    uint32_t run[16];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) run[k] = 0;
    for (uint32_t s0 = 0; s0 < n; s0 += 64) {
        const uint32_t s = s0 + l;
        const uint32_t len = s < n ? lens[s] : 0;
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t k = 1; k < 16; ++k) {
            const uint64_t b = __ballot(len == k);
            if (len == k) rank = run[k] + __popcll(b & ((1ull << l) - 1));
            run[k] += __popcll(b);
        }
        if (len) {
            uint32_t off = 0, fc = 0;
#pragma unroll
            for (uint32_t k = 1; k < 16; ++k)
                if (k == len) { off = offs[k]; fc = first[k]; }
            c.sym[off + rank] = (uint16_t)s;
            const uint32_t rev = bitrev(fc + rank, len);
            if (len <= rb) {
                const uint32_t e = len | (kind == 0 ? lit_entry_bits(s) : dist_entry_bits(s));
                for (uint32_t j = rev; j < R; j += 1u << len) T[j] = e;
            } else {
                atomicMax(&T[rev & (R - 1)], len);
            }
        }
    }
    wave_sync();
    // subtables: lane l places the long prefixes among root entries
    // [l * R / 64, (l + 1) * R / 64)
    const uint32_t per = R / 64;
    uint32_t mine = 0;
    for (uint32_t j = 0; j < per; ++j) {
        const uint32_t v = T[l * per + j];
        if (is_marker(v, rb)) mine += 1u << (v - rb);
    }
    const uint32_t incl = wave_incl_scan_u32(mine);
    const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
    const bool fits = R + total <= cap;
    uint32_t base = R + incl - mine;
    for (uint32_t j = 0; j < per; ++j) {
        const uint32_t v = T[l * per + j];
        if (is_marker(v, rb)) {
            T[l * per + j] = fits ? (kIndirect | ((v - rb) << 11) | (base << 16)) : 0;
            base += 1u << (v - rb);
        }
    }
    wave_sync();
    if (!fits) return 0;
    const uint32_t ntot = offs[16];
    for (uint32_t k = offs[rb + 1] + l; k < ntot; k += 64) {
        uint32_t len = rb + 1, o = 0, fc = 0;
#pragma unroll
        for (uint32_t q = 1; q < 16; ++q)
            if (q > rb && offs[q] <= k) { len = q; o = offs[q]; fc = first[q]; }
        const uint32_t s = c.sym[k];
        const uint32_t rev = bitrev(fc + (k - o), len);
        const uint32_t root = T[rev & (R - 1)];
        const uint32_t sb = (root >> 11) & 15, sbase = root >> 16;
        const uint32_t e = len | (kind == 0 ? lit_entry_bits(s) : dist_entry_bits(s));
        for (uint32_t j = rev >> rb; j < (1u << sb); j += 1u << (len - rb)) T[sbase + j] = e;
    }
    wave_sync();
    return 0;
}

// One lane's table lookup on the bit pattern b (LSB first): the entry, with
// its code length in [3:0]; 0 where no code is defined.
MDZ_DEV uint32_t lookup2(const uint32_t* T, uint32_t rb, const Code& c, uint32_t kind, uint64_t b) {
    uint32_t e = T[(uint32_t)b & ((1u << rb) - 1)];
    if (e & kIndirect) {
        const uint32_t sb = (e >> 11) & 15;
        e = T[(e >> 16) + ((uint32_t)(b >> rb) & ((1u << sb) - 1))];
    }
    if ((e & 15) == 0) {                        // canonical decode (RFC 1951 3.2.2 order)
        int code = 0, fst = 0, index = 0;
        e = 0;
        for (uint32_t len = 1; len < 16; ++len) {
            code |= (int)(b & 1);
            b >>= 1;
            const int count = c.cnt[len];
            if (code - count < fst) {
                const uint32_t s = c.sym[index + (code - fst)];
                e = len | (kind == 0 ? lit_entry_bits(s) : dist_entry_bits(s));
                break;
            }
            index += count;
            fst += count;
            fst <<= 1;
            code <<= 1;
        }
    }
    return e;
}

// 64 input bits from relative bit position p of the staged round input.
MDZ_DEV uint64_t peek64(const uint32_t* in, uint32_t p) {
    const uint32_t w = p >> 5, sh = p & 31;
    const uint64_t lo = (uint64_t)in[pad16(w)] | ((uint64_t)in[pad16(w + 1)] << 32);
    const uint64_t hi = in[pad16(w + 2)];
    return sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
}

enum : uint32_t { kSymOk = 0, kSymEob = 1, kSymErr = 2 };

// One lane decodes the literal/length (and distance) symbol at relative bit
// p; `lim` is the stream's end in the same coordinates.  Returns kSymOk with
// *sym packed (literal: the byte; match: 1<<31 | (len-3)<<15 | (dist-1)) and
// *used its bits, kSymEob with *used the code's bits, or kSymErr with *sym
// the status, by zlib's checks in zlib's order.
MDZ_DEV uint32_t decode_sym(const InflateLds& S, uint32_t p, uint32_t lim, uint32_t* sym,
                            uint32_t* used) {
    uint64_t b = peek64(S.in, p);
    const uint32_t e = lookup2(S.lit, kLitRoot, S.lit_c, 0, b);
    const uint32_t el = e & 15;
    const uint32_t kind = el ? (e >> 8) & 3 : (uint32_t)kBad;
    uint32_t q = p + el;
    *used = el;
    if (q > lim || (kind == kBad && q + 1 > lim)) { *sym = MDZ_ST_INPUT_OVERRUN; return kSymErr; }
    if (kind == kBad) { *sym = (uint32_t)MDZ_ST_BAD_SYMBOL; return kSymErr; }
    if (kind == kEob) return kSymEob;
    if (kind == kLit) { *sym = e >> 16; return kSymOk; }
    b >>= el;
    const uint32_t xb = (e >> 4) & 15;
    const uint32_t len = (e >> 16) + ((uint32_t)b & ((1u << xb) - 1));
    b >>= xb;
    q += xb;
    const uint32_t d = lookup2(S.dist, kDistRoot, S.dist_c, 1, b);
    const uint32_t dl = d & 15;
    q += dl;
    if (!dl || ((d >> 8) & 3) == kBad) {
        *sym = q + 1 > lim ? (uint32_t)MDZ_ST_INPUT_OVERRUN : (uint32_t)MDZ_ST_BAD_DISTANCE;
        return kSymErr;
    }
    b >>= dl;
    const uint32_t dx = (d >> 4) & 15;
    const uint32_t dist = (d >> 16) + ((uint32_t)b & ((1u << dx) - 1));
    q += dx;
    if (q > lim) { *sym = MDZ_ST_INPUT_OVERRUN; return kSymErr; }
    *used = q - p;
    *sym = 0x80000000u | ((len - 3) << 15) | (dist - 1);
    return kSymOk;
}

struct Reader {
    const uint8_t* src;
    uint64_t n;          // stream bytes
    uint64_t wbyte;      // stream byte offset of the current window's lane-0 dword start (aligned)
    uint32_t win, nxt;   // current and next 256-byte windows (one dword per lane)
    uint32_t wpos;       // next dword of `win` to append
    uint64_t bb;         // bit buffer (LSB first)
    uint32_t nb;         // valid bits in bb
    uint64_t cbits;      // bits consumed since the stream start

    // dword k of the aligned window at absolute address a (bytes outside the
    // stream read as 0, never touched)
    MDZ_DEV uint32_t load_dword(uint64_t a) const {
        const uint64_t s = (uint64_t)(uintptr_t)src, e = s + n;
        if (a >= s && a + 4 <= e) return *reinterpret_cast<const uint32_t*>((uintptr_t)a);
        uint32_t w = 0;
        for (uint32_t k = 0; k < 4; ++k)
            if (a + k >= s && a + k < e) w |= (uint32_t)*reinterpret_cast<const uint8_t*>((uintptr_t)(a + k)) << (8 * k);
        return w;
    }
    MDZ_DEV uint32_t load_window(uint64_t abase) const {
        return load_dword(abase + 4 * (uint64_t)lane_id());
    }
    // positions the reader at stream byte q
    MDZ_DEV void seek(uint64_t q) {
        const uint64_t a = ((uint64_t)(uintptr_t)src + q) & ~3ull;
        win = load_window(a);
        nxt = load_window(a + 256);
        wbyte = a;
        const uint32_t skew = (uint32_t)(((uint64_t)(uintptr_t)src + q) & 3);
        const uint32_t w0 = __builtin_amdgcn_readlane(win, 0);
        bb = w0 >> (8 * skew);
        nb = 32 - 8 * skew;
        wpos = 1;
        cbits = 8 * q;
    }
    // positions the reader at stream bit q
    MDZ_DEV void seek_bit(uint64_t q) {
        seek(q >> 3);
        drop((uint32_t)(q & 7));
    }
    MDZ_DEV void refill() {
        if (nb > 32) return;
        const uint32_t w = __builtin_amdgcn_readlane(win, wpos);
        bb |= (uint64_t)w << nb;
        nb += 32;
        if (++wpos == 64) {
            win = nxt;
            wbyte += 256;
            nxt = load_window(wbyte + 256);
            wpos = 0;
        }
    }
    MDZ_DEV uint32_t peek(uint32_t k) const { return (uint32_t)(bb & ((1ull << k) - 1)); }
    MDZ_DEV void drop(uint32_t k) { bb >>= k; nb -= k; cbits += k; }
    MDZ_DEV uint32_t bits(uint32_t k) {             // k <= 32, after refill
        const uint32_t v = peek(k);
        drop(k);
        return v;
    }
    MDZ_DEV bool over() const { return cbits > 8 * n; }
};

MDZ_DEV uint32_t ring_at(uint64_t p) { return (uint32_t)(p % kRing); }

// The item queue between a stream's two waves.  A wave publishes a count
// (items published, or items finished with) after a workgroup-scope release
// that orders every lane's LDS reads and writes of the item before it; the
// other wave reads the count with an acquire, so once it sees the count it
// sees the item (or may overwrite the item's slot).
MDZ_DEV void q_publish(uint32_t* ctr, uint32_t v) {
    wave_sync();
    if (lane_id() == 0) __hip_atomic_store(ctr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    wave_sync();
}
MDZ_DEV uint32_t q_read(uint32_t* ctr) {
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
}
// Waits until *ctr >= want (mod 2^32: the two lie within 2^31).  False if
// the stream is halted first, or after `limit` polls (the wave then halts
// the stream itself).
MDZ_DEV bool q_wait(uint32_t* ctr, uint32_t want, uint32_t* halt, uint32_t limit) {
    for (uint32_t n = 0;; ++n) {
        const uint32_t v = q_read(ctr);
        if ((int32_t)(v - want) >= 0) break;             // counts run mod 2^32
        if (q_read(halt)) return false;
        if (n >= limit) {
            if (lane_id() == 0) __hip_atomic_store(halt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            wave_sync();
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}

// Trailer checksums over the output, 1 KiB blocks per wave step: lane l holds
// bytes 16l..16l+15 of a block.  A final block of t < 1024 bytes is read as
// 1024 - t zero bytes followed by its t bytes; leading zeros change neither
// the CRC residue nor the Adler-32 sums, and the lane weights stay fixed.
constexpr uint32_t kCkBlock = 1024;

struct CrcTables {
    uint32_t lane[64];       // x^(8 * 16 * (63 - l)): bytes of the block after lane l's
    uint32_t pow2[11];       // x^(8 * 2^j)
};

constexpr CrcTables make_crc_tables() {
    CrcTables t{};
    for (uint32_t l = 0; l < 64; ++l) t.lane[l] = crc_xpow(8ull * 16 * (63 - l));
    for (uint32_t j = 0; j < 11; ++j) t.pow2[j] = crc_xpow(8ull << j);
    return t;
}

__constant__ CrcTables kCrcT = make_crc_tables();


// CRC-32 (crc) or Adler-32 of p[0, n); every lane returns it.
MDZ_DEV uint32_t stream_checksum(const uint8_t* p, uint64_t n, bool crc) {
    const uint32_t l = lane_id();
    const uint32_t lane_x = kCrcT.lane[l];
    const bool aligned = ((uintptr_t)p & 3) == 0;
    uint32_t A = 1, B = 0, R = 0xFFFFFFFFu;
    for (uint64_t b0 = 0; b0 < n; b0 += kCkBlock) {
        const uint32_t t = (uint32_t)min<uint64_t>(n - b0, kCkBlock);
        const uint8_t* q = p + b0 - (kCkBlock - t);     // virtual block start
        uint32_t w[4];
        if (t == kCkBlock && aligned) {
            const uint32_t* q4 = reinterpret_cast<const uint32_t*>(q + 16 * l);
#pragma unroll
            for (int k = 0; k < 4; ++k) w[k] = q4[k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w[k] = 0;
                for (uint32_t j = 0; j < 4; ++j) {
                    const uint32_t v = 16 * l + 4 * k + j;
                    if (v >= kCkBlock - t) w[k] |= (uint32_t)q[v] << (8 * j);
                }
            }
        }
        if (crc) {
            uint32_t r = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) r = crc_dword(r, w[k]);
            r = crc_mulmod(r, lane_x);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) r ^= __shfl_xor(r, o, kWave);
            uint32_t shift = kCrcT.pow2[10];
            if (t != kCkBlock) {
                shift = 0x80000000u;
                for (uint32_t j = 0; j < 10; ++j)
                    if (t & (1u << j)) shift = crc_mulmod(shift, kCrcT.pow2[j]);
            }
            R = crc_mulmod(R, shift) ^ r;
        } else {
            // s2 weights each byte by its distance to the lane's segment end;
            // 16 (63 - l) more bytes of the block follow the segment
            uint32_t s1 = 0, s2 = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t b = (w[k] >> (8 * j)) & 0xFF;
                    s1 += b;
                    s2 += (16 - (4 * k + j)) * b;
                }
            const uint32_t bs1 = wave_sum_u32(s1);
            const uint32_t bs2 = wave_sum_u32(s2 + s1 * 16 * (63 - l));   // < 2^29
            B = (B + t * A + bs2 % kAdlerMod) % kAdlerMod;
            A = (A + bs1) % kAdlerMod;
        }
    }
    return crc ? ~R : (B << 16) | A;
}

// Live work queue (mdz_inflate_live_*): live[0] is the number of published
// items (the host raises it after writing their descriptors), live[1] is set
// once no more will come, live[2] is set here when a worker gives up idle,
// live[3] is the number of polls (about 15 us apart) before it does.  All
// live in host memory the host writes while the kernel runs.
constexpr uint32_t kLivePollsPerMs = 64;

// Waits until item q is published and returns q, or returns ~0u once the
// queue is closed without it, or after live[3] polls (live[2] records that:
// the host then knows items past this point will not be decoded).
MDZ_DEV uint32_t live_wait(uint32_t* live, uint32_t q) {
    const uint32_t polls = __hip_atomic_load(&live[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    for (uint32_t it = 0; it < polls; ++it) {
        if (q < __hip_atomic_load(&live[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM)) return q;
        if (__hip_atomic_load(&live[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM))   // closed: the count is final
            return q < __hip_atomic_load(&live[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) ? q : ~0u;
        for (int k = 0; k < 4; ++k) __builtin_amdgcn_s_sleep(127);   // ~8K cycles each
    }
    __hip_atomic_store(&live[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return ~0u;
}

// Two waves per stream.  Wave 1, the producer, parses block headers, builds
// the tables and decodes the symbols of one round at a time into the LDS
// item queue; wave 0, the consumer, turns each published item (a round's
// list, or a stored block) into output in the LDS window and stores it to
// HBM.  The queue holds kSlots items, so the producer may decode up to two
// rounds ahead of the one being written out; the waves meet at a workgroup
// barrier only when a stream starts and when it ends.
__global__ __launch_bounds__(128) void inflate_kernel(const mdz_inflate_in_t* __restrict__ ins,
                                                      mdz_inflate_status_t* __restrict__ outs,
                                                      const uint32_t* __restrict__ order,
                                                      uint32_t count, int parse_hdr,
                                                      uint32_t* __restrict__ cursor,
                                                      mdz_inflate_cost_t* __restrict__ cost,
                                                      uint32_t* live, uint32_t wait_polls) {
    extern __shared__ __align__(16) uint8_t smem[];
    InflateLds& S = *reinterpret_cast<InflateLds*>(smem);
    HdrScratch& H = S.u.h;
    const uint32_t l = lane_id();
    const bool producer = threadIdx.x >= 64;

    for (;;) {
        // next stream: one per block (plain) or from the shared cursor (persistent)
        if (threadIdx.x == 0) {
            const uint32_t q = cursor ? atomicAdd(cursor, 1u) : blockIdx.x;
            const uint32_t lq = live && q < count ? live_wait(live, q) : q;
            if (live && lq < count) {
                // the host wrote this descriptor after the kernel started: read
                // it past every cache with system-scope loads, after the acquire
                // that saw it published (a plain uniform load could come from
                // the scalar cache, filled earlier from the same line)
                const uint64_t* d = reinterpret_cast<const uint64_t*>(ins + lq);
                for (int k = 0; k < 4; ++k)
                    S.ctrl.desc[k] = __hip_atomic_load(d + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            S.ctrl.qi = lq;
            S.ctrl.halt = 0;
            S.ctrl.n_pub = 0;
            S.ctrl.n_fin = 0;
            S.ctrl.sym_fin = 0;
        }
        __syncthreads();
        const uint32_t qi = S.ctrl.qi;
        if (qi >= count) return;
        const uint32_t si = order ? order[qi] : qi;
        if (si >= count) {                         // a bad order entry names no stream
            if (!cursor) return;
            __syncthreads();
            continue;
        }
        mdz_inflate_in_t in;
        if (live) {
            in.src = reinterpret_cast<const void*>(S.ctrl.desc[0]);
            in.src_bytes = S.ctrl.desc[1];
            in.dst = reinterpret_cast<void*>(S.ctrl.desc[2]);
            in.dst_bytes = S.ctrl.desc[3];
        } else {
            in = ins[si];
        }
        const int hdr = parse_hdr & 3;

        if (producer) {
            // ------------------------------------------------------ producer
            Reader r;
            r.src = reinterpret_cast<const uint8_t*>(in.src);
            r.n = in.src_bytes;
            int st = MDZ_ST_OK;
            if (hdr == MDZ_HDR_GZIP) {
                // the gzip prepass left the header length (or -1) in the status
                // slot; as cuDF, the footer is cut off and a rejected header
                // reports -1 with nothing written
                const mdz_inflate_status_t pre = outs[si];
                r.n = r.n >= 8 ? r.n - 8 : 0;
                if (pre.status != MDZ_ST_OK) {
                    st = MDZ_ST_BAD_BLOCK_TYPE;
                } else {
                    r.src += pre.bytes_written;
                    r.n -= pre.bytes_written;
                }
            }
            r.seek(0);
            if (st == MDZ_ST_OK && hdr == MDZ_HDR_ZLIB) {
                r.refill();
                const uint32_t cmf = r.bits(8), flg = r.bits(8);
                if (r.over()) st = MDZ_ST_INPUT_OVERRUN;
                else if ((cmf & 15) != 8 || (cmf >> 4) > 7 || (flg & 0x20) || ((cmf << 8) | flg) % 31)
                    st = MDZ_ST_BAD_HEADER;
            }
            uint64_t P = r.cbits;                  // next bit of the Deflate data
            INF_PHASE_START();
            bool need_hdr = true, final = false, pdone = false;
            uint32_t fixed_built = 0, n_stored = 0, n_fixed = 0, n_dyn = 0;
            uint32_t sym_pos = 0, sym_at = 0;      // symbols listed (running count) and their ring end
            const uintptr_t src_addr = (uintptr_t)r.src;
            for (uint32_t ik = 0; !pdone; ++ik) {
                // item ik goes to slot ik % kSlots once the consumer has finished
                // with item ik - kSlots; a halted stream gets no more items
                if (q_read(&S.ctrl.halt)) break;
                INF_PHASE(11);
                const bool free = q_wait(&S.ctrl.n_fin, ik >= kSlots ? ik + 1 - kSlots : 0, &S.ctrl.halt, wait_polls);
                INF_PHASE(9);                      // producer waiting for the consumer
                if (!free) break;
                {
                    Item it = {kItemNone, 0, MDZ_ST_OK, 0, 0, 0};
                    const uint32_t b = ik % kSlots;
                    if (st != MDZ_ST_OK) {
                        it.status = st;
                        it.last = 1;
                    } else if (need_hdr) {
                        r.seek_bit(P);
                        r.refill();
                        final = r.bits(1);
                        const uint32_t type = r.bits(2);
                        if (r.over()) st = MDZ_ST_INPUT_OVERRUN;
                        else if (type == 3) st = MDZ_ST_BAD_BLOCK_TYPE;
                        else if (type == 0) {                           // stored
                            ++n_stored;
                            r.drop(r.nb & 7);
                            r.refill();
                            const uint32_t len = r.bits(16), nlen = r.bits(16);
                            if (r.over()) st = MDZ_ST_INPUT_OVERRUN;
                            else if (len != (~nlen & 0xffffu)) st = MDZ_ST_STORED_LEN;
                            else {
                                const uint64_t q = r.cbits / 8;
                                const uint64_t avail = q <= r.n ? r.n - q : 0;
                                it.kind = kItemStored;
                                it.src = q;
                                it.take = min<uint64_t>(len, avail);
                                if (it.take < len) { it.status = MDZ_ST_INPUT_OVERRUN; it.last = 1; }
                                else if (final) it.last = 1;
                                P = 8 * (q + len);
                            }
                        } else if (type == 1) {                         // fixed tables
                            ++n_fixed;
                            if (!fixed_built) {
                                for (uint32_t k = l; k < 320; k += 64)
                                    H.lens[k] = k < 144 ? 8 : k < 256 ? 9 : k < 280 ? 7 : k < 288 ? 8 : 5;
                                wave_sync();
                                build_table2(S.lit_c, S.lit, kLitRoot, kLitCap, H.lens, 288, 0);
                                build_table2(S.dist_c, S.dist, kDistRoot, kDistCap, H.lens + 288, 32, 1);
                                fixed_built = 1;
                            }
                            need_hdr = false;
                            P = r.cbits;
                        } else {                                        // dynamic tables
                            ++n_dyn;
                            fixed_built = 0;
                            r.refill();
                            const uint32_t nlen = r.bits(5) + 257, ndist = r.bits(5) + 1, ncode = r.bits(4) + 4;
                            if (r.over()) st = MDZ_ST_INPUT_OVERRUN;
                            else if (nlen > 286 || ndist > 30) st = MDZ_ST_BAD_COUNTS;
                            if (st == MDZ_ST_OK) {
                                for (uint32_t k = l; k < 19; k += 64) H.lens[k] = 0;
                                wave_sync();
                                for (uint32_t i = 0; i < ncode; ++i) {
                                    r.refill();
                                    const uint32_t v = r.bits(3);
                                    if (l == 0) H.lens[kClOrder[i]] = (uint8_t)v;
                                }
                                wave_sync();
                                if (r.over()) st = MDZ_ST_INPUT_OVERRUN;
                                else if (build_code(H.cl, H.cl_lut, 7, H.lens, 19, 2)) st = MDZ_ST_BAD_CODE_LENS;
                            }
                            if (st == MDZ_ST_OK) {
                                // code lengths for nlen + ndist symbols (lens[] reused from 0)
                                uint32_t idx = 0, prev = 0;
                                while (idx < nlen + ndist) {
                                    r.refill();
                                    const uint32_t e = __builtin_amdgcn_readfirstlane(H.cl_lut[r.peek(7)]);
                                    r.drop(e & 15);
                                    const uint32_t sym = e >> 16;
                                    uint32_t rep = 0, val = 0;
                                    if (sym < 16) { val = sym; rep = 1; prev = sym; }
                                    else if (sym == 16) { val = prev; rep = 3 + r.bits(2); }
                                    else if (sym == 17) { rep = 3 + r.bits(3); prev = 0; }
                                    else { rep = 11 + r.bits(7); prev = 0; }
                                    if (r.over()) { st = MDZ_ST_INPUT_OVERRUN; break; }
                                    if (sym == 16 && idx == 0) { st = MDZ_ST_BAD_REPEAT; break; }
                                    if (idx + rep > nlen + ndist) { st = MDZ_ST_BAD_REPEAT; break; }
                                    for (uint32_t j = l; j < rep; j += 64) {
                                        const uint32_t t = idx + j;
                                        H.lens[t < nlen ? t : 288 + (t - nlen)] = (uint8_t)val;
                                    }
                                    idx += rep;
                                }
                            }
                            if (st == MDZ_ST_OK) {
                                for (uint32_t k = nlen + l; k < 288; k += 64) H.lens[k] = 0;
                                for (uint32_t k = 288 + ndist + l; k < 320; k += 64) H.lens[k] = 0;
                                wave_sync();
                                if (H.lens[256] == 0) st = MDZ_ST_NO_EOB;
                                else if (build_table2(S.lit_c, S.lit, kLitRoot, kLitCap, H.lens, nlen, 0))
                                    st = MDZ_ST_BAD_LITLEN;
                                else if (build_table2(S.dist_c, S.dist, kDistRoot, kDistCap, H.lens + 288, ndist, 1))
                                    st = MDZ_ST_BAD_DIST;
                            }
                            need_hdr = false;
                            P = r.cbits;
                        }
                        if (st != MDZ_ST_OK) { it.status = st; it.last = 1; }
                    }
                    INF_PHASE(0);
                    if (it.kind == kItemNone && !it.last) {
                        // ---- one round of symbols.  Lane t takes the kSpan input bits
                        // from R + t*kSpan (R: the round's first true symbol) and
                        // decodes from there as if a symbol started at its first bit,
                        // noting each symbol start it meets (an end-of-block or
                        // invalid code restarts its decode one bit later).  Huffman
                        // decoding resynchronises: the true symbol sequence, followed
                        // past the end of lane t's span, soon meets a start a later
                        // lane noted, and from there that lane's sequence is the true
                        // one.  The lanes the true sequence passes through form a
                        // chain; each lists its piece of the sequence, in order.
                        const uint64_t byte = P >> 3;
                        const uintptr_t A = (src_addr + byte) & ~(uintptr_t)3;
                        const uint32_t rel = (uint32_t)(8 * (src_addr + byte - A) + (P & 7));
                        const uint64_t limb = 8 * (uint64_t)(src_addr + r.n - A);
                        const uint32_t lim = limb > 0xF0000000ull ? 0xF0000000u : (uint32_t)limb;
                        const uint64_t s_lo = (uint64_t)(uintptr_t)r.src;
                        if (A >= s_lo && A + 4ull * kStageWords <= s_lo + r.n) {
                            // the whole round inside the stream: every load issued
                            // before the first LDS store (one memory latency a round)
                            constexpr uint32_t kPer = (kStageWords + 63) / 64;
                            uint32_t v[kPer];
#pragma unroll
                            for (uint32_t i = 0; i < kPer; ++i) {
                                const uint32_t k = l + 64 * i;
                                v[i] = k < kStageWords ? *reinterpret_cast<const uint32_t*>(A + 4ull * k) : 0u;
                            }
#pragma unroll
                            for (uint32_t i = 0; i < kPer; ++i) {
                                const uint32_t k = l + 64 * i;
                                if (k < kStageWords) S.in[pad16(k)] = v[i];
                            }
                        } else {
                            for (uint32_t k = l; k < kStageWords; k += 64) S.in[pad16(k)] = r.load_dword(A + 4 * k);
                        }
                        wave_sync();
                        INF_PHASE(5);

                        // phase 1: each lane's own span (positions past the stream end
                        // start no symbol; the true path meets them only as an overrun)
                        const uint32_t s_me = rel + kSpan * l, s_nx = s_me + kSpan;
                        uint32_t bm[kSpan / 32];
#pragma unroll
                        for (uint32_t k = 0; k < kSpan / 32; ++k) bm[k] = 0;
                        uint32_t p = s_me, sym = 0, used = 0;
                        uint32_t evp[kEvents], nev = 0;  // positions of the first kEvents events (restarts)
#pragma unroll
                        for (uint32_t k = 0; k < kEvents; ++k) evp[k] = ~0u;
                        bool at_end = false;
                        while (p < s_nx) {
                            if (p >= lim) { at_end = true; break; }
                            const uint32_t k = p - s_me;
#pragma unroll
                            for (uint32_t w = 0; w < kSpan / 32; ++w)
                                if ((k >> 5) == w) bm[w] |= 1u << (k & 31);
                            const uint32_t rc = decode_sym(S, p, lim, &sym, &used);
                            if (rc != kSymOk) {
#pragma unroll
                                for (uint32_t k = 0; k < kEvents; ++k)
                                    if (nev == k) evp[k] = p;
                                ++nev;
                            }
                            p += rc == kSymOk ? used : 1;
                        }
                        if (p < s_nx) p = s_nx;
#pragma unroll
                        for (uint32_t k = 0; k < kSpan / 32; ++k) S.u.bm[pad16((kSpan / 32) * l + k)] = bm[k];
                        wave_sync();
                        INF_PHASE(6);

                        // phase 2: hand-off.  From its span's end the lane follows its
                        // sequence until it meets a start noted by a later lane g (in
                        // the span it has reached, at most kReach spans ahead).  stop:
                        // the sequence has its end of block or an error at h; lost: no
                        // noted start met in reach.  Lane 63 hands off to the next round.
                        uint32_t h = p, g = 64, steps = 0;
                        bool stop = false, lost = false;
                        if (l < 63) {
                            for (;;) {
                                const uint32_t j = (h - rel) >> kSpanLog;
                                if (j > l + kReach || j > 63) { lost = true; break; }
                                const uint32_t k = h - (rel + kSpan * j);
                                if ((S.u.bm[pad16((kSpan / 32) * j + (k >> 5))] >> (k & 31)) & 1) { g = j; break; }
                                const uint32_t rc = decode_sym(S, h, lim, &sym, &used);
                                if (rc != kSymOk) { stop = true; break; }
                                h += used;
                                ++steps;
                            }
                        }
                        INF_PHASE(7);
                        // the true chain: lane 0, then g(0), g(g(0)), ...; each member's
                        // piece starts where its predecessor handed off
                        uint32_t a = rel;
                        uint64_t chain = 0;
                        for (uint32_t cur = 0;;) {
                            chain |= 1ull << cur;
                            const uint32_t nx = __builtin_amdgcn_readlane(g, cur);
                            if (nx > 63) break;
                            const uint32_t hc = __builtin_amdgcn_readlane(h, cur);
                            if (l == nx) a = hc;
                            cur = nx;
                        }
                        const bool member = (chain >> l) & 1;

                        // phase 3: count each member's piece [a, h) and find the event
                        // that ends it, if any.  From a (one of its noted starts) the
                        // lane's own sequence is the true one, so up to the first event
                        // at or after a its noted starts are the piece's symbols, and
                        // past its span the hand-off steps are.  Events come in
                        // position order, so the first tracked one at or after a is
                        // that event; a lane whose kEvents tracked events all lie
                        // before a (and that met more), or that met the stream's end,
                        // decodes its piece to count it.
                        uint32_t cnt = 0, ev = kSymOk, evst = 0, evpos = 0, evlen = 0;
                        if (member) {
                            uint32_t x = ~0u;
#pragma unroll
                            for (uint32_t k = kEvents; k-- > 0;)
                                if (evp[k] >= a) x = evp[k];
                            bool need_ev = false;
                            if (at_end || (nev > kEvents && x == ~0u)) {
                                p = a;
                                while (p < h) {
                                    const uint32_t rc = decode_sym(S, p, lim, &sym, &used);
                                    if (rc != kSymOk) { ev = rc; evst = sym; evpos = p; evlen = used; break; }
                                    ++cnt;
                                    p += used;
                                }
                                if (ev == kSymOk && p != h) { ev = kSymErr; evst = (uint32_t)MDZ_ST_INTERNAL; evpos = p; }
                                if (ev == kSymOk && stop) { evpos = h; need_ev = true; }
                            } else {
                                // noted starts in [a, min(x, span end))
                                const uint32_t lo = a - s_me, hi = x < s_nx ? x - s_me : kSpan;
#pragma unroll
                                for (uint32_t w = 0; w < kSpan / 32; ++w) {
                                    const uint32_t b0 = 32 * w;
                                    uint32_t m = bm[w];
                                    if (lo > b0) m = lo - b0 >= 32 ? 0u : m & (~0u << (lo - b0));
                                    if (hi < b0 + 32) m = hi <= b0 ? 0u : m & (~0u >> (b0 + 32 - hi));
                                    cnt += __popc(m);
                                }
                                if (x < s_nx) {
                                    evpos = x;
                                    need_ev = true;
                                } else {
                                    cnt += steps;
                                    if (stop) { evpos = h; need_ev = true; }
                                }
                            }
                            if (need_ev) {
                                ev = decode_sym(S, evpos, lim, &sym, &used);
                                evst = sym;
                                evlen = used;
                            }
                        }
                        // the round ends at the first member with an event or a lost hand-off
                        const uint64_t endm = __ballot(member && (ev != kSymOk || lost));
                        const uint32_t top = 63 - (uint32_t)__builtin_clzll(chain);
                        uint32_t last = endm ? (uint32_t)__builtin_ctzll(endm) : top;
                        const uint32_t incl = wave_incl_scan_u32(member && l <= last ? cnt : 0);
                        // keep whole pieces within the symbol list's capacity
                        const uint64_t over = __ballot(member && l <= last && incl > kSymCap);
                        const bool cut = over != 0;
                        if (cut) {   // the last member below the first that overflows (lane 0's piece always fits)
                            const uint64_t below = chain & ((1ull << __builtin_ctzll(over)) - 1);
                            last = 63 - (uint32_t)__builtin_clzll(below);
                        }
                        const uint32_t nsym = __builtin_amdgcn_readlane(incl, last);
                        const uint32_t r_ev = cut ? kSymOk : __builtin_amdgcn_readlane(ev, last);
                        const uint32_t r_evst = __builtin_amdgcn_readlane(evst, last);
                        const uint32_t r_evpos = __builtin_amdgcn_readlane(evpos, last);
                        const uint32_t r_evlen = __builtin_amdgcn_readlane(evlen, last);
                        const uint32_t r_h = __builtin_amdgcn_readlane(h, last);
                        INF_PHASE(8);

                        // phase 4: list the true symbols in order, once the ring has
                        // room for them (the consumer frees a finished item's)
                        const bool room = q_wait(&S.ctrl.sym_fin, sym_pos + nsym - kSymRing, &S.ctrl.halt, wait_polls);
                        INF_PHASE(9);
                        bool astray = false;
                        if (room && member && l <= last) {
                            uint32_t o = sym_wrap(sym_at + incl - cnt);
                            p = a;
                            for (uint32_t i = 0; i < cnt; ++i) {
                                decode_sym(S, p, lim, &sym, &used);
                                S.sym[o] = sym;
                                o = o + 1 == kSymRing ? 0 : o + 1;
                                p += used;
                            }
                            // the listed symbols must end where the piece does
                            astray = p != (ev != kSymOk ? evpos : h);
                        }
                        it.kind = kItemRound;
                        it.nsym = nsym;
                        it.sym0 = sym_at;
                        if (!room) {                            // halted, or the wait ran out
                            it.nsym = 0;
                            st = MDZ_ST_INTERNAL;
                            it.status = st;
                            it.last = 1;
                        } else if (__ballot(astray)) {          // never expected
                            it.nsym = 0;
                            st = MDZ_ST_INTERNAL;
                            it.status = st;
                            it.last = 1;
                        } else if (r_ev == kSymEob) {
                            P += (r_evpos + r_evlen) - rel;
                            need_hdr = true;
                            if (final) it.last = 1;
                        } else if (r_ev == kSymErr) {
                            st = (int)r_evst;
                            it.status = st;
                            it.last = 1;
                            P += r_evpos - rel;
                        } else {
                            P += r_h - rel;
                        }
                        sym_pos += it.nsym;
                        sym_at = sym_wrap(sym_at + it.nsym);
                        INF_PHASE(1);
                    }
                    if (l == 0) S.item[b] = it;
                    if (it.last) pdone = true;
                    q_publish(&S.ctrl.n_pub, ik + 1);
                }
            }
            // block counts and Deflate input consumed, for the cost log and the
            // trailer check; the consumer reads them after the closing barrier
            if (l == 0) {
                S.ctrl.n_stored = n_stored;
                S.ctrl.n_fixed = n_fixed;
                S.ctrl.n_dyn = n_dyn;
                S.ctrl.consumed = P;
            }
            __syncthreads();                       // the stream's closing barrier
        } else {
            // ------------------------------------------------------ consumer
            uint8_t* const dst = reinterpret_cast<uint8_t*>(in.dst);
            const uint64_t cap = in.dst_bytes;
            uint64_t pos = 0, flushed = 0;             // output produced / flushed to dst
            uint32_t n_sym = 0, n_match = 0;
            uint32_t sym_done = 0;                     // symbols of the items finished (mod 2^32)
            int st = MDZ_ST_OK;
            const uint64_t t0 = clock64();

            // flushes [flushed, upto) in dword stores at dword-aligned dst
            // addresses; before the last flush it stops at such an address, so
            // only the first flush has head bytes and only the last tail bytes
            const uint32_t dmis = (uint32_t)((uintptr_t)dst & 3);
            auto flush = [&](uint64_t upto, bool all) {
                uint64_t lim = min(upto, cap);
                if (!all) {
                    const uint64_t cut = (lim + dmis) & 3;
                    lim = lim >= cut ? lim - cut : 0;
                }
                if (lim <= flushed) return;
                // at most one batch (< kRing bytes) is pending: ring indices are
                // the flush start's plus an offset, wrapped once; kRing % 4 == 0,
                // so ring index and output offset agree mod 4
                const uint32_t rf = ring_at(flushed);
                auto at = [&](uint64_t p) {
                    const uint32_t x = rf + (uint32_t)(p - flushed);
                    return x >= kRing ? x - kRing : x;
                };
                const uint64_t h = min(lim, flushed + ((4 - ((dmis + flushed) & 3)) & 3));
                for (uint64_t p = flushed + l; p < h; p += 64) dst[p] = S.ring[at(p)];
                const uint64_t nw = (lim - h) / 4;
                if (dmis == 0) {                       // h % 4 == 0: aligned ring dwords
                    for (uint64_t w = l; w < nw; w += 64) {
                        const uint64_t p = h + 4 * w;
                        *reinterpret_cast<uint32_t*>(dst + p) = *reinterpret_cast<const uint32_t*>(&S.ring[at(p)]);
                    }
                } else {                               // ring bytes gathered into aligned dst dwords
                    for (uint64_t w = l; w < nw; w += 64) {
                        const uint64_t p = h + 4 * w;
                        *reinterpret_cast<uint32_t*>(dst + p) =
                            (uint32_t)S.ring[at(p)] | ((uint32_t)S.ring[at(p + 1)] << 8) |
                            ((uint32_t)S.ring[at(p + 2)] << 16) | ((uint32_t)S.ring[at(p + 3)] << 24);
                    }
                }
                for (uint64_t p = h + 4 * nw + l; p < lim; p += 64) dst[p] = S.ring[at(p)];
                flushed = lim;
            };

            const uint8_t* src = reinterpret_cast<const uint8_t*>(in.src);
            if (hdr == MDZ_HDR_GZIP && outs[si].status == MDZ_ST_OK) src += outs[si].bytes_written;
            INF_PHASE_START();
            for (uint32_t ik = 0;; ++ik) {
                INF_PHASE(12);
                const bool have = q_wait(&S.ctrl.n_pub, ik + 1, &S.ctrl.halt, wait_polls);
                INF_PHASE(10);                     // consumer waiting for the producer
                // the producer halts a stream only when a wait of its own ran out
                if (!have) { st = MDZ_ST_INTERNAL; break; }
                {
                    const uint32_t b = ik % kSlots;
                    const Item it = S.item[b];
                    if (it.kind == kItemStored) {
                        const uint64_t q = it.src, take = it.take;
                        for (uint64_t done = 0; done < take;) {
                            const uint64_t k = min<uint64_t>(take - done, kBatchOut);
                            // lane l moves bytes [32l, 32l + 32) of the piece: all its
                            // loads are issued before the first store
                            const uint8_t* const from = src + q + done;
                            const uint32_t rb = ring_at(pos);
                            // (a byte past the piece loads byte 0 and is stored to the
                            // lane's sink: no branch per byte)
                            uint8_t* const sink = reinterpret_cast<uint8_t*>(&S.sink[l]);
                            uint8_t v[32];
#pragma unroll
                            for (uint32_t j = 0; j < 32; ++j) {
                                const uint32_t o = 32 * l + j;
                                v[j] = from[o < k ? o : 0];
                            }
#pragma unroll
                            for (uint32_t j = 0; j < 32; ++j) {
                                const uint32_t o = 32 * l + j, x = rb + o;
                                uint8_t* const t = o < k ? &S.ring[x >= kRing ? x - kRing : x] : sink;
                                *t = v[j];
                            }
                            wave_sync();
                            const uint64_t room = pos < cap ? cap - pos : 0;
                            pos += min(k, room);
                            flush(pos, false);
                            wave_sync();
                            done += k;
                            if (k > room) { st = MDZ_ST_OUTPUT_OVERRUN; break; }
                        }
                        INF_PHASE(4);
                    } else if (it.kind == kItemRound) {
                        // ---- emit the list in batches (at most 64 symbols or about
                        // kBatchOut bytes each): literals, then matches in order.  The
                        // consumer is one wave and its LDS operations execute in order,
                        // so a match reads what earlier literals and matches of the
                        // batch stored without a wait in between; ring positions are
                        // kept below 2 * kRing and wrapped with one compare.
                        const uint32_t nsym = it.nsym;
                        for (uint32_t b0 = 0; b0 < nsym && st == MDZ_ST_OK;) {
                            const uint32_t i = b0 + l;
                            const uint32_t x = i < nsym ? S.sym[sym_wrap(it.sym0 + i)] : 0;
                            const bool have = i < nsym, ref = have && (x >> 31);
                            const uint32_t rec_len = !have ? 0 : ref ? ((x >> 15) & 255) + 3 : 1;
                            const uint32_t rec_val = ref ? (x & 32767) + 1 : (x & 255);
                            const uint32_t oin = wave_incl_scan_u32(rec_len);
                            const uint32_t rec_off = oin - rec_len;
                            // batch rules, in stream order: a symbol joins while fewer than
                            // kBatchOut bytes are queued; a distance past the output start
                            // is an error before the symbol; an overrun ends the batch with
                            // the overrunning symbol included
                            const uint64_t full = __ballot(have && rec_off >= kBatchOut);
                            const uint64_t bad = __ballot(ref && rec_val > pos + rec_off);
                            const uint64_t ovr = __ballot(have && pos + oin > cap);
                            uint32_t n = min(nsym - b0, 64u);
                            if (full) n = min(n, (uint32_t)__builtin_ctzll(full));
                            if (ovr) n = min(n, (uint32_t)__builtin_ctzll(ovr) + 1);
                            if (bad && (uint32_t)__builtin_ctzll(bad) < n) {
                                n = (uint32_t)__builtin_ctzll(bad);
                                st = MDZ_ST_BAD_DISTANCE;
                            }
                            const uint32_t bout = n ? __builtin_amdgcn_readlane(oin, n - 1) : 0;
                            n_sym += n;
                            n_match += __popcll(__ballot(ref && l < n));
                            const uint64_t room = pos < cap ? cap - pos : 0;
                            const uint32_t rb = ring_at(pos);          // ring index of the batch start
                            auto wrap = [](uint32_t v) { return v >= kRing ? v - kRing : v; };
                            // each lane's ring destination, and for a match its source
                            // start (dist <= 32768 < kRing), computed for all symbols at
                            // once so the serial match loop only reads them
                            const uint32_t my_d0 = wrap(rb + rec_off);
                            const uint32_t my_s0 = wrap(my_d0 + kRing - (ref ? rec_val : 1u));
                            if (l < n && !ref) S.ring[my_d0] = (uint8_t)rec_val;
                            __builtin_amdgcn_wave_barrier();
                            uint64_t mm = __ballot(l < n && ref);
                            while (mm) {
                                const uint32_t m = __builtin_ctzll(mm);
                                mm &= mm - 1;
                                const uint32_t len = __builtin_amdgcn_readlane(rec_len, m);
                                const uint32_t dist = __builtin_amdgcn_readlane(rec_val, m);
                                const uint32_t d0 = __builtin_amdgcn_readlane(my_d0, m);   // destination start
                                const uint32_t s0 = __builtin_amdgcn_readlane(my_s0, m);   // source start
                                if (dist >= len) {
                                    for (uint32_t j = l; j < len; j += 64) S.ring[wrap(d0 + j)] = S.ring[wrap(s0 + j)];
                                } else {
                                    // overlapping copy: byte j repeats source byte j mod dist;
                                    // j < 258, so a float reciprocal and one correction step
                                    // give the exact quotient
                                    const float rd = __builtin_amdgcn_rcpf((float)dist);
                                    for (uint32_t j = l; j < len; j += 64) {
                                        uint32_t qj = (uint32_t)((float)j * rd);
                                        if (qj * dist > j) --qj;
                                        else if ((qj + 1) * dist <= j) ++qj;
                                        S.ring[wrap(d0 + j)] = S.ring[wrap(s0 + (j - qj * dist))];
                                    }
                                }
                                __builtin_amdgcn_wave_barrier();
                            }
                            if (bout > room) {
                                pos += room;
                                flush(pos, true);
                                st = MDZ_ST_OUTPUT_OVERRUN;
                                break;
                            }
                            pos += bout;
                            flush(pos, false);
                            wave_sync();
                            b0 += n;
                        }
                        INF_PHASE(2);
                    }
                    if (st == MDZ_ST_OK && it.status != MDZ_ST_OK) st = it.status;
                    if (st != MDZ_ST_OK || it.last) {
                        // the stream is finished: a producer still decoding, or
                        // waiting for a free slot, stops
                        if (l == 0) __hip_atomic_store(&S.ctrl.halt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        wave_sync();
                        break;
                    }
                    // the item's symbols are free for the producer's next lists
                    sym_done += it.nsym;
                    wave_sync();
                    if (l == 0) {
                        __hip_atomic_store(&S.ctrl.sym_fin, sym_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        __hip_atomic_store(&S.ctrl.n_fin, ik + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    wave_sync();
                }
            }
            flush(pos, true);
            __syncthreads();                       // the stream's closing barrier
            // a live queue's host reads the status while the kernel runs: the
            // output (and the status fields) must reach memory first
            if (live) __threadfence_system();
            if (l == 0) {
                mdz_inflate_status_t o;
                o.bytes_written = pos;
                if (cost) {
                    mdz_inflate_cost_t c;
                    c.cycles = clock64() - t0;
                    c.symbols = n_sym;
                    c.matches = n_match;
                    c.stored_blocks = S.ctrl.n_stored;
                    c.fixed_blocks = S.ctrl.n_fixed;
                    c.dynamic_blocks = S.ctrl.n_dyn;
                    c.reserved = 0;
                    cost[si] = c;
                }
                o.status = (uint32_t)st;
                // Deflate input consumed through the last block (from the start
                // of the Deflate data), for the trailer check; saturates
                o.reserved = (uint32_t)min<uint64_t>((S.ctrl.consumed + 7) / 8, 0xFFFFFFFFull);
                if (live) {
                    outs[si].bytes_written = o.bytes_written;
                    outs[si].reserved = o.reserved;
                    __threadfence_system();
                    __hip_atomic_store(&outs[si].status, o.status, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                } else {
                    outs[si] = o;
                }
            }
        }
        if (!cursor) return;
    }
}

__global__ void reset_cursor_kernel(uint32_t* c) { *c = 0; }

// gzip framing, before the decode: each stream's header length by cuDF's
// rules, left in its status slot (bytes_written = length, status = 0, or
// status = -1 for a rejected header) for inflate_kernel to read.  Kept out of
// the decode kernel: any framing code there slowed its symbol loop.
__global__ __launch_bounds__(256) void gzip_prepass_kernel(const mdz_inflate_in_t* __restrict__ ins,
                                                           mdz_inflate_status_t* __restrict__ outs, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const mdz_inflate_in_t in = ins[i];
    const int64_t h = gzip_header_len(reinterpret_cast<const uint8_t*>(in.src), in.src_bytes);
    mdz_inflate_status_t o;
    o.bytes_written = h < 0 ? 0 : (uint64_t)h;
    o.status = h < 0 ? (uint32_t)MDZ_ST_BAD_BLOCK_TYPE : (uint32_t)MDZ_ST_OK;
    o.reserved = 0;
    outs[i] = o;
}

// Trailer check, after the decode, one wave per stream that decoded with
// status 0: the zlib Adler-32 (big-endian, right after the Deflate data,
// whose length inflate_kernel left in `reserved`) or the gzip CRC-32 and
// ISIZE (the member's last 8 bytes) against the output written.
__global__ __launch_bounds__(64) void trailer_check_kernel(const mdz_inflate_in_t* __restrict__ ins,
                                                           mdz_inflate_status_t* __restrict__ outs, int hdr) {
    const uint32_t si = blockIdx.x;
    const mdz_inflate_status_t o = outs[si];
    if (o.status != (uint32_t)MDZ_ST_OK) return;
    const mdz_inflate_in_t in = ins[si];
    const uint8_t* s0 = reinterpret_cast<const uint8_t*>(in.src);
    const uint8_t* dst = reinterpret_cast<const uint8_t*>(in.dst);
    int st = MDZ_ST_OK;
    if (hdr == MDZ_HDR_ZLIB) {
        const uint64_t tb = (uint64_t)o.reserved;   // the 2-byte header counts in the Deflate reader
        if (o.reserved == 0xFFFFFFFFu || tb + 4 > in.src_bytes) {
            st = MDZ_ST_INPUT_OVERRUN;
        } else {
            const uint32_t want = ((uint32_t)s0[tb] << 24) | ((uint32_t)s0[tb + 1] << 16) |
                                  ((uint32_t)s0[tb + 2] << 8) | s0[tb + 3];
            if (stream_checksum(dst, o.bytes_written, false) != want) st = MDZ_ST_BAD_CHECKSUM;
        }
    } else {
        const uint8_t* f = s0 + in.src_bytes - 8;
        const uint32_t crc = f[0] | ((uint32_t)f[1] << 8) | ((uint32_t)f[2] << 16) | ((uint32_t)f[3] << 24);
        const uint32_t isz = f[4] | ((uint32_t)f[5] << 8) | ((uint32_t)f[6] << 16) | ((uint32_t)f[7] << 24);
        if (isz != (uint32_t)o.bytes_written || stream_checksum(dst, o.bytes_written, true) != crc)
            st = MDZ_ST_BAD_CHECKSUM;
    }
    if (lane_id() == 0 && st != MDZ_ST_OK) outs[si].status = (uint32_t)st;
}

// Stored streams: the first min(src_bytes, dst_bytes) bytes of each stream
// go to its output slot (cuDF's copy_uncompressed_kernel, reference
// include/cudf_deflate.cuh:1349-1397).  kCopySlices workgroups share a
// stream; each copies a contiguous range of 16-byte (or dword, or byte)
// units, by the widest unit at which src and dst are co-aligned.
constexpr uint32_t kCopySlices = 16, kCopyThreads = 256;

MDZ_DEV void copy_one(const mdz_inflate_in_t in, uint32_t slice) {
    const uint64_t n = min(in.src_bytes, in.dst_bytes);
    const uint8_t* src = reinterpret_cast<const uint8_t*>(in.src);
    uint8_t* dst = reinterpret_cast<uint8_t*>(in.dst);
    const uintptr_t mis = (uintptr_t)src ^ (uintptr_t)dst;
    const uint32_t unit = (mis & 15) == 0 ? 16 : (mis & 3) == 0 ? 4 : 1;
    const uint64_t head = min<uint64_t>((unit - ((uintptr_t)dst & (unit - 1))) & (unit - 1), n);
    const uint64_t units = (n - head) / unit;
    const uint64_t per = (units + kCopySlices - 1) / kCopySlices;
    const uint64_t u0 = min(units, per * slice), u1 = min(units, u0 + per);
    const uint32_t t = threadIdx.x;
    if (slice == 0)
        for (uint64_t i = t; i < head; i += kCopyThreads) dst[i] = src[i];
    if (slice == kCopySlices - 1)
        for (uint64_t i = head + units * unit + t; i < n; i += kCopyThreads) dst[i] = src[i];
    if (unit == 16) {
        const uint4* s4 = reinterpret_cast<const uint4*>(src + head);
        uint4* d4 = reinterpret_cast<uint4*>(dst + head);
        for (uint64_t u = u0 + t; u < u1; u += kCopyThreads) d4[u] = s4[u];
    } else if (unit == 4) {
        const uint32_t* s4 = reinterpret_cast<const uint32_t*>(src + head);
        uint32_t* d4 = reinterpret_cast<uint32_t*>(dst + head);
        for (uint64_t u = u0 + t; u < u1; u += kCopyThreads) d4[u] = s4[u];
    } else {
        for (uint64_t u = u0 + t; u < u1; u += kCopyThreads) dst[head + u] = src[head + u];
    }
}

constexpr uint32_t kCopyMaxStreamsPerLaunch = 65536;

__global__ __launch_bounds__(kCopyThreads) void copy_uncompressed_kernel(const mdz_inflate_in_t* __restrict__ ins,
                                                                          uint32_t count) {
    const uint32_t slice = blockIdx.x % kCopySlices;
    for (uint32_t si = blockIdx.x / kCopySlices; si < count; si += gridDim.x / kCopySlices)
        copy_one(ins[si], slice);
}

// The queue's poll limit for a launch: kWaitPolls, or a lower value from
// MDZ_INFLATE_WAIT_POLLS (read at each launch; a test sets it to make the
// waves give up, and it can never raise the limit).
uint32_t wait_polls() {
    const char* v = std::getenv("MDZ_INFLATE_WAIT_POLLS");
    if (!v || !*v) return kWaitPolls;
    char* end = nullptr;
    const unsigned long x = std::strtoul(v, &end, 10);
    return end && *end == 0 && x < kWaitPolls ? (uint32_t)x : kWaitPolls;
}

}  // namespace

int inflate_slots(uint32_t* slots) {
    const size_t smem = sizeof(InflateLds);
    hipError_t e = set_dynamic_lds((const void*)inflate_kernel, smem);
    if (e != hipSuccess) return record_hip(e, "inflate_kernel LDS");
    int dev = 0, per_cu = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess ||
        (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)inflate_kernel, 128,
                                                          smem)) != hipSuccess)
        return record_hip(e, "inflate occupancy");
    *slots = (uint32_t)(per_cu * cus);
    return MDZ_OK;
}

// Launches the persistent kernel on a live queue (descriptors, statuses and
// the control words in host memory the host writes while it runs).
int inflate_live_launch(const mdz_inflate_in_t* in, mdz_inflate_status_t* out, uint32_t cap, int parse_hdr,
                        uint32_t blocks, uint32_t* cursor, uint32_t* live, hipStream_t s) {
    const size_t smem = sizeof(InflateLds);
    hipError_t e = set_dynamic_lds((const void*)inflate_kernel, smem);
    if (e != hipSuccess) return record_hip(e, "inflate_kernel LDS");
    hipLaunchKernelGGL(inflate_kernel, dim3(blocks), dim3(128), smem, s, in, out, nullptr, cap, parse_hdr,
                       cursor, nullptr, live, wait_polls());
    return record_hip(hipGetLastError(), "inflate_kernel live launch");
}

int inflate_batch(const mdz_inflate_in_t* in, mdz_inflate_status_t* out, uint32_t count,
                  int parse_hdr, const uint32_t* order, uint32_t persistent_blocks,
                  uint32_t* cursor, mdz_inflate_cost_t* cost, hipStream_t s) {
    if ((!in || !out) && count) return MDZ_ERR_INVALID;
    if (!inflate_hdr_ok(parse_hdr)) return MDZ_ERR_INVALID;
    if (count == 0) return MDZ_OK;
    const bool persistent = cursor != nullptr;
    const size_t smem = sizeof(InflateLds);
    hipError_t e = set_dynamic_lds((const void*)inflate_kernel, smem);
    if (e != hipSuccess) return record_hip(e, "inflate_kernel LDS");
    uint32_t grid = count;
    if (persistent) {
        if (!persistent_blocks) {
            int rc = inflate_slots(&persistent_blocks);
            if (rc) return rc;
        }
        grid = min(persistent_blocks, count);
        hipLaunchKernelGGL(reset_cursor_kernel, dim3(1), dim3(1), 0, s, cursor);
        if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "inflate cursor reset");
    }
    const int hdr = parse_hdr & 3;
    if (hdr == MDZ_HDR_GZIP) {
        hipLaunchKernelGGL(gzip_prepass_kernel, dim3((count + 255) / 256), dim3(256), 0, s, in, out, count);
        if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "gzip_prepass_kernel launch");
    }
    hipLaunchKernelGGL(inflate_kernel, dim3(grid), dim3(128), smem, s, in, out, order, count,
                       parse_hdr, persistent ? cursor : nullptr, cost, nullptr, wait_polls());
    if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "inflate_kernel launch");
    if ((parse_hdr & MDZ_HDR_VERIFY) && hdr) {
        hipLaunchKernelGGL(trailer_check_kernel, dim3(count), dim3(64), 0, s, in, out, hdr);
        if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "trailer_check_kernel launch");
    }
    return MDZ_OK;
}

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[kScratchDevices];   // scratch's address on each device
}  // namespace

// Looks up scratch's address on the current device, once per device (the
// device is current during the lookup, so the address is that device's copy).
int scratch_address(uint32_t** p) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return record_hip(e, "hipGetDevice");
    if (dev < 0 || dev >= kScratchDevices)
        return set_error(MDZ_ERR_UNSUPPORTED, "device %d: no-allocation entry points cover devices 0..%d",
                         dev, kScratchDevices - 1);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* q = nullptr;
        e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return record_hip(e, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = static_cast<uint32_t*>(q);
    }
    *p = g_scratch[dev];
    return MDZ_OK;
}

int copy_uncompressed(const mdz_inflate_in_t* in, uint32_t count, hipStream_t s) {
    if (!in && count) return MDZ_ERR_INVALID;
    if (count == 0) return MDZ_OK;
    const uint32_t grid = min(count, kCopyMaxStreamsPerLaunch) * kCopySlices;
    hipLaunchKernelGGL(copy_uncompressed_kernel, dim3(grid), dim3(kCopyThreads), 0, s, in, count);
    return record_hip(hipGetLastError(), "copy_uncompressed_kernel launch");
}

}  // namespace mdz

extern "C" int mdz_inflate_batch_sched_noalloc(const mdz_inflate_in_t* in, mdz_inflate_status_t* out,
                                               int count, int parse_hdr, const uint32_t* order,
                                               uint32_t blocks, mdz_inflate_cost_t* cost,
                                               void* stream) {
    if (count < 0 || !mdz::inflate_hdr_ok(parse_hdr)) return MDZ_ERR_INVALID;
    uint32_t* cursor = nullptr;
    const int rc = mdz::scratch_address(&cursor);
    if (rc) return rc;
    // the cursor is cleared by inflate_batch's reset kernel before the launch
    return mdz::inflate_batch(in, out, (uint32_t)count, parse_hdr, order, blocks, cursor, cost,
                              reinterpret_cast<hipStream_t>(stream));
}

// ------------------------------------------------------------ live queue --
// mdz_inflate_live_*: one persistent inflate launch that decodes streams the
// host publishes while it runs (descriptors, statuses and the control words
// {published, closed, expired, idle polls} in pinned host memory the GPU
// reads and writes coherently).  The kernel's workers claim items in order
// and wait for them; each ends when the queue is closed and drained, or
// after idle_ms without its item (8 s by default).
struct mdz_inflate_live {
    mdz_inflate_in_t* ins = nullptr;
    mdz_inflate_status_t* outs = nullptr;
    uint32_t* ctrl = nullptr;
    uint32_t* cursor = nullptr;
    uint32_t cap = 0, published = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
};

static void live_free(mdz_inflate_live* q) {
    if (q->done) (void)hipEventDestroy(q->done);
    if (q->stream) (void)hipStreamDestroy(q->stream);
    if (q->cursor) (void)hipFree(q->cursor);
    if (q->ins) (void)hipHostFree(q->ins);
    if (q->outs) (void)hipHostFree(q->outs);
    if (q->ctrl) (void)hipHostFree(q->ctrl);
    delete q;
}

extern "C" int mdz_inflate_live_open(uint32_t capacity, int parse_hdr, uint32_t blocks, uint32_t idle_ms,
                                     mdz_inflate_live_t** out) {
    if (!out || capacity == 0 || capacity > 0x7fffffffu)
        return mdz::set_error(MDZ_ERR_INVALID, "inflate live: capacity %u", capacity);
    *out = nullptr;
    if (parse_hdr != MDZ_HDR_RAW && parse_hdr != MDZ_HDR_ZLIB)
        return mdz::set_error(MDZ_ERR_UNSUPPORTED, "inflate live: raw or zlib framing only (parse_hdr %d)",
                              parse_hdr);
    mdz_inflate_live* q = new mdz_inflate_live();
    q->cap = capacity;
    hipError_t e;
    // coherent (fine-grained) pinned memory: each side sees the other's
    // writes while the kernel runs
    const unsigned fl = hipHostMallocCoherent;
    if ((e = hipHostMalloc((void**)&q->ins, sizeof(mdz_inflate_in_t) * capacity, fl)) != hipSuccess ||
        (e = hipHostMalloc((void**)&q->outs, sizeof(mdz_inflate_status_t) * capacity, fl)) != hipSuccess ||
        (e = hipHostMalloc((void**)&q->ctrl, 64, fl)) != hipSuccess ||
        (e = hipMalloc((void**)&q->cursor, 4)) != hipSuccess) {
        live_free(q);
        return mdz::set_error(MDZ_ERR_NOMEM, "inflate live: allocation failed (%s)", hipGetErrorString(e));
    }
    memset(q->ins, 0, sizeof(mdz_inflate_in_t) * capacity);
    for (uint32_t i = 0; i < capacity; ++i) {
        q->outs[i].bytes_written = 0;
        q->outs[i].status = (uint32_t)MDZ_ST_LIVE_PENDING;
        q->outs[i].reserved = 0;
    }
    memset(q->ctrl, 0, 64);
    const uint64_t polls = (uint64_t)(idle_ms ? idle_ms : 8000u) * mdz::kLivePollsPerMs;
    q->ctrl[3] = polls > 0xffffffffull ? 0xffffffffu : (uint32_t)polls;
    uint32_t grid = blocks;
    if (!grid) {                                   // a quarter of the co-resident slots
        int rc = mdz::inflate_slots(&grid);
        if (rc) { live_free(q); return rc; }
        grid = grid / 4 ? grid / 4 : 1;
    }
    if ((e = hipMemset(q->cursor, 0, 4)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&q->done, hipEventDisableTiming)) != hipSuccess) {
        live_free(q);
        return mdz::record_hip(e, "inflate live setup");
    }
    int rc = mdz::inflate_live_launch(q->ins, q->outs, capacity, parse_hdr, grid, q->cursor, q->ctrl, q->stream);
    if (rc) { live_free(q); return rc; }
    q->launched = true;
    if ((e = hipEventRecord(q->done, q->stream)) != hipSuccess) {
        __atomic_store_n(&q->ctrl[1], 1u, __ATOMIC_RELEASE);   // let the workers end before freeing
        (void)hipStreamSynchronize(q->stream);
        live_free(q);
        return mdz::record_hip(e, "inflate live event");
    }
    *out = q;
    return MDZ_OK;
}

extern "C" int mdz_inflate_live_submit(mdz_inflate_live_t* q, const mdz_inflate_in_t* in, uint32_t n,
                                       uint32_t* first) {
    if (!q || (!in && n)) return mdz::set_error(MDZ_ERR_INVALID, "inflate live submit: null argument");
    if (__atomic_load_n(&q->ctrl[1], __ATOMIC_ACQUIRE))
        return mdz::set_error(MDZ_ERR_INVALID, "inflate live submit: queue closed");
    if (n > q->cap - q->published)
        return mdz::set_error(MDZ_ERR_INVALID, "inflate live submit: %u items past capacity %u", n, q->cap);
    if (n) memcpy(q->ins + q->published, in, sizeof(mdz_inflate_in_t) * n);
    if (first) *first = q->published;
    q->published += n;
    __atomic_store_n(&q->ctrl[0], q->published, __ATOMIC_RELEASE);   // descriptors before the count
    return MDZ_OK;
}

extern "C" const mdz_inflate_status_t* mdz_inflate_live_statuses(const mdz_inflate_live_t* q) {
    return q ? q->outs : nullptr;
}

extern "C" int mdz_inflate_live_expired(mdz_inflate_live_t* q) {
    if (!q) return mdz::set_error(MDZ_ERR_INVALID, "inflate live: null queue");
    if (__atomic_load_n(&q->ctrl[2], __ATOMIC_ACQUIRE)) return 1;
    return hipEventQuery(q->done) == hipSuccess ? 1 : 0;          // ended before the close
}

extern "C" int mdz_inflate_live_close(mdz_inflate_live_t* q, mdz_inflate_status_t* statuses) {
    if (!q) return mdz::set_error(MDZ_ERR_INVALID, "inflate live close: null queue");
    __atomic_store_n(&q->ctrl[1], 1u, __ATOMIC_RELEASE);
    const hipError_t e = hipEventSynchronize(q->done);
    if (statuses && q->published) memcpy(statuses, q->outs, sizeof(mdz_inflate_status_t) * q->published);
    live_free(q);
    if (e != hipSuccess) return mdz::record_hip(e, "inflate live close");
    return MDZ_OK;
}
