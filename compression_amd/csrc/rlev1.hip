// RLE v1 decoder for gfx950.
//
// Replaces the reference's reader-warp / decoder-warp inflate kernel
// (include/rle_v1/decoder.h:426-470, decoder_warp :275-423).  Token grammar
// (decoder.h:296-411): int8 h < 0 -> -h LEB128 literals; h >= 0 -> int8 delta,
// LEB128 base, h+3 values base + i*delta (64-bit arithmetic, truncated to T).
//
// Pseudo-varints.  Cut a lane stream after every LEB128 terminator (a byte
// with bit 7 clear); call the pieces pseudo-varints.  Every token starts at
// a cut: a literal header has bit 7 set, so it joins the first literal's
// pseudo-varint and a group of n literals spans exactly n of them; a run
// header (bit 7 clear) is one by itself, and the delta byte is one too (bit 7
// clear) or joins the base value's (bit 7 set), so a run spans 3 or 2.
//
// One workgroup (16 wave64s) per chunk; a chunk fills most of a CU's LDS, so
// the workgroup is as wide as the hardware allows:
//   1. setup and staging: the chunk's interleaved words are scattered into
//      per-lane LDS regions (the same code as RLE v2);
//   2. cut table: 32 threads per lane count the lane's terminators, then list
//      where each pseudo-varint starts: cut[t][0] = 0, cut[t][i+1] = one past
//      the i-th terminator (16-bit entries);
//   3. walk: 32 threads of wave 0, one per lane, follow the token chain in
//      pseudo-varint steps (one table read and one data read per token),
//      check each header as the oracle does, list every token (first
//      pseudo-varint, first value, header) in LDS after the cut table, and
//      publish per lane how many values the listed tokens cover;
//   4. decode, overlapped with the walk: the other waves (and wave 0 once it
//      is done) claim units of 32 values of a lane in value order, wait until
//      the lane's walk has passed the unit, find the unit's first token by
//      binary search, load its tokens (run bases decoded here) and decode all
//      32 values without branches: a literal straight from its pseudo-varint
//      (start and length from the cut table), a run value as base + i*delta.
//      Every load of a unit is independent of the others, so they overlap.
//      The unit's bytes are stored as P-byte pieces, eight at a time.
// If a lane has more tokens than its list holds, the chunk is walked again
// with a checkpoint per value group and decoded in P-byte pieces in output
// order, each stepping from its group's checkpoint to its token.
// A lane's status is its first error in stream order.  Header errors (walk)
// and value errors (decode) carry the key (value index, step) and are
// combined with an atomic minimum; the chunk reports its lowest failing lane,
// as the oracle does.  Values after a lane's first error are unspecified.
// Decode waits end when the lane's walk passes the unit or ends, which every
// walk does; the wait is bounded anyway, and an expired wait reports
// MDZ_ST_INTERNAL instead of hanging.
// A chunk whose lanes do not fit the LDS budget, with a lane stream over
// 65535 bytes, or whose cut table does not fit, is decoded by a single-wave
// decoder that streams the payload through an LDS window.
#include <type_traits>

#include "lanes_dev.h"
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

enum : uint32_t { kHdr = 0, kDelta = 1, kBase = 2, kRun = 3, kLit = 4, kDone = 5 };
constexpr uint32_t kNT = 1024;           // threads per workgroup
constexpr uint32_t kGroups = 64;         // value groups (walk checkpoints) per lane
constexpr uint32_t kCutThreads = kNT / 32;   // threads per lane building the cut table
constexpr uint32_t kNoErr = 0xffffffffu;
constexpr uint32_t kUnit = 32;           // values per decode unit (token path)
constexpr uint32_t kSpinMax = 1u << 22;  // decode-unit wait bound (about 0.2 s)
constexpr uint32_t kMT = 4;              // tokens a unit loads at once

template <int EB> struct VarintSpec;
template <> struct VarintSpec<1> { static constexpr uint32_t kMax = 2; };
template <> struct VarintSpec<2> { static constexpr uint32_t kMax = 3; };
template <> struct VarintSpec<4> { static constexpr uint32_t kMax = 5; };
template <> struct VarintSpec<8> { static constexpr uint32_t kMax = 10; };

// Gather the 7-bit groups of the 8 bytes of x into one integer.
MDZ_DEV uint64_t gather7x8(uint64_t x) {
    return (x & 0x7full) | ((x >> 1) & (0x7full << 7)) | ((x >> 2) & (0x7full << 14)) |
           ((x >> 3) & (0x7full << 21)) | ((x >> 4) & (0x7full << 28)) |
           ((x >> 5) & (0x7full << 35)) | ((x >> 6) & (0x7full << 42)) |
           ((x >> 7) & (0x7full << 49));
}

// LEB128 decode from up to 16 peeked bytes (lo = bytes 0..7, hi = 8..15) of
// which `avail` are stream bytes.  Returns the length, or 0 with *err set:
// input overrun when the stream ends inside the value, malformed symbol when
// the value is longer than kMax bytes.  Mirrors the oracle's leb128().
template <int EB>
MDZ_DEV uint32_t varint_decode(uint64_t lo, uint64_t hi, uint32_t avail, uint64_t& val,
                               int32_t& err) {
    constexpr uint32_t kMax = VarintSpec<EB>::kMax;
    uint32_t len;
    if constexpr (kMax <= 8) {
        constexpr uint64_t tmask = kMax == 8 ? 0x8080808080808080ull
                                            : ((1ull << (8 * kMax)) - 1) & 0x8080808080808080ull;
        const uint64_t term = ~lo & tmask;
        len = term ? (__builtin_ctzll(term) >> 3) + 1 : 0;
    } else {
        const uint64_t tl = ~lo & 0x8080808080808080ull, th = ~hi & 0x8080ull;
        len = tl ? (__builtin_ctzll(tl) >> 3) + 1 : th ? 9 + (__builtin_ctzll(th) >> 3) : 0;
    }
    if (len == 0 || len > avail) {
        err = (len == 0 && avail >= kMax) ? MDZ_ST_BAD_SYMBOL : MDZ_ST_INPUT_OVERRUN;
        return 0;
    }
    uint64_t v = gather7x8(lo);
    if (len < 8) {
        v &= (1ull << (7 * len)) - 1;
    } else if constexpr (kMax > 8) {
        if (len > 8) {
            uint64_t h = (hi & 0x7f) | ((hi >> 1) & (0x7full << 7));
            h &= (1ull << (7 * (len - 8))) - 1;
            v |= h << 56;
        }
    }
    val = v;
    return len;
}

// ------------------------------------------------ windowed single wave ----
// Fallback for chunks whose payload exceeds the LDS budget: wave 0 stages
// the raw interleaved payload through an LDS ring; lanes 0..31 decode their
// streams from it, reading their words through the lane-address table; when
// every running lane waits on data beyond the window, the window slides to
// the lowest word still needed.  Returns the chunk status.
template <int EB, int SW>
MDZ_DEV int32_t rlev1_windowed(const u32x4* __restrict__ src, LaneTable& T, uint32_t cl,
                               uint32_t* ring, uint32_t ring_words, uint8_t* chunk_out,
                               uint64_t limit, uint32_t nvals, uint32_t line) {
    constexpr int NB = EB == 8 ? 16 : 8;
    constexpr uint32_t kMaxV = VarintSpec<EB>::kMax;
    constexpr uint32_t kRefillAt = NB - 4;
    const uint32_t l = lane_id(), t = l & 31;
    const bool decoder = l < 32;
    const uint32_t W = (cl + 3) / 4;
    const uint32_t chunk_words = (T.total + 3) & ~3u;

    LaneCursor cur;
    cur.init(T, t, decoder ? W : 0);
    ByteBuf<NB> bb;
    LineEmitter<EB, SW> em;
    em.init(chunk_out, limit, t, line, 0);
    uint32_t mode = decoder ? kHdr : kDone;
    uint32_t rem = 0, vo = 0;
    uint64_t rv = 0;
    int64_t rd = 0;
    int32_t err = MDZ_ST_OK;
    const uint32_t last_bytes = cl - 4 * (W ? W - 1 : 0);

    uint32_t hi = min(chunk_words, ring_words);
    ring_stage(ring, ring_words, src, 0, hi);
    wave_sync();
    const uint32_t rmask = ring_words - 1;
    for (;;) {
        while (mode != kDone) {
            while (bb.nb <= kRefillAt && cur.r < cur.W) {
                uint32_t a = cur.addr();
                if (a >= hi) break;
                bb.append(ring[a & rmask], cur.r + 1 == cur.W ? last_bytes : 4);
                cur.advance(T, t);
            }
            const uint32_t need = mode == kRun ? 0 : (mode == kLit || mode == kBase) ? kMaxV : 1;
            if (bb.nb < need && cur.r < cur.W) break;   // waits on the window
            if (mode == kRun) {
                em.push(rv);
                rv += (uint64_t)rd;
                ++vo;
                if (--rem == 0) mode = kHdr;
            } else if (mode == kLit || mode == kBase) {
                uint64_t v;
                uint32_t n = varint_decode<EB>(bb.lo(), bb.hi(), bb.nb, v, err);
                if (n == 0) { mode = kDone; break; }
                bb.consume(n);
                if (mode == kLit) {
                    em.push(v);
                    ++vo;
                    if (--rem == 0) mode = kHdr;
                } else {
                    rv = v;
                    mode = kRun;
                }
            } else if (mode == kHdr) {
                if (bb.nb == 0) { err = MDZ_ST_INPUT_OVERRUN; mode = kDone; break; }
                uint32_t h = (uint32_t)(bb.lo() & 0xff);
                bb.consume(1);
                if (h & 0x80) { rem = 256 - h; mode = kLit; }
                else { rem = h + 3; mode = kDelta; }
                if (rem > nvals - vo) { err = MDZ_ST_OUTPUT_OVERRUN; mode = kDone; break; }
            } else {   // kDelta
                if (bb.nb == 0) { err = MDZ_ST_INPUT_OVERRUN; mode = kDone; break; }
                rd = (int64_t)(int8_t)(bb.lo() & 0xff);
                bb.consume(1);
                mode = kBase;
            }
            if (vo == nvals) mode = kDone;
        }
        uint32_t needw = (mode != kDone && cur.r < cur.W) ? cur.addr() : 0xffffffffu;
        uint32_t lo = wave_min_u32(needw);
        if (lo == 0xffffffffu) break;
        lo &= ~3u;
        uint32_t nhi = min(chunk_words, lo + ring_words);
        wave_sync();
        if (nhi > hi) ring_stage(ring, ring_words, src, max(hi, lo), nhi);
        hi = max(hi, nhi);
        wave_sync();
    }
    uint64_t bad = __ballot(err != MDZ_ST_OK);
    if (!bad) return MDZ_ST_OK;
    return __builtin_amdgcn_readlane(err, (int)__builtin_ctzll(bad));
}


// ---------------------------------------------------------------- tables --
struct V1Tables {
    uint32_t lbase[32];      // lane region byte offsets in the data area (16-aligned)
    uint32_t cl[32];         // lane stream bytes
    uint32_t lerr[32];       // first error of each lane: (4 * value + step) << 8 | status byte
    uint32_t sW[32];         // interleave segments (lane_setup)
    uint32_t S[32];
    uint32_t M[32];
    uint32_t nterm[32];      // terminators of each lane stream
    uint32_t cut0[32];       // first cut-table entry of each lane
    uint32_t vstop[32];      // values the walk covered (later values are not decoded)
    uint32_t ntok[32];       // tokens listed per lane (token path)
    uint32_t vpub[32];       // values the lane's listed tokens cover so far (token path)
    uint32_t done[32];       // the lane's walk has ended: vstop and ntok are final
    uint32_t total;          // words of all lanes
    uint32_t fits;           // lane regions fit the data budget
    uint32_t extent_ok;      // lanes lie inside the chunk's payload extent
    uint32_t used;           // bytes of the lane regions
    uint32_t fast;           // the chunk takes the table path
    uint32_t tok_off;        // byte offset of the token lists in LDS
    uint32_t cap;            // tokens each lane's list holds
    uint32_t overflow;       // a lane has more tokens than its list holds
    uint32_t claimed;        // decode units claimed, all lanes (token path)
    uint32_t stuck;          // a decode wait expired (never expected)
    uint32_t pad2[6];
    union {
        struct {                                 // staging
            uint32_t dest[32][32];
            uint8_t segof[kSegBlocks];
        } st;
        uint32_t ck[32][kGroups];                // checkpoint path: token k | v << 16
        LaneTable lt;                            // windowed decoder
    };
};
static_assert(sizeof(V1Tables) % 16 == 0, "data area alignment");

MDZ_DEV uint32_t err_key(uint32_t v, uint32_t step, int32_t code) {
    return ((4 * v + step) << 8) | ((uint32_t)code & 0xffu);
}

// Two bytes at byte p of a 4-byte-aligned LDS region (byte p in bits 0..7).
MDZ_DEV uint32_t lds_u16_at(const uint8_t* base, uint32_t p) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(base) + (p >> 2);
    return alignbyte(w[1], w[0], p & 3) & 0xffffu;
}

// Value of the LEB128 of `len` bytes (1 <= len <= its maximum) at byte a of
// a 4-byte-aligned LDS region.  Bits past 64 are dropped, as the oracle does.
template <int EB>
MDZ_DEV uint64_t leb_at(const uint8_t* base, uint32_t a, uint32_t len) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(base) + (a >> 2);
    const uint32_t k = a & 3;
    if constexpr (EB <= 4) {
        const uint32_t w0 = w[0], w1 = w[1];
        const uint32_t lo = alignbyte(w1, w0, k);
        uint32_t v = (lo & 0x7fu) | ((lo >> 1) & 0x3f80u) | ((lo >> 2) & 0x1fc000u) |
                     ((lo >> 3) & 0xfe00000u);
        if (len < 4) v &= (1u << (7 * len)) - 1u;
        else if (len > 4) v |= ((w1 >> (8 * k)) & 0x7fu) << 28;
        return v;
    } else {
        const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
        const uint64_t lo = ((uint64_t)alignbyte(w2, w1, k) << 32) | alignbyte(w1, w0, k);
        uint64_t v = gather7x8(lo);
        if (len < 8) {
            v &= (1ull << (7 * len)) - 1;
        } else if (len > 8) {
            const uint32_t hi = alignbyte(w3, w2, k);
            uint64_t h = (hi & 0x7fu) | ((hi >> 1) & (0x7fu << 7));
            h &= (1ull << (7 * (len - 8))) - 1;
            v |= h << 56;
        }
        return v;
    }
}

// Pseudo-varint j of a lane read as a LEB128 value, skipping its first byte
// when `skip` (a literal or delta byte joined to it).  j < nterm: complete,
// its length is cut[j+1] - start; j == nterm: the lane's unterminated tail.
// Returns the status the oracle's LEB128 reader gives there.
template <int EB>
MDZ_DEV int32_t pseudo_value(const uint8_t* base, const uint16_t* cut, uint32_t j, uint32_t nterm,
                             uint32_t cl, uint32_t skip, uint64_t& val) {
    constexpr uint32_t kMax = VarintSpec<EB>::kMax;
    const uint32_t a = cut[j] + skip;
    if (j >= nterm) {
        val = 0;
        return cl - a >= kMax ? MDZ_ST_BAD_SYMBOL : MDZ_ST_INPUT_OVERRUN;
    }
    const uint32_t len = cut[j + 1] - a;
    if (len > kMax) { val = 0; return MDZ_ST_BAD_SYMBOL; }
    val = leb_at<EB>(base, a, len);
    return MDZ_ST_OK;
}

// One token of a lane: pseudo-varint k starts it; its first two bytes are
// h (header) and, for a run, the delta byte.
struct Token {
    uint32_t k, v;           // first pseudo-varint, first value
    uint32_t hd;             // header | delta << 8
    MDZ_DEV bool lit() const { return hd & 0x80u; }
    MDZ_DEV uint32_t count() const { return lit() ? 256u - (hd & 0xffu) : (hd & 0xffu) + 3u; }
    MDZ_DEV uint32_t next_k() const { return k + (lit() ? count() : (hd & 0x8000u) ? 2u : 3u); }
};

// Writes one output piece of P bytes (words w, byte i of the piece in
// w[i / 4] bits 8 * (i % 4)) at chunk offset o, clipped to `limit`.
template <int P>
MDZ_DEV void store_piece(uint8_t* out, uint32_t o, uint64_t limit, const uint32_t (&w)[4]) {
    uint8_t* p = out + o;
    if (o + P <= limit) {
        if constexpr (P == 16) *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
        else if constexpr (P == 8) *reinterpret_cast<uint2*>(p) = make_uint2(w[0], w[1]);
        else if constexpr (P == 4) *reinterpret_cast<uint32_t*>(p) = w[0];
        else if constexpr (P == 2) *reinterpret_cast<uint16_t*>(p) = (uint16_t)w[0];
        else *p = (uint8_t)w[0];
    } else {
#pragma unroll
        for (uint32_t i = 0; i < (uint32_t)P; ++i)      // unrolled: w stays in registers
            if (o + i < limit) p[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
    }
}

// Decodes the piece of lane t holding values [v0, v0 + P / EB).  The walk
// covered values below vs; `ck` is the token covering value (v0 >> gsh) << gsh.
// Value errors go to lerr[t] as an atomic minimum.
template <int EB, int P>
MDZ_DEV void decode_piece(const uint8_t* base, const uint16_t* cut, uint32_t nterm, uint32_t cl,
                          uint32_t ck, uint32_t v0, uint32_t vs, uint32_t* lerr,
                          uint32_t (&w)[4]) {
    constexpr int NV = P / EB;
    constexpr uint64_t vmask = EB == 8 ? ~0ull : ((1ull << (8 * EB)) - 1);
    Token tk;
    tk.k = ck & 0xffffu;
    tk.v = ck >> 16;
    tk.hd = lds_u16_at(base, cut[tk.k]);
    while (tk.v + tk.count() <= v0) {             // tokens between the checkpoint and v0
        const uint32_t n = tk.count();
        tk.k = tk.next_k();
        tk.v += n;
        tk.hd = lds_u16_at(base, cut[tk.k]);
    }
    uint64_t rbase = 0;
    bool have_base = false;
    int32_t e = MDZ_ST_OK;
    uint32_t ekey = kNoErr;
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = 0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const uint32_t vi = v0 + i;
        if (vi >= vs) break;
        if (vi >= tk.v + tk.count()) {            // the next token starts inside the piece
            const uint32_t n = tk.count();
            tk.k = tk.next_k();
            tk.v += n;
            tk.hd = lds_u16_at(base, cut[tk.k]);
            have_base = false;
        }
        uint64_t val;
        if (tk.lit()) {
            const uint32_t j = tk.k + (vi - tk.v);
            if (j > nterm) break;                 // past the lane's unterminated tail (an
                                                  // earlier value reports the error)
            e = pseudo_value<EB>(base, cut, j, nterm, cl, j == tk.k, val);
            if (e != MDZ_ST_OK && ekey == kNoErr) ekey = err_key(vi, 2, e);
        } else {
            if (!have_base) {
                const bool joined = tk.hd & 0x8000u;    // delta byte joined to the base's
                e = pseudo_value<EB>(base, cut, tk.k + (joined ? 1 : 2), nterm, cl, joined, rbase);
                if (e != MDZ_ST_OK && ekey == kNoErr) ekey = err_key(tk.v, 2, e);
                rbase &= vmask;
                have_base = true;
            }
            const int64_t d = (int64_t)(int8_t)(tk.hd >> 8);
            val = rbase + (uint64_t)(vi - tk.v) * (uint64_t)d;
        }
        val &= vmask;
        if constexpr (EB == 8) {
            w[2 * i] = (uint32_t)val;
            w[2 * i + 1] = (uint32_t)(val >> 32);
        } else if constexpr (EB == 4) {
            w[i] = (uint32_t)val;
        } else {
            w[(i * EB) >> 2] |= (uint32_t)val << (8 * ((i * EB) & 3));
        }
    }
    if (ekey != kNoErr) atomicMin(lerr, ekey);
}

// Run base per listed token: the value type, 64-bit for 8-byte elements.
template <int EB> using BaseT = typename std::conditional<EB == 8, uint64_t, uint32_t>::type;

// Status of literal pseudo-varint j read from byte a0 with `len` bytes (the
// oracle's LEB128 reader): j == nterm is the unterminated tail of the lane.
template <int EB>
MDZ_DEV int32_t lit_status(uint32_t j, uint32_t nterm, uint32_t cl, uint32_t a0, uint32_t len) {
    constexpr uint32_t kMax = VarintSpec<EB>::kMax;
    return j >= nterm ? (cl - a0 >= kMax ? MDZ_ST_BAD_SYMBOL : MDZ_ST_INPUT_OVERRUN)
                      : (len > kMax ? MDZ_ST_BAD_SYMBOL : MDZ_ST_OK);
}

// Places value i of a piece into its words (w starts zeroed).
template <int EB>
MDZ_DEV void put_value(uint32_t (&w)[4], uint32_t i, BaseT<EB> x) {
    if constexpr (EB == 8) {
        w[2 * i] = (uint32_t)x;
        w[2 * i + 1] = (uint32_t)(x >> 32);
    } else if constexpr (EB == 4) {
        w[i] = (uint32_t)x;
    } else {
        w[(i * EB) >> 2] |= (uint32_t)x << (8 * ((i * EB) & 3));
    }
}

// One listed token reduced to what a value needs: its first value tv, a
// literal flag, and A, D: a literal value i is pseudo-varint i - A, a run
// value is A + i * D.  Passed by value so the per-value selection stays in
// registers.
template <int EB> struct TokVal {
    BaseT<EB> A, D;
    uint32_t tv;
    uint32_t lit;
    uint32_t berr;           // a run's base: its LEB128 status (0 for a literal group)
};

// Token idx of a lane's list (idx >= nt: none; its first value reads as vs).
// A run's base is decoded here from its pseudo-varint.
template <int EB>
MDZ_DEV TokVal<EB> tok_val(const uint8_t* base, const uint16_t* cut, uint32_t nterm, uint32_t cl,
                           const uint32_t* kvs, const uint16_t* hds, uint32_t nt, uint32_t idx,
                           uint32_t vs) {
    typedef BaseT<EB> VT;
    const uint32_t ic = min(idx, nt - 1);
    const uint32_t kv = kvs[ic], h = hds[ic], k = kv & 0xffffu;
    TokVal<EB> r;
    r.tv = idx < nt ? kv >> 16 : vs;
    r.lit = h & 0x80u;
    const VT d = (VT)(int64_t)(int8_t)(h >> 8);
    // run: the delta byte is its own pseudo-varint (bit 7 clear) or joins the base's
    const uint32_t joined = (h >> 15) & 1u;
    uint64_t b = 0;
    const int32_t be = pseudo_value<EB>(base, cut, k + 2 - joined, nterm, cl, joined, b);
    r.berr = r.lit ? (uint32_t)MDZ_ST_OK : (uint32_t)be;
    r.A = r.lit ? (VT)(r.tv - k) : (VT)((VT)b - (VT)r.tv * d);
    r.D = d;
    return r;
}

template <int EB>
MDZ_DEV TokVal<EB> pick(bool c, TokVal<EB> x, TokVal<EB> y) {
    TokVal<EB> r;
    r.A = c ? x.A : y.A;
    r.D = c ? x.D : y.D;
    r.tv = c ? x.tv : y.tv;
    r.lit = c ? x.lit : y.lit;
    r.berr = c ? x.berr : y.berr;
    return r;
}

// Value vi of a lane from its token m (selected by the caller): a literal
// read from its pseudo-varint, or a run value.  `ok` is false for a literal
// past the lane's unterminated tail (an earlier value reports that error).
// `ek` is the error key the value reports: a literal's LEB128 status at vi,
// or a run base's status at the run's first value; kNoErr if none.
template <int EB>
MDZ_DEV BaseT<EB> unit_value(TokVal<EB> m, uint32_t vi, const uint8_t* base, const uint16_t* cut,
                             uint32_t nterm, uint32_t cl, bool& ok, uint32_t& ek) {
    typedef BaseT<EB> VT;
    ok = true;
    if (!m.lit) {
        ek = m.berr != (uint32_t)MDZ_ST_OK ? err_key(m.tv, 2, (int32_t)m.berr) : kNoErr;
        return m.A + (VT)vi * m.D;
    }
    // pseudo-varint j; reads clamped to the lane's table
    const uint32_t j = vi - (uint32_t)m.A;
    const uint32_t a0 = cut[min(j, nterm)] + (m.tv == vi);
    const uint32_t len = cut[min(j + 1, nterm)] - a0;
    ok = j <= nterm;
    const int32_t le = lit_status<EB>(j, nterm, cl, a0, len);
    ek = ok && le != MDZ_ST_OK ? err_key(vi, 2, le) : kNoErr;
    return (VT)leb_at<EB>(base, a0, len);
}

// Decodes values [V, V + n) of lane t (n <= kUnit, V a multiple of kUnit)
// from its token list and stores them as P-byte pieces.
// i0 = the token covering V; tokens are taken kMT = 4 at a time (T0..T3,
// and e = the first value after them).  A piece inside the current batch
// decodes its values unrolled (their loads overlap); a piece reaching past
// it (rare) goes value by value, loading batches as needed.
template <int EB, int P>
MDZ_DEV void decode_unit(const uint8_t* base, const uint16_t* cut, const uint32_t* kvs,
                         const uint16_t* hds, uint32_t nt, uint32_t i0, uint32_t nterm,
                         uint32_t cl, uint32_t V, uint32_t n, uint32_t vs, uint32_t* lerr,
                         uint8_t* out, uint64_t limit, uint32_t t, uint32_t line) {
    static_assert(kMT == 4, "four tokens per batch");
    typedef BaseT<EB> VT;
    constexpr uint32_t NV = P / EB;                 // values per piece
    constexpr VT vmask = (VT)(~0ull >> (64 - 8 * EB));
    const uint32_t vend = min(V + n, vs);
    uint32_t ekey = kNoErr, ib = i0;
    TokVal<EB> T0, T1, T2, T3;
    uint32_t e;
    auto load = [&]() {
        T0 = tok_val<EB>(base, cut, nterm, cl, kvs, hds, nt, ib, vs);
        T1 = tok_val<EB>(base, cut, nterm, cl, kvs, hds, nt, ib + 1, vs);
        T2 = tok_val<EB>(base, cut, nterm, cl, kvs, hds, nt, ib + 2, vs);
        T3 = tok_val<EB>(base, cut, nterm, cl, kvs, hds, nt, ib + 3, vs);
        e = ib + 4 < nt ? kvs[ib + 4] >> 16 : vs;
    };
    auto value = [&](uint32_t vi, bool& ok, uint32_t& ek) {
        const TokVal<EB> m = pick<EB>(vi >= T3.tv, T3, pick<EB>(vi >= T2.tv, T2,
                                      pick<EB>(vi >= T1.tv, T1, T0)));
        return unit_value<EB>(m, vi, base, cut, nterm, cl, ok, ek) & vmask;
    };
    load();
    // Groups of up to 8 pieces (one 128-byte line of 16-byte pieces) are kept
    // in registers and stored together, so each line is written within a few
    // cycles (lines written piece by piece over a whole unit are evicted from
    // L2 half-written).
    constexpr uint32_t NP = kUnit * EB / P;         // pieces per unit
    constexpr uint32_t NG = NP < 8 ? NP : 8;        // pieces per group
    const uint32_t lb = V * EB;
    uint32_t r = lb / line, off = lb - r * line;
    for (uint32_t g0 = 0; g0 < NP && g0 * NV < n; g0 += NG) {
        uint32_t ub[NG][4];
#pragma unroll
        for (uint32_t gq = 0; gq < NG; ++gq) {
            const uint32_t q = g0 + gq, v0 = V + q * NV;
            uint32_t (&w)[4] = ub[gq];
            w[0] = w[1] = w[2] = w[3] = 0;
            if (q * NV >= n) continue;
            if (min(v0 + NV, vend) <= e) {
#pragma unroll
                for (uint32_t i = 0; i < NV; ++i) {
                    const uint32_t vi = v0 + i;
                    bool ok;
                    uint32_t ek;
                    const VT val = value(vi, ok, ek);
                    const bool act = vi < vend;
                    ekey = act ? min(ekey, ek) : ekey;
                    put_value<EB>(w, i, act && ok ? val : (VT)0);
                }
            } else {
                // value by value; the piece's bytes shift in from the top of w
                for (uint32_t i = 0; i < NV; ++i) {
                    const uint32_t vi = v0 + i;
                    while (vi >= e && vi < vend) {
                        ib += kMT;
                        load();
                    }
                    bool ok;
                    uint32_t ek;
                    const VT val = value(vi, ok, ek);
                    const bool act = vi < vend;
                    ekey = act ? min(ekey, ek) : ekey;
                    const VT x = act && ok ? val : (VT)0;
                    if constexpr (EB == 8) {
                        w[0] = w[2]; w[1] = w[3];
                        w[2] = (uint32_t)x; w[3] = (uint32_t)(x >> 32);
                    } else if constexpr (EB == 4) {
                        w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = (uint32_t)x;
                    } else {
                        w[0] = alignbyte(w[1], w[0], EB); w[1] = alignbyte(w[2], w[1], EB);
                        w[2] = alignbyte(w[3], w[2], EB);
                        w[3] = (w[3] >> (8 * EB)) | ((uint32_t)x << (32 - 8 * EB));
                    }
                }
                // the piece now sits in the top P bytes of w: move it to the bottom
                if constexpr (P == 8) { w[0] = w[2]; w[1] = w[3]; }
                else if constexpr (P == 4) { w[0] = w[3]; }
                else if constexpr (P == 2) { w[0] = w[3] >> 16; }
                else if constexpr (P == 1) { w[0] = w[3] >> 24; }
            }
        }
#pragma unroll
        for (uint32_t gq = 0; gq < NG; ++gq) {
            if ((g0 + gq) * NV >= n) break;
            store_piece<P>(out, (r * 32 + t) * line + off, limit, ub[gq]);
            off += P;
            if (off == line) { off = 0; ++r; }
        }
    }
    if (ekey != kNoErr) atomicMin(lerr, ekey);
}

// ------------------------------------------------------------ the kernel --
// `lds_bytes` is the dynamic LDS size; `data_cap` bounds the lane regions
// for the table path; `ring_words` (a power of two) sizes the window of the
// streaming decoder; `tok_limit` bounds the per-lane token lists.
template <int EB, int P>
__global__ __launch_bounds__(kNT) void rlev1_decode_kernel(
    const uint8_t* __restrict__ payload, const uint64_t* __restrict__ blk_off,
    const uint64_t* __restrict__ col_len, uint64_t n_chunks, uint64_t chunk_bytes,
    uint32_t line, uint8_t* __restrict__ dst, uint64_t dst_bytes, int32_t* __restrict__ status,
    uint32_t data_cap, uint32_t ring_words, uint32_t lds_bytes, uint32_t tok_limit) {
    extern __shared__ __align__(16) uint32_t smem[];
    V1Tables& tb = *reinterpret_cast<V1Tables*>(smem);
    uint8_t* const data = reinterpret_cast<uint8_t*>(smem) + sizeof(V1Tables);

    const uint64_t c = blockIdx.x;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    MDZ_PHASE_START();
    const uint32_t nvals = (uint32_t)(chunk_bytes / 32 / EB);
    const uint64_t out_base = c * chunk_bytes;
    const uint64_t limit = dst_bytes > out_base ? min(dst_bytes - out_base, chunk_bytes) : 0;
    const uint64_t c_off = blk_off[c], c_end = blk_off[c + 1];
    const u32x4* src = reinterpret_cast<const u32x4*>(payload + c_off);

    // ---- 1. lane geometry (wave 0)
    if (wave == 0) {
        const uint32_t used = lane_setup<32>(tb, col_len, c, 0, data_cap, c_off, c_end);
        const uint32_t cl = (uint32_t)col_len[c * 32 + (l & 31)];
        const bool big = __ballot(cl > 0xffffu) != 0 || nvals > 0xffffu;
        if (l == 0) {
            tb.used = used;
            tb.fast = !big;
        }
    }
    __syncthreads();
    if (!tb.extent_ok) {
        if (tid == 0 && status) status[c] = MDZ_ST_INPUT_OVERRUN;
        return;
    }
    const uint32_t used = tb.used;
    uint16_t* const cut = reinterpret_cast<uint16_t*>(data + used + 16);
    const uint32_t cut_cap = (lds_bytes - (uint32_t)sizeof(V1Tables) - used - 16) / 2;
    bool fast = tb.fits && tb.fast;
    if (fast) {
        stage_tables(tb, tid, kNT);
        __syncthreads();
        MDZ_PHASE(0);
        stage_lanes(tb, reinterpret_cast<uint32_t*>(data), src, tid, kNT);
        __syncthreads();
        MDZ_PHASE(1);

        // ---- 2. cut table: thread j of lane t takes a contiguous share of its blocks
        const uint32_t t = tid / kCutThreads, j = tid % kCutThreads;
        const uint32_t cl = tb.cl[t], nb = (cl + 15) / 16;
        const uint32_t b0 = j * nb / kCutThreads, b1 = (j + 1) * nb / kCutThreads;
        const u32x4* blk = reinterpret_cast<const u32x4*>(data + tb.lbase[t]);
        auto term_mask = [&](uint32_t b, uint32_t i, uint32_t x) {
            uint32_t m = ~x & 0x80808080u;
            const int rest = (int)cl - (int)(16 * b + 4 * i);   // stream bytes in this word
            if (rest < 4) m = rest <= 0 ? 0u : m & ((1u << (8 * rest)) - 1u);
            return m;
        };
        uint32_t n = 0;
        for (uint32_t b = b0; b < b1; ++b) {
            const u32x4 x = blk[b];
#pragma unroll
            for (int i = 0; i < 4; ++i) n += __popc(term_mask(b, i, x[i]));
        }
        uint32_t incl = n;
#pragma unroll
        for (int o = 1; o < (int)kCutThreads; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, kCutThreads);
            if (j >= (uint32_t)o) incl += y;
        }
        if (j == kCutThreads - 1) tb.nterm[t] = incl;
        __syncthreads();
        if (wave == 0) {
            const uint32_t nt = l < 32 ? tb.nterm[l] + 1 : 0;
            const uint32_t ex = wave_incl_scan_u32(nt) - nt;
            if (l < 32) tb.cut0[l] = ex;
            if (l < 32) {
                tb.vpub[l] = 0;
                tb.done[l] = 0;
            }
            if (l == 31) {
                tb.fast = ex + nt <= cut_cap;
                // token lists after the cut table: as many tokens per lane as fit
                const uint32_t off = (uint32_t)sizeof(V1Tables) + used + 16 + ((2 * (ex + nt) + 15) & ~15u);
                constexpr uint32_t rec = 4 + 2;
                tb.tok_off = off;
                tb.cap = min(off < lds_bytes ? (lds_bytes - off) / (32 * rec) : 0u, tok_limit);
                tb.overflow = 0;
                tb.claimed = 0;
                tb.stuck = 0;
            }
        }
        __syncthreads();
        fast = tb.fast;
        if (fast) {
            uint16_t* ct = cut + tb.cut0[t];
            if (j == 0) ct[0] = 0;
            uint32_t idx = incl - n + 1;
            for (uint32_t b = b0; b < b1; ++b) {
                const u32x4 x = blk[b];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t m = term_mask(b, i, x[i]);
                    const uint32_t pos = 16 * b + 4 * i + 1;
#pragma unroll
                    for (int j = 0; j < 4; ++j)            // predicated, no data-dependent loop
                        if (m & (0x80u << (8 * j))) ct[idx++] = (uint16_t)(pos + j);
                }
            }
        }
        __syncthreads();
        MDZ_PHASE(2);
    }
    if (!fast) {
        // ---- streaming decoder (wave 0)
        if (wave == 0) {
            const uint32_t cl = tb.cl[l & 31];
            lane_table_build(tb.lt, (cl + 3) / 4);
            int32_t st = rlev1_windowed<EB, P>(src, tb.lt, cl, reinterpret_cast<uint32_t*>(data),
                                              ring_words, dst + out_base, limit, nvals, line);
            if (l == 0 && status) status[c] = st;
        }
        return;
    }

    // ---- 3-4. Token path: wave 0 walks, one thread per lane stream, and
    //      lists every token after the cut table (tkv = first pseudo-varint |
    //      first value << 16, thd = header | delta << 8), as many per lane as
    //      fit, publishing per lane the values its listed tokens cover.  The
    //      other waves (and wave 0 once its walk ends) claim units of kUnit
    //      values in value order, wait until their lane's walk has passed the
    //      unit, and decode it.  If a lane has more tokens than its list
    //      holds, the walk flags the chunk, every wait ends, and the chunk is
    //      walked again with a checkpoint per value group (checkpoint path).
    const uint32_t upl = (nvals + kUnit - 1) / kUnit;      // decode units per lane
    const uint32_t cap = tb.cap;
    uint32_t* const tkv = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(smem) + tb.tok_off);
    uint16_t* const thd = reinterpret_cast<uint16_t*>(tkv + 32 * cap);
    uint8_t* const out = dst + out_base;
    bool tokens = cap >= kMT;
    if (tokens) {
        if (wave == 0) {
            if (l < 32) {
                const uint32_t t = l;
                const uint8_t* base = data + tb.lbase[t];
                const uint16_t* ct = cut + tb.cut0[t];
                uint32_t* kvl = tkv + t * cap;
                uint16_t* hdl = thd + t * cap;
                const uint32_t nterm = tb.nterm[t], cl = tb.cl[t];
                uint32_t k = 0, v = 0, i = 0, err = kNoErr;
                while (v < nvals && k <= nterm) {
                    const uint32_t s = ct[k];
                    if (s >= cl) { err = err_key(v, 0, MDZ_ST_INPUT_OVERRUN); break; }
                    Token tk;
                    tk.k = k;
                    tk.v = v;
                    tk.hd = lds_u16_at(base, s);
                    const uint32_t cnt = tk.count();
                    if (cnt > nvals - v) { err = err_key(v, 0, MDZ_ST_OUTPUT_OVERRUN); break; }
                    if (!tk.lit() && s + 1 >= cl) { err = err_key(v, 1, MDZ_ST_INPUT_OVERRUN); break; }
                    if (i == cap) { tb.overflow = 1; break; }
                    kvl[i] = k | v << 16;
                    hdl[i] = (uint16_t)tk.hd;
                    ++i;
                    k = tk.next_k();
                    v += cnt;
                    // publish: the token's LDS writes land before the count
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __hip_atomic_store(&tb.ntok[t], i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_store(&tb.vpub[t], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if (err != kNoErr) atomicMin(&tb.lerr[t], err);
                tb.ntok[t] = i;
                tb.vstop[t] = v;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __hip_atomic_store(&tb.done[t], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            wave_sync();
        }
        MDZ_PHASE(3);
        // Consumers: a wave claims 64 units at a time in value order (unit u
        // = lane u % 32, values kUnit * (u / 32) ..), each thread waits until
        // its lane's walk has covered its unit, then decodes it.
        const uint32_t units = 32 * upl;
        for (;;) {
            uint32_t u0 = 0;
            if (l == 0) u0 = atomicAdd(&tb.claimed, 64u);
            u0 = __builtin_amdgcn_readfirstlane(u0);
            if (u0 >= units) break;
            const uint32_t u = u0 + l;
            if (u >= units) continue;
            const uint32_t t = u & 31, V = (u >> 5) * kUnit, n = min(kUnit, nvals - V);
            bool fin = false;
            for (uint32_t it = 0;; ++it) {
                if (__hip_atomic_load(&tb.done[t], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                    fin = true;
                    break;
                }
                if (__hip_atomic_load(&tb.vpub[t], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= V + n)
                    break;
                if (it >= kSpinMax) { tb.stuck = 1; break; }
                __builtin_amdgcn_s_sleep(2);
            }
            if (tb.overflow || tb.stuck) continue;
            // values the lane's listed tokens cover; ntok read after vpub
            const uint32_t vs = fin ? tb.vstop[t] : __hip_atomic_load(&tb.vpub[t], __ATOMIC_ACQUIRE,
                                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            const uint32_t nt = __hip_atomic_load(&tb.ntok[t], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (V >= vs || nt == 0) continue;
            const uint32_t* kvs = tkv + t * cap;
            // the token covering V: the last with first value <= V
            uint32_t lo = 0, hi = nt - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if ((kvs[mid] >> 16) <= V) lo = mid;
                else hi = mid - 1;
            }
            decode_unit<EB, P>(data + tb.lbase[t], cut + tb.cut0[t], kvs, thd + t * cap, nt, lo,
                               tb.nterm[t], tb.cl[t], V, n, vs, &tb.lerr[t], out, limit, t, line);
        }
        __syncthreads();
        MDZ_PHASE(4);
        tokens = !tb.overflow;
        if (tb.overflow) {                        // the checkpoint path redoes the chunk
            if (tid < 32) tb.lerr[tid] = kNoErr;
            __syncthreads();
        }
    }
    if (!tokens) {
        uint32_t gsh = 0;                            // groups of G = 2^gsh values, at most kGroups
        while (((nvals + (1u << gsh) - 1) >> gsh) > kGroups) ++gsh;
        if (wave == 0 && l < 32) {
            const uint32_t t = l;
            const uint8_t* base = data + tb.lbase[t];
            const uint16_t* ct = cut + tb.cut0[t];
            const uint32_t nterm = tb.nterm[t], cl = tb.cl[t], G = 1u << gsh;
            uint32_t k = 0, v = 0, err = kNoErr;
            while (v < nvals && k <= nterm) {
                const uint32_t s = ct[k];
                if (s >= cl) { err = err_key(v, 0, MDZ_ST_INPUT_OVERRUN); break; }
                Token tk;
                tk.k = k;
                tk.v = v;
                tk.hd = lds_u16_at(base, s);
                const uint32_t cnt = tk.count();
                if (cnt > nvals - v) { err = err_key(v, 0, MDZ_ST_OUTPUT_OVERRUN); break; }
                if (!tk.lit() && s + 1 >= cl) { err = err_key(v, 1, MDZ_ST_INPUT_OVERRUN); break; }
                for (uint32_t g = (v + G - 1) >> gsh; (g << gsh) < v + cnt; ++g) tb.ck[t][g] = k | v << 16;
                k = tk.next_k();
                v += cnt;
            }
            tb.vstop[t] = v;
            tb.lerr[t] = err;
        }
        __syncthreads();
        MDZ_PHASE(4);

        // P-byte pieces in output order, each from its group's checkpoint
        {
            const float inv_line = __builtin_amdgcn_rcpf((float)line);
            const uint32_t npieces = (uint32_t)(chunk_bytes / P);
            for (uint32_t q = tid; q < npieces; q += kNT) {
                const uint32_t o = q * P;      // pieces past `limit` are decoded for the status only
                uint32_t R = (uint32_t)((float)o * inv_line);     // line of the piece
                if (R * line > o) --R;
                else if ((R + 1) * line <= o) ++R;
                const uint32_t t = R & 31, v0 = ((R >> 5) * line + (o - R * line)) / EB;
                const uint32_t vs = tb.vstop[t];
                if (v0 >= vs) continue;
                uint32_t w[4];
                decode_piece<EB, P>(data + tb.lbase[t], cut + tb.cut0[t], tb.nterm[t], tb.cl[t],
                                    tb.ck[t][v0 >> gsh], v0, vs, &tb.lerr[t], w);
                store_piece<P>(out, o, limit, w);
            }
        }
    }
    __syncthreads();
    MDZ_PHASE(5);
    if (tid == 0 && status) {
        int32_t st = MDZ_ST_OK;
        for (int t = 0; t < 32; ++t) {
            const uint32_t e = tb.lerr[t];
            if (e != kNoErr) { st = (int32_t)(int8_t)(e & 0xffu); break; }
        }
        status[c] = tb.stuck ? MDZ_ST_INTERNAL : st;
    }
}

template <int EB, int P>
hipError_t launch(const mdz_lanes_t& in, void* dst, uint64_t dst_bytes, int32_t* status,
                  uint32_t lds, uint32_t data_cap, uint32_t ring_words, uint32_t tok_limit,
                  hipStream_t s) {
    auto k = rlev1_decode_kernel<EB, P>;
    hipError_t e = set_dynamic_lds((const void*)k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((uint32_t)in.n_chunks), dim3(kNT), lds, s,
                       (const uint8_t*)in.payload, in.blk_off, in.col_len, in.n_chunks,
                       in.chunk_bytes, in.lane_line, (uint8_t*)dst, dst_bytes, status, data_cap,
                       ring_words, lds, tok_limit);
    return hipGetLastError();
}

template <int EB>
hipError_t launch_eb(const mdz_lanes_t& in, void* dst, uint64_t dst_bytes, int32_t* status,
                     uint32_t lds, uint32_t data_cap, uint32_t ring_words, uint32_t tok_limit,
                     uint32_t p, hipStream_t s) {
#define MDZ_V1_LAUNCH(P) launch<EB, P>(in, dst, dst_bytes, status, lds, data_cap, ring_words, tok_limit, s)
    if (p == 16) return MDZ_V1_LAUNCH(16);
    if (p == 8) return MDZ_V1_LAUNCH(8);
    if constexpr (EB <= 4)
        if (p == 4) return MDZ_V1_LAUNCH(4);
    if constexpr (EB <= 2)
        if (p == 2) return MDZ_V1_LAUNCH(2);
    if constexpr (EB == 1)
        if (p == 1) return MDZ_V1_LAUNCH(1);
#undef MDZ_V1_LAUNCH
    return hipErrorInvalidValue;
}

}  // namespace

int rlev1_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                     const mdz_lanes_tune_t* tune, hipStream_t s) {
    if (!in || !in->payload || !in->blk_off || !in->col_len || (!dst && dst_bytes))
        return MDZ_ERR_INVALID;
    const uint32_t eb = in->elem_bytes, line = in->lane_line;
    const uint64_t chunk = in->chunk_bytes;
    if (!(eb == 1 || eb == 2 || eb == 4 || eb == 8) || line == 0 || line % eb || chunk % 32 ||
        (chunk / 32) % line || chunk / 32 / eb == 0 || chunk / 32 / eb > 0xffffffu)
        return MDZ_ERR_INVALID;
    if (in->n_chunks == 0) return MDZ_OK;
    if (in->n_chunks > 0x7fffffffull) return MDZ_ERR_UNSUPPORTED;
    if (((uintptr_t)in->payload & 15) != 0) return MDZ_ERR_INVALID;
    // LDS: the tables, the lane regions and the cut table (at most one entry
    // per payload byte, plus one per lane).  With a bound on one chunk's
    // payload the allocation is sized for it, else the whole budget is taken.
    constexpr uint32_t kLdsMax = 160 * 1024;
    const uint32_t fixed = (uint32_t)sizeof(V1Tables) + 16;
    uint32_t lds = kLdsMax;
    if (tune && tune->max_chunk_bytes) {
        const uint64_t regions = ((uint64_t)tune->max_chunk_bytes + 32 * 12 + 15) & ~15ull;
        const uint64_t want = fixed + regions + 2 * ((uint64_t)tune->max_chunk_bytes + 32) + 16;
        if (want < lds) lds = (uint32_t)((want + 255) & ~255ull);
    }
    uint32_t ring_words = 0;
    if (tune && tune->ring_words) {          // a small window, for testing the streaming path
        if (tune->ring_words < 256 || (tune->ring_words & (tune->ring_words - 1)) ||
            tune->ring_words * 4ull > kLdsMax - fixed)
            return MDZ_ERR_INVALID;
        ring_words = tune->ring_words;
        lds = max(lds, fixed + ring_words * 4);
    }
    lds = max(lds, fixed + 4096);
    uint32_t data_cap = lds - fixed - 64;
    if (ring_words) {
        data_cap = min(data_cap, ring_words * 4);
    } else {                                 // the largest power of two in the data area
        ring_words = 256;
        while (ring_words * 8ull <= lds - fixed) ring_words *= 2;
    }
    // widest piece that divides the line and the output alignment
    uint32_t p = 16;
    while (p > eb && ((line % p) || ((uintptr_t)dst % p))) p >>= 1;
    const uint32_t tok_limit = tune && tune->max_lane_tokens ? tune->max_lane_tokens : 0xffffffffu;
    hipError_t e;
    switch (eb) {
    case 1: e = launch_eb<1>(*in, dst, dst_bytes, status, lds, data_cap, ring_words, tok_limit, p, s); break;
    case 2: e = launch_eb<2>(*in, dst, dst_bytes, status, lds, data_cap, ring_words, tok_limit, p, s); break;
    case 4: e = launch_eb<4>(*in, dst, dst_bytes, status, lds, data_cap, ring_words, tok_limit, p, s); break;
    default: e = launch_eb<8>(*in, dst, dst_bytes, status, lds, data_cap, ring_words, tok_limit, p, s); break;
    }
    return record_hip(e, "rlev1_decode_kernel launch");
}

}  // namespace mdz
