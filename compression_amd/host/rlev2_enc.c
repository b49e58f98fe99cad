/* RLE v2 encoders (host side): the lane-transposed format, and the plain
 * per-chunk format at the end of the file.  The lane encoder restates the
 * "theirs" side of the reference's unresolved merge:
 *   include/rlev2/utils.h:6-69       width maps, constants (INPUT_T = u64)
 *   include/rlev2/encoder.h:83-785   determineEncoding and the four writers
 *   include/rlev2/encoder_trans.h:14-289  per-lane block_encode loop
 *   include/rlev2/encoder_trans.h:291-349 4-byte round-robin transpose
 * Values are unsigned (no zigzag for values); deltas are raw u64
 * differences. */
#include "mdz_host.h"

#include <stdlib.h>
#include <string.h>

#define MIN_REPEAT 3
#define MAX_LIT 128
#define MAX_SHORT_REPEAT 10

/* utils.h:36-59 */
static const uint8_t DEC_W[32] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16,
                                  17, 18, 19, 20, 21, 22, 23, 24, 26, 28, 30, 32, 40, 48, 56, 64};

static uint8_t enc_w(uint32_t bits) {       /* __BIT_WIDTH_ENCODE_MAP */
    if (bits <= 1) return 0;
    if (bits <= 24) return (uint8_t)(bits - 1);
    if (bits <= 26) return 24;
    if (bits <= 28) return 25;
    if (bits <= 30) return 26;
    if (bits <= 32) return 27;
    if (bits <= 40) return 28;
    if (bits <= 48) return 29;
    if (bits <= 56) return 30;
    return 31;
}
static uint8_t closest_bit(uint32_t bits) {  /* __CLOSEST_FIXED_BIT_MAP */
    if (bits <= 1) return 1;
    if (bits <= 24) return (uint8_t)bits;
    if (bits <= 26) return 26;
    if (bits <= 28) return 28;
    if (bits <= 30) return 30;
    if (bits <= 32) return 32;
    if (bits <= 40) return 40;
    if (bits <= 48) return 48;
    if (bits <= 56) return 56;
    return 64;
}
static uint8_t closest_aligned_bit(uint32_t bits) {  /* __CLOSEST_ALIGNED_FIXED_BIT_MAP */
    if (bits <= 1) return 1;
    if (bits <= 2) return 2;
    if (bits <= 4) return 4;
    if (bits <= 8) return 8;
    if (bits <= 16) return 16;
    if (bits <= 24) return 24;
    if (bits <= 32) return 32;
    if (bits <= 40) return 40;
    if (bits <= 48) return 48;
    if (bits <= 56) return 56;
    return 64;
}
/* find_closes_num_bits (encoder.h:103-114): closest fixed width of the
 * value's bit length (values are unsigned, so never "negative"). */
static uint8_t closest_num_bits(uint64_t v) {
    uint32_t c = 0;
    while (v) { v >>= 1; ++c; }
    return closest_bit(c);
}

typedef struct {
    uint8_t* out; uint64_t pos, cap;
    uint64_t lit[MAX_LIT], delta[MAX_LIT];
    uint32_t n, fix, var;
    int fixed_delta;
    uint8_t delta_bits;
    int head;       /* plain format: the HEAD side's int64 determineEncoding */
} enc_t;

static void put(enc_t* e, uint8_t b) {
    if (e->pos < e->cap) e->out[e->pos] = b;
    e->pos++;
}

/* write_aligned_ints (encoder.h:124-165) */
static void write_aligned(enc_t* e, const uint64_t* in, uint32_t len, uint32_t bits) {
    if (bits < 8) {
        const uint8_t mask = (uint8_t)((1u << bits) - 1);
        const uint32_t hops = 8 / bits, rem = len % hops, full = len - rem;
        for (uint32_t i = 0; i < full; i += hops) {
            uint8_t b = 0;
            for (uint32_t j = 0; j < hops; ++j) b |= (uint8_t)((in[i + j] & mask) << (8 - (j + 1) * bits));
            put(e, b);
        }
        if (rem) {
            int sh = 8 - (int)bits;
            uint8_t b = 0;
            for (uint32_t i = full; i < len; ++i) { b |= (uint8_t)((in[i] & mask) << sh); sh -= (int)bits; }
            put(e, b);
        }
    } else {
        const int bytes = (int)bits / 8;
        for (uint32_t i = 0; i < len; ++i)
            for (int j = bytes - 1; j >= 0; --j) put(e, (uint8_t)(in[i] >> (j * 8)));
    }
}

/* write_unaligned_ints (encoder.h:167-216): MSB-first bit packing.  A
 * 64-bit field keeps all its bits (the device shift the reference encoder
 * ran with saturates, so its mask is all ones). */
static void write_unaligned(enc_t* e, const uint64_t* in, uint32_t len, uint32_t bits) {
    uint32_t left = 8;
    uint8_t cur = 0;
    for (uint32_t i = 0; i < len; ++i) {
        uint64_t v = in[i];
        uint32_t todo = bits;
        while (todo > left) {
            cur |= (uint8_t)(v >> (todo - left));
            todo -= left;
            v &= todo >= 64 ? ~0ull : ((1ull << todo) - 1);
            put(e, cur);
            cur = 0;
            left = 8;
        }
        left -= todo;
        cur |= (uint8_t)(v << left);
        if (left == 0) { put(e, cur); cur = 0; left = 8; }
    }
    if (left != 8) put(e, cur);
}

static void compute_hist(const uint64_t* d, uint32_t len, uint8_t hist[32]) {
    memset(hist, 0, 32);
    for (uint32_t i = 0; i < len; ++i) hist[enc_w(closest_num_bits(d[i]))]++;
}

/* compute_percentile_bit_width (encoder.h:225-236) */
static uint8_t percentile_bits(const uint8_t hist[32], uint32_t len, uint32_t p) {
    p = 100 - p;
    if (p > 100) p = 100;
    int32_t plen = p == 0 ? 0 : (int32_t)((len * p - 1) / 100 + 1);
    for (int i = 31; i >= 0; --i) {
        plen -= hist[i];
        if (plen < 0) return DEC_W[i];
    }
    return 0;
}

static void write_varint_u(enc_t* e, uint64_t v) {   /* writeVulong (encoder.h:603-615) */
    for (;;) {
        if ((v & ~0x7full) == 0) { put(e, (uint8_t)v); return; }
        put(e, (uint8_t)(0x80 | (v & 0x7f)));
        v >>= 7;
    }
}

/* writeShortRepeatValues (encoder.h:537-569) */
static void write_short_repeat(enc_t* e) {
    const uint64_t v = e->lit[0];
    const uint32_t vbits = closest_num_bits(v), vbytes = (vbits - 1) / 8 + 1;
    put(e, (uint8_t)(((vbytes - 1) << 3) | (e->fix - MIN_REPEAT)));
    for (int i = (int)vbytes - 1; i >= 0; --i) put(e, (uint8_t)(v >> (i * 8)));
    e->fix = 0; e->n = 0; e->var = 0;
}

/* writeDirectValues (encoder.h:571-598) */
static void write_direct(enc_t* e) {
    uint8_t hist[32];
    compute_hist(e->lit, e->n, hist);
    const uint32_t bits = closest_aligned_bit(percentile_bits(hist, e->n, 100));
    const uint32_t len = e->var - 1;
    put(e, (uint8_t)(0x40 | (enc_w(bits) << 1) | ((len & 0x100) >> 8)));
    put(e, (uint8_t)(len & 0xff));
    write_aligned(e, e->lit, e->n, bits);
    e->var = 0; e->n = 0;
}

/* writeDeltaValues (encoder.h:622-697) */
static void write_delta(enc_t* e) {
    uint32_t len, code = 0, bits = closest_aligned_bit(e->delta_bits);
    if (e->fixed_delta) {
        if (e->fix > MIN_REPEAT) { len = e->fix - 1; e->fix = 0; }
        else { len = e->var - 1; e->var = 0; }
    } else {
        if (bits == 1) bits = 2;
        code = enc_w(bits);
        len = e->var - 1;
        e->var = 0;
    }
    put(e, (uint8_t)(0xC0 | (code << 1) | ((len & 0x100) >> 8)));
    put(e, (uint8_t)(len & 0xff));
    write_varint_u(e, e->lit[0]);
    const int64_t d0 = (int64_t)e->delta[0];
    write_varint_u(e, ((uint64_t)d0 << 1) ^ (uint64_t)(d0 >> 63));       /* writeVslong */
    if (!e->fixed_delta) write_aligned(e, e->delta + 1, e->n - 2, bits);
    for (uint32_t i = 0; i < e->n; ++i) e->delta[i] = 0;
    e->n = 0;
    e->var = 0;
}

/* preparePatchedBlob1 + writePatchedBasedValues (encoder.h:244-364,699-785) */
static void write_patched(enc_t* e, uint64_t lit_min, uint64_t* red, uint8_t b95, uint8_t b100) {
    uint64_t mask = (1ull << b95) - 1;
    uint32_t pw = closest_bit((uint32_t)(b100 - b95));
    if (pw == 64) { pw = 56; b95 = 8; mask = (1ull << b95) - 1; }
    uint64_t gaps[MAX_LIT], patches[MAX_LIT], list[MAX_LIT + 8];
    uint32_t np = 0, prev = 0, max_gap = 0;
    for (uint32_t i = 0; i < e->n; ++i) {
        if (red[i] > mask) {
            const uint32_t gap = i - prev;
            if (gap > max_gap) max_gap = gap;
            prev = i;
            gaps[np] = gap;
            patches[np++] = red[i] >> b95;
            red[i] &= mask;
        }
    }
    uint32_t plen = np;
    uint32_t pgw = (max_gap == 0 && plen != 0) ? 1 : closest_num_bits(max_gap);
    if (pgw > 8) { pgw = 8; plen += max_gap == 511 ? 2 : 1; }
    uint32_t nl = 0;
    for (uint32_t i = 0, gi = 0; i < plen && gi < np; ++i, ++gi) {
        uint64_t g = gaps[gi];
        while (g > 255) { list[nl++] = 255ull << pw; ++i; g -= 255; }
        list[nl++] = (g << pw) | patches[gi];
    }
    const uint32_t varlen = e->var - 1;
    const uint32_t vbits = closest_num_bits(lit_min) + 1, vbytes = (vbits - 1) / 8 + 1;
    put(e, (uint8_t)(0x80 | (enc_w(b95) << 1) | ((varlen & 0x100) >> 8)));
    put(e, (uint8_t)(varlen & 0xff));
    put(e, (uint8_t)(((vbytes - 1) << 5) | enc_w(pw)));
    put(e, (uint8_t)(((pgw - 1) << 5) | plen));
    for (int i = (int)vbytes - 1; i >= 0; --i) put(e, (uint8_t)(i >= 8 ? 0 : lit_min >> (i * 8)));
    write_unaligned(e, red, e->n, closest_bit(b95));
    write_unaligned(e, list, plen, closest_bit(pgw + pw));
    e->var = 0;
    e->n = 0;
}

/* determineEncoding (encoder.h:366-535, theirs) */
static void determine(enc_t* e) {
    if (e->n <= MIN_REPEAT) { write_direct(e); return; }
    int inc = 1, dec = 1;
    e->fixed_delta = 1;
    uint64_t mn = e->lit[0], mx = e->lit[0];
    const uint64_t d_init = e->lit[1] - e->lit[0];
    uint64_t max_delta = 0;
    for (uint32_t i = 1; i < e->n; ++i) {
        const uint64_t l1 = e->lit[i], l0 = e->lit[i - 1], cd = l1 - l0;
        if (l1 < mn) mn = l1;
        if (l1 > mx) mx = l1;
        inc &= l0 <= l1;
        dec &= l0 >= l1;
        e->fixed_delta &= cd == d_init;
        /* HEAD (encoder.h:414): deltas[i-1] = abs(currDelta) over int64; the
         * plain encoder takes values in [0, 2^63), where int64 order is the
         * u64 order and every difference is exact */
        const uint64_t ad = e->head && (int64_t)cd < 0 ? 0 - cd : cd;
        e->delta[i - 1] = ad;
        if (ad > max_delta) max_delta = ad;
    }
    e->delta[0] = e->lit[1] - e->lit[0];
    /* is_safe_subtract(max, min) with INT8_MIN: false only when min == 0
     * and max <= 128 (encoder.h:93-97, B5).  The HEAD side subtracts
     * INT64_MIN (encoder.h:85-89): with min == 0 that wraps to INT64_MIN in
     * the device arithmetic the encoder ran with, so values in [0, 2^63)
     * are always safe. */
    if (!e->head && !(mn > 0 || mx > mn + 128)) { write_direct(e); return; }
    if (mn == mx || e->fixed_delta) { write_delta(e); return; }
    if (d_init != 0) {
        e->delta_bits = closest_num_bits(max_delta);
        if (inc || dec) { write_delta(e); return; }
    }
    uint8_t hist[32];
    compute_hist(e->lit, e->n, hist);
    const uint32_t b100 = percentile_bits(hist, e->n, 100), b90 = percentile_bits(hist, e->n, 90);
    if ((uint32_t)(b100 - b90) > 1) {
        uint64_t red[MAX_LIT];
        for (uint32_t i = 0; i < e->n; ++i) red[i] = e->lit[i] - mn;
        compute_hist(red, e->n, hist);
        const uint8_t r95 = percentile_bits(hist, e->n, 95), r100 = percentile_bits(hist, e->n, 100);
        if (r100 != r95) { write_patched(e, mn, red, r95, r100); return; }
    }
    write_direct(e);
}

/* block_encode main loop (encoder_trans.h:82-196, theirs) over one lane's
 * values. */
static uint64_t run_encode(const uint64_t* vals, uint64_t nv, uint8_t* out, uint64_t cap, int head) {
    enc_t* e = (enc_t*)calloc(1, sizeof(enc_t));
    e->out = out; e->cap = cap; e->head = head;
    uint64_t prev_delta = 0;
    for (uint64_t k = 0; k < nv; ++k) {
        const uint64_t val = vals[k];
        if (e->n == 0) { e->lit[e->n++] = val; e->fix = 1; e->var = 1; continue; }
        if (e->n == 1) {
            e->lit[e->n++] = val;
            prev_delta = val - e->lit[0];
            if (val == e->lit[0]) { e->fix = 2; e->var = 0; } else { e->fix = 0; e->var = 2; }
            continue;
        }
        const uint64_t cur_delta = val - e->lit[e->n - 1];
        if (prev_delta == 0 && cur_delta == 0) {
            e->lit[e->n++] = val;
            if (e->var > 0) e->fix = 2;
            e->fix++;
            if (e->fix >= MIN_REPEAT && e->var > 0) {
                e->n -= MIN_REPEAT;
                e->var -= MIN_REPEAT - 1;
                determine(e);
                prev_delta = 0;
                for (uint32_t i = 0; i < MIN_REPEAT; ++i) { e->lit[i] = val; e->delta[i] = 0; }
                e->n = MIN_REPEAT;
            } else if (e->fix >= MIN_REPEAT) {
                e->delta[0] = 0;
            }
            if (e->fix == MAX_LIT) {
                /* B19: the reference enters determineEncoding here with
                 * var_runlen == 0; for 128 zeros the INT8_MIN safety test
                 * sends the run to DIRECT with length var_runlen-1 = -1 (a
                 * header claiming 512 values over 128 bits of data, which
                 * its decoder's uint8 length reads as 0).  The run's
                 * literal count is the intended length. */
                if (e->var == 0) e->var = e->n;
                determine(e);
            }
            continue;
        }
        if (e->fix >= MIN_REPEAT) {
            if (e->fix <= MAX_SHORT_REPEAT) {
                write_short_repeat(e);
            } else {
                e->fixed_delta = 1;
                write_delta(e);
            }
        }
        if (e->fix > 0 && e->fix < MIN_REPEAT && val != e->lit[e->n - 1]) {
            e->var = e->fix;
            e->fix = 0;
        }
        if (e->n == 0) {
            e->lit[e->n++] = val; e->fix = 1; e->var = 1;
        } else {
            prev_delta = val - e->lit[e->n - 1];
            e->lit[e->n++] = val;
            e->var++;
            if (e->var == MAX_LIT) determine(e);
        }
    }
    if (e->n != 0) {
        if (e->var != 0) {
            determine(e);
        } else if (e->fix != 0) {
            if (e->fix < MIN_REPEAT) {
                e->var = e->fix; e->fix = 0;
                determine(e);
            } else if (e->fix <= MAX_SHORT_REPEAT) {
                write_short_repeat(e);
            } else {
                e->fixed_delta = 1;
                write_delta(e);
            }
        }
    }
    uint64_t len = e->pos;
    free(e);
    return len;
}

uint64_t mdz_rlev2_lane_encode(const uint64_t* vals, uint64_t nv, uint8_t* out, uint64_t cap) {
    return run_encode(vals, nv, out, cap, 0);
}

/* ------------------------------------------------------ plain format -- */
/* block_encode (encoder.h:808-958, HEAD side: int64 values, CHUNK_SIZE
 * bytes per chunk) over one chunk's values: the same state machine as the
 * lane encoder, with the HEAD determineEncoding (absolute deltas, INT64_MIN
 * safety test).  Two HEAD slips are not kept: a long fixed run written after
 * a variable run reuses that run's stale deltas[0] (theirs resets it, as the
 * lane encoder does), and writePatchedBasedValues prints a line. */
uint64_t mdz_rlev2_plain_chunk_encode(const int64_t* vals, uint64_t nv, uint8_t* out, uint64_t cap) {
    return run_encode((const uint64_t*)vals, nv, out, cap, 1);
}

typedef struct {
    const int64_t* in; uint64_t in_bytes, chunk;
    uint8_t** streams; uint64_t* len; int bad;
} plain2_ctx;

static void plain2_task(void* vctx, uint64_t c) {
    plain2_ctx* x = (plain2_ctx*)vctx;
    const uint64_t start = c * x->chunk / 8;
    const uint64_t stop = ((c + 1) * x->chunk < x->in_bytes ? (c + 1) * x->chunk : x->in_bytes) / 8;
    const uint64_t nv = stop > start ? stop - start : 0;
    /* worst case per 128 values: a 2-byte DIRECT header and 64-bit values */
    const uint64_t cap = nv * 8 + (nv / 128 + 2) * 4 + 64;
    uint8_t* buf = (uint8_t*)malloc(cap);
    const uint64_t n = mdz_rlev2_plain_chunk_encode(x->in + start, nv, buf, cap);
    if (n > cap) x->bad = 1;
    x->streams[c] = buf;
    x->len[c] = n;
}

/* compress_gpu (encoder.h:1003-1058): [u32 n+1][u64 ptr[n+1]][u64 in_bytes]
 * [chunk streams back to back], ptr the prefix sums with ptr[0] = 0.  The
 * last in_bytes % 8 bytes are not encoded (each chunk ends at its last whole
 * int64). */
int mdz_rlev2_plain_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** out,
                           uint64_t* out_bytes, int threads) {
    if (chunk == 0 || chunk % 8 || in_bytes == 0 || !out || !out_bytes) return -1;
    const uint64_t n = (in_bytes - 1) / chunk + 1;
    if (n + 1 > 0xffffffffull) return -1;
    for (uint64_t i = 0; i < in_bytes / 8; ++i)
        if (((const int64_t*)in)[i] < 0) return -2;       /* values in [0, 2^63) only */
    plain2_ctx x = {(const int64_t*)in, in_bytes, chunk, (uint8_t**)calloc(n, sizeof(uint8_t*)),
                    (uint64_t*)calloc(n, sizeof(uint64_t)), 0};
    mdz_parallel_for(n, threads, plain2_task, &x);
    uint64_t tot = 0;
    for (uint64_t c = 0; c < n; ++c) tot += x.len[c];
    const uint64_t hdr = 4 + 8 * (n + 1) + 8;
    uint8_t* o = x.bad ? NULL : (uint8_t*)malloc(hdr + tot);
    if (o) {
        const uint32_t np = (uint32_t)(n + 1);
        memcpy(o, &np, 4);
        uint64_t acc = 0;
        memcpy(o + 4, &acc, 8);
        for (uint64_t c = 0; c < n; ++c) {
            memcpy(o + hdr + acc, x.streams[c], x.len[c]);
            acc += x.len[c];
            memcpy(o + 4 + 8 * (c + 1), &acc, 8);
        }
        memcpy(o + 4 + 8 * (n + 1), &in_bytes, 8);
    }
    for (uint64_t c = 0; c < n; ++c) free(x.streams[c]);
    free(x.streams);
    free(x.len);
    if (!o) return -3;
    *out = o;
    *out_bytes = hdr + tot;
    return 0;
}

/* ---------------------------------------------------------- container -- */
typedef struct {
    const uint8_t* in; uint64_t in_bytes, chunk; int eb; uint32_t ru;
    uint8_t** bufs; uint64_t* col_len; uint64_t* chunk_len;
} enc_ctx;

/* Lane t's values: chunk elements k*32*RU + t*RU + j (encoder_trans.h:42-43,94-108),
 * zero-extended from eb bytes; the input is zero-padded to whole chunks. */
static void gather(const enc_ctx* x, uint64_t c, int t, uint64_t* v, uint64_t nv) {
    const uint64_t base = c * x->chunk;
    for (uint64_t q = 0; q < nv; ++q) {
        const uint64_t el = (q / x->ru) * 32 * x->ru + (uint64_t)t * x->ru + q % x->ru;
        const uint64_t off = base + el * (uint64_t)x->eb;
        uint64_t val = 0;
        if (off + (uint64_t)x->eb <= x->in_bytes) memcpy(&val, x->in + off, (size_t)x->eb);
        v[q] = val;
    }
}

static void enc_task(void* vctx, uint64_t c) {
    enc_ctx* x = (enc_ctx*)vctx;
    const uint64_t nv = x->chunk / (uint64_t)x->eb / 32;
    uint64_t* v = (uint64_t*)malloc(nv * 8);
    const uint64_t cap = nv * 10 + 64;
    uint64_t total = 0;
    for (int t = 0; t < 32; ++t) {
        gather(x, c, t, v, nv);
        uint8_t* s = (uint8_t*)malloc(cap);
        const uint64_t len = mdz_rlev2_lane_encode(v, nv, s, cap);
        x->bufs[c * 32 + t] = s;
        x->col_len[c * 32 + t] = len;
        total += (len + 3) / 4 * 4;
    }
    x->chunk_len[c] = (total + 127) / 128 * 128;
    free(v);
}

static void interleave(uint8_t* dst, uint8_t* const* lanes, const uint64_t* len) {
    uint64_t words[32], maxw = 0, o = 0;
    for (int t = 0; t < 32; ++t) {
        words[t] = (len[t] + 3) / 4;
        if (words[t] > maxw) maxw = words[t];
    }
    for (uint64_t r = 0; r < maxw; ++r)
        for (int t = 0; t < 32; ++t)
            if (r < words[t])
                for (int b = 0; b < 4; ++b) {
                    const uint64_t s = r * 4 + (uint64_t)b;
                    dst[o++] = s < len[t] ? lanes[t][s] : 0;
                }
}

/* compress_gpu_transpose (encoder_trans.h:428-537).  blk_off has n+1
 * entries: offsets 0..n-1 of each chunk, and blk_off[n] = in_bytes (the
 * reference stores the uncompressed size there, :490-492).  *payload_bytes
 * is the true payload length. */
int mdz_rlev2_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, int eb,
                         uint32_t read_unit, uint8_t** payload, uint64_t* payload_bytes,
                         uint64_t** col_len, uint64_t** blk_off, uint64_t* n_chunks, int threads) {
    if (!(eb == 1 || eb == 2 || eb == 4 || eb == 8) || read_unit == 0 || chunk % ((uint64_t)eb * 32) ||
        (chunk / (uint64_t)eb / 32) % read_unit || in_bytes == 0)
        return -1;
    const uint64_t n = (in_bytes + chunk - 1) / chunk;
    enc_ctx x = {in, in_bytes, chunk, eb, read_unit, (uint8_t**)calloc(n * 32, sizeof(uint8_t*)),
                 (uint64_t*)malloc(n * 32 * 8), (uint64_t*)malloc(n * 8)};
    mdz_parallel_for(n, threads, enc_task, &x);
    uint64_t* off = (uint64_t*)malloc((n + 1) * 8);
    off[0] = 0;
    for (uint64_t c = 0; c < n; ++c) off[c + 1] = off[c] + x.chunk_len[c];
    uint8_t* pay = (uint8_t*)calloc(off[n] ? off[n] : 1, 1);
    for (uint64_t c = 0; c < n; ++c) {
        interleave(pay + off[c], x.bufs + c * 32, x.col_len + c * 32);
        for (int t = 0; t < 32; ++t) free(x.bufs[c * 32 + t]);
    }
    *payload_bytes = off[n];
    off[n] = in_bytes;
    free(x.bufs);
    free(x.chunk_len);
    *payload = pay; *col_len = x.col_len; *blk_off = off; *n_chunks = n;
    return 0;
}

