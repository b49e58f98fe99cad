/* mdz oracle: RLE v2 (ORC integer RLE v2, lane-transposed).  TEST
 * INFRASTRUCTURE ONLY.
 *
 * Restates the "theirs" side of the reference's unresolved merge:
 *   include/rlev2/utils.h:6-69       width maps, constants (INPUT_T = u64)
 *   include/rlev2/encoder.h:83-785   determineEncoding and the four writers
 *   include/rlev2/encoder_trans.h:14-289  per-lane block_encode loop
 *   include/rlev2/encoder_trans.h:291-349 4-byte round-robin transpose
 *   include/rlev2/decoder_new.h:354-611   decoder_warp
 * Values are unsigned (no zigzag for values); deltas are raw u64
 * differences; PATCHED_BASE patches are applied with a 64-bit mask (the
 * intended value; decoder_new.h:533 uses a 32-bit mask, B6). */
#include "mdz_oracle.h"

#include <stdlib.h>
#include <string.h>

#define MIN_REPEAT 3

/* __CLOSEST_FIXED_BIT_MAP (utils.h:36-59): the patch-entry width */
static uint8_t closest_bit(uint32_t bits) {
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

/* utils.h:36-59 */
static const uint8_t DEC_W[32] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16,
                                  17, 18, 19, 20, 21, 22, 23, 24, 26, 28, 30, 32, 40, 48, 56, 64};

/* ----------------------------------------------------------- decoder -- */
typedef struct { const uint8_t* in; uint64_t len, p; } rd_t;

static int rd_byte(rd_t* r, uint8_t* b) {
    if (r->p >= r->len) return MDZ_ST_INPUT_OVERRUN;
    *b = r->in[r->p++];
    return MDZ_ST_OK;
}
/* MSB-first bit reader over whole bytes of one run (decoder_new.h:457-474). */
typedef struct { uint32_t left; uint8_t cur; } bits_t;
static int rd_bits(rd_t* r, bits_t* bs, uint32_t n, uint64_t* out) {
    uint64_t res = 0;
    while (n > bs->left) {
        res = bs->left ? (res << bs->left) | (bs->cur & ((1u << bs->left) - 1)) : res;
        n -= bs->left;
        int st = rd_byte(r, &bs->cur);
        if (st) return st;
        bs->left = 8;
    }
    if (n) {
        res = (n >= 64 ? 0 : res << n);
        bs->left -= n;
        res |= (uint64_t)((bs->cur >> bs->left) & ((1u << n) - 1));
    }
    *out = res;
    return MDZ_ST_OK;
}
static int rd_uvarint(rd_t* r, uint64_t* out) {
    uint64_t v = 0;
    for (int sh = 0;; sh += 7) {
        uint8_t b;
        int st = rd_byte(r, &b);
        if (st) return st;
        if (sh < 64) v |= (uint64_t)(b & 0x7f) << sh;
        if (!(b & 0x80)) break;
        if (sh >= 63) return MDZ_ST_BAD_SYMBOL;
    }
    *out = v;
    return MDZ_ST_OK;
}

/* Decodes one lane stream into nv values (truncated to eb bytes).  Stops
 * once nv values are out; a run longer than the values left is an output
 * overrun. */
int mdz_ref_rlev2_lane_decode(const uint8_t* in, uint64_t in_len, int eb, uint64_t* vals,
                              uint64_t nv) {
    rd_t r = {in, in_len, 0};
    const uint64_t vmask = eb == 8 ? ~0ull : ((1ull << (8 * eb)) - 1);
    uint64_t k = 0;
    int st;
    uint8_t b;
    while (k < nv) {
        uint8_t first;
        if ((st = rd_byte(&r, &first))) return st;
        switch (first & 0xC0) {
        case 0x00: {                                   /* SHORT_REPEAT :439-449 */
            uint32_t nb = ((first >> 3) & 7) + 1, cnt = (first & 7) + MIN_REPEAT;
            uint64_t v = 0;
            for (uint32_t i = 0; i < nb; ++i) {
                if ((st = rd_byte(&r, &b))) return st;
                v = (v << 8) | b;
            }
            if (cnt > nv - k) return MDZ_ST_OUTPUT_OVERRUN;
            for (uint32_t i = 0; i < cnt; ++i) vals[k++] = v & vmask;
        } break;
        case 0x40: {                                   /* DIRECT :450-475 */
            const uint32_t w = DEC_W[(first >> 1) & 0x1f];
            if ((st = rd_byte(&r, &b))) return st;
            const uint32_t len = (((uint32_t)first & 1) << 8 | b) + 1;
            if (len > nv - k) return MDZ_ST_OUTPUT_OVERRUN;
            bits_t bs = {0, 0};
            for (uint32_t i = 0; i < len; ++i) {
                uint64_t v;
                if ((st = rd_bits(&r, &bs, w, &v))) return st;
                vals[k++] = v & vmask;
            }
        } break;
        case 0xC0: {                                   /* DELTA :476-519 */
            const uint32_t code = (first >> 1) & 0x1f, w = DEC_W[code];
            if ((st = rd_byte(&r, &b))) return st;
            const uint32_t len = (((uint32_t)first & 1) << 8 | b) + 1;
            if (len > nv - k) return MDZ_ST_OUTPUT_OVERRUN;
            if (len < 2) return MDZ_ST_BAD_SYMBOL;
            uint64_t base, zd;
            if ((st = rd_uvarint(&r, &base))) return st;
            if ((st = rd_uvarint(&r, &zd))) return st;
            const uint64_t delta = (zd >> 1) ^ (0 - (zd & 1));
            uint64_t cur = base & vmask, dt = delta & vmask;
            vals[k++] = cur;
            cur = (cur + dt) & vmask;
            vals[k++] = cur;
            /* multiplier = (DATA_TYPE)delta > 0 ? 1 : -1 on the unsigned type */
            const int neg = dt == 0;
            bits_t bs = {0, 0};
            for (uint32_t i = 2; i < len; ++i) {
                if (code != 0) {
                    uint64_t v;
                    if ((st = rd_bits(&r, &bs, w, &v))) return st;
                    cur = (cur + (neg ? 0 - v : v)) & vmask;
                } else {
                    cur = (cur + dt) & vmask;
                }
                vals[k++] = cur;
            }
        } break;
        default: {                                     /* PATCHED_BASE :520-607 */
            const uint32_t w = DEC_W[(first >> 1) & 0x1f];
            if ((st = rd_byte(&r, &b))) return st;
            const uint32_t len = (((uint32_t)first & 1) << 8 | b) + 1;
            uint8_t third, fourth;
            if ((st = rd_byte(&r, &third))) return st;
            if ((st = rd_byte(&r, &fourth))) return st;
            const uint32_t bw = ((third >> 5) & 7) + 1, pw = DEC_W[third & 0x1f];
            const uint32_t pgw = ((fourth >> 5) & 7) + 1, pll = fourth & 0x1f;
            if (len > nv - k) return MDZ_ST_OUTPUT_OVERRUN;
            uint64_t base = 0;
            for (uint32_t i = 0; i < bw; ++i) {
                if ((st = rd_byte(&r, &b))) return st;
                base = (base << 8) | b;
            }
            const uint64_t k0 = k;
            bits_t bs = {0, 0};
            for (uint32_t i = 0; i < len; ++i) {
                uint64_t v;
                if ((st = rd_bits(&r, &bs, w, &v))) return st;
                vals[k++] = v + base;       /* masked after patching */
            }
            const uint32_t cfb = closest_bit(pw + pgw);
            bits_t ps = {0, 0};
            uint64_t gap = 0;
            const uint64_t pmask = pw >= 64 ? ~0ull : ((1ull << pw) - 1);
            for (uint32_t i = 0; i < pll; ++i) {
                uint64_t v;
                if ((st = rd_bits(&r, &ps, cfb, &v))) return st;
                gap += pw >= 64 ? 0 : v >> pw;
                if (gap >= len) return MDZ_ST_BAD_SYMBOL;
                uint64_t* o = &vals[k0 + gap];
                *o = (((*o - base) | ((v & pmask) << w)) + base);
            }
            for (uint64_t i = k0; i < k; ++i) vals[i] &= vmask;
        } break;
        }
    }
    return MDZ_ST_OK;
}

typedef struct {
    const uint8_t* pay; const uint64_t* off; const uint64_t* cl;
    uint64_t chunk; int eb; uint32_t ru; uint8_t* out; uint64_t out_bytes; int32_t* status;
} dec_ctx;

static void dec_task(void* vctx, uint64_t c) {
    dec_ctx* x = (dec_ctx*)vctx;
    const uint64_t* cl = x->cl + c * 32;
    const uint8_t* src = x->pay + x->off[c];
    uint64_t words[32], maxw = 0, need = 0;
    for (int t = 0; t < 32; ++t) {
        words[t] = (cl[t] + 3) / 4;
        need += words[t] * 4;
        if (words[t] > maxw) maxw = words[t];
    }
    uint8_t* lanes = (uint8_t*)malloc(need + 16);
    uint64_t lbase[32], acc = 0, o = 0;
    for (int t = 0; t < 32; ++t) { lbase[t] = acc; acc += words[t] * 4; }
    for (uint64_t r = 0; r < maxw; ++r)
        for (int t = 0; t < 32; ++t)
            if (r < words[t]) { memcpy(lanes + lbase[t] + r * 4, src + o, 4); o += 4; }
    const uint64_t nv = x->chunk / (uint64_t)x->eb / 32;
    uint64_t* v = (uint64_t*)malloc(nv * 8);
    int st = MDZ_ST_OK;
    for (int t = 0; t < 32; ++t) {
        memset(v, 0, nv * 8);
        const int s = mdz_ref_rlev2_lane_decode(lanes + lbase[t], cl[t], x->eb, v, nv);
        if (s != MDZ_ST_OK && st == MDZ_ST_OK) st = s;
        for (uint64_t q = 0; q < nv; ++q) {
            const uint64_t el = (q / x->ru) * 32 * x->ru + (uint64_t)t * x->ru + q % x->ru;
            const uint64_t dst = c * x->chunk + el * (uint64_t)x->eb;
            if (dst + (uint64_t)x->eb <= x->out_bytes) memcpy(x->out + dst, &v[q], (size_t)x->eb);
            else if (dst < x->out_bytes) memcpy(x->out + dst, &v[q], (size_t)(x->out_bytes - dst));
        }
    }
    free(v);
    free(lanes);
    x->status[c] = st;
}

/* decompress_gpu (decoder_trans.h:615-681); blk_off[n] is not read. */
int mdz_ref_rlev2_decode(const uint8_t* payload, const uint64_t* blk_off, const uint64_t* col_len,
                         uint64_t n_chunks, uint64_t chunk, int eb, uint32_t read_unit, uint8_t* out,
                         uint64_t out_bytes, int32_t* status, int threads) {
    if (!(eb == 1 || eb == 2 || eb == 4 || eb == 8) || read_unit == 0 ||
        chunk % ((uint64_t)eb * 32) || (chunk / (uint64_t)eb / 32) % read_unit)
        return MDZ_ST_BAD_HEADER;
    dec_ctx x = {payload, blk_off, col_len, chunk, eb, read_unit, out, out_bytes, status};
    mdz_parallel_for(n_chunks, threads, dec_task, &x);
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (status[c] != MDZ_ST_OK) return status[c];
    return 0;
}

/* ------------------------------------------------------ plain format -- *
 * block_decode (include/rlev2/decoder.h:190-222) and its four run decoders
 * (:28-188): one run stream per chunk of int64 values, read until the
 * chunk's bytes are consumed.  Each run is checked in this order before any
 * of its values is written: header bytes present (input overrun), DELTA
 * length >= 2 and varints of at most 10 bytes (malformed symbol), values fit
 * the chunk (output overrun), data and patch list present (input overrun),
 * patch positions inside the run (malformed symbol).  The reference checks
 * none of these.  Three places follow the intended values rather than the
 * reference's arithmetic (the plain encoder never relies on them):
 *   - PATCHED_BASE patches are applied with a full pw-bit mask; decoder.h:176
 *     builds it in 16 bits, which drops patch bits at positions >= 16;
 *   - a packed-delta run writes its len values only; decoder.h:117-124 runs
 *     its prefix loop over len entries from the third value on, so it
 *     rewrites the two values after the run (later runs overwrite them; after
 *     a chunk's last run they land in the next chunk);
 *   - a patch entry shifted past 63 bits contributes nothing.
 * DELTA runs ascend when the zigzag delta base is positive and descend
 * otherwise (decoder.h:115), with absolute packed deltas. */
static int rd_plain_varint(rd_t* r, uint64_t* out) {
    uint64_t v = 0;
    for (uint32_t i = 0;; ++i) {
        uint8_t b;
        if (i == 10) return MDZ_ST_BAD_SYMBOL;
        if (r->p >= r->len) return MDZ_ST_INPUT_OVERRUN;
        b = r->in[r->p++];
        v |= (uint64_t)(b & 0x7f) << (7 * i);           /* a 10th byte keeps bit 0 */
        if (!(b & 0x80)) break;
    }
    *out = v;
    return MDZ_ST_OK;
}

int mdz_ref_rlev2_plain_chunk_decode(const uint8_t* in, uint64_t in_len, uint64_t* vals, uint64_t cap,
                                     uint64_t* n_out) {
    rd_t r = {in, in_len, 0};
    uint64_t k = 0;
    int st = MDZ_ST_OK;
    while (r.p < in_len) {
        const uint64_t p0 = r.p;
        const uint8_t first = in[r.p++];
        const uint32_t kind = first >> 6;
        if (kind == 0) {                                   /* SHORT_REPEAT :28-40 */
            const uint32_t nb = ((first >> 3) & 7) + 1, cnt = (first & 7) + MIN_REPEAT;
            if (p0 + 1 + nb > in_len) { st = MDZ_ST_INPUT_OVERRUN; break; }
            if (cnt > cap - k) { st = MDZ_ST_OUTPUT_OVERRUN; break; }
            uint64_t v = 0;
            for (uint32_t i = 0; i < nb; ++i) v = (v << 8) | in[r.p++];
            for (uint32_t i = 0; i < cnt; ++i) vals[k++] = v;
            continue;
        }
        if (p0 + 2 > in_len) { st = MDZ_ST_INPUT_OVERRUN; break; }
        const uint32_t w = DEC_W[(first >> 1) & 0x1f], code = (first >> 1) & 0x1f;
        const uint32_t len = (((uint32_t)first & 1) << 8 | in[r.p++]) + 1;
        if (kind == 1) {                                   /* DIRECT :134-141 */
            if (len > cap - k) { st = MDZ_ST_OUTPUT_OVERRUN; break; }
            if (r.p + ((uint64_t)len * w + 7) / 8 > in_len) { st = MDZ_ST_INPUT_OVERRUN; break; }
            bits_t bs = {0, 0};
            for (uint32_t i = 0; i < len; ++i) {
                uint64_t v;
                rd_bits(&r, &bs, w, &v);
                vals[k++] = v;
            }
        } else if (kind == 3) {                            /* DELTA :94-132 */
            if (len < 2) { st = MDZ_ST_BAD_SYMBOL; break; }
            uint64_t base, zz;
            if ((st = rd_plain_varint(&r, &base))) break;
            if ((st = rd_plain_varint(&r, &zz))) break;
            if (len > cap - k) { st = MDZ_ST_OUTPUT_OVERRUN; break; }
            const uint64_t nd = code ? len - 2 : 0;
            if (r.p + (nd * w + 7) / 8 > in_len) { st = MDZ_ST_INPUT_OVERRUN; break; }
            const uint64_t bd = (zz >> 1) ^ (0 - (zz & 1));
            uint64_t cur = base;
            vals[k++] = cur;
            cur += bd;
            vals[k++] = cur;
            const int up = (int64_t)bd > 0;
            bits_t bs = {0, 0};
            for (uint32_t i = 2; i < len; ++i) {
                if (code) {
                    uint64_t d;
                    rd_bits(&r, &bs, w, &d);
                    cur = up ? cur + d : cur - d;
                } else {
                    cur += bd;
                }
                vals[k++] = cur;
            }
        } else {                                           /* PATCHED_BASE :143-188 */
            if (p0 + 4 > in_len) { st = MDZ_ST_INPUT_OVERRUN; break; }
            const uint8_t third = in[r.p++], fourth = in[r.p++];
            const uint32_t bw = ((third >> 5) & 7) + 1, pw = DEC_W[third & 0x1f];
            const uint32_t pgw = ((fourth >> 5) & 7) + 1, pll = fourth & 0x1f;
            const uint32_t cfb = closest_bit(pw + pgw);
            if (len > cap - k) { st = MDZ_ST_OUTPUT_OVERRUN; break; }
            const uint64_t need = bw + ((uint64_t)len * w + 7) / 8 + ((uint64_t)pll * cfb + 7) / 8;
            if (r.p + need > in_len) { st = MDZ_ST_INPUT_OVERRUN; break; }
            uint64_t base = 0;
            for (uint32_t i = 0; i < bw; ++i) base = (base << 8) | in[r.p++];
            uint64_t tmp[512];
            bits_t bs = {0, 0};
            for (uint32_t i = 0; i < len; ++i) rd_bits(&r, &bs, w, &tmp[i]);
            bits_t ps = {0, 0};
            uint64_t gap = 0;
            const uint64_t pmask = pw >= 64 ? ~0ull : ((1ull << pw) - 1);
            for (uint32_t i = 0; i < pll && st == MDZ_ST_OK; ++i) {
                uint64_t e;
                rd_bits(&r, &ps, cfb, &e);
                gap += pw >= 64 ? 0 : e >> pw;
                if (gap >= len) { st = MDZ_ST_BAD_SYMBOL; break; }
                tmp[gap] |= w >= 64 ? 0 : (e & pmask) << w;
            }
            if (st) break;
            for (uint32_t i = 0; i < len; ++i) vals[k++] = tmp[i] + base;
        }
    }
    *n_out = k;
    return st;
}

typedef struct {
    const uint8_t* data; const uint64_t* ptr; uint64_t data_bytes, chunk;
    uint8_t* out; uint64_t out_bytes; int32_t* status;
} plain_dec_ctx;

static void plain_dec_task(void* vctx, uint64_t c) {
    plain_dec_ctx* x = (plain_dec_ctx*)vctx;
    const uint64_t a = x->ptr[c], b = x->ptr[c + 1];
    if (a > b || b > x->data_bytes) { x->status[c] = MDZ_ST_BAD_HEADER; return; }
    const uint64_t cap = x->chunk / 8;
    uint64_t* v = (uint64_t*)malloc((cap ? cap : 1) * 8);
    uint64_t n = 0;
    x->status[c] = mdz_ref_rlev2_plain_chunk_decode(x->data + a, b - a, v, cap, &n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t o = c * x->chunk + 8 * i;
        if (o + 8 <= x->out_bytes) memcpy(x->out + o, &v[i], 8);
        else if (o < x->out_bytes) memcpy(x->out + o, &v[i], (size_t)(x->out_bytes - o));
    }
    free(v);
}

/* decompress_gpu (decoder.h:227-262) over a container in memory:
 * [u32 n+1][u64 ptr[n+1]][u64 raw_bytes][streams]; chunk c's values go to
 * byte c*chunk of out (out_bytes = raw_bytes normally).  Returns -1 for a
 * container whose header does not fit, else the first failing chunk's
 * status (0 when all are good). */
int mdz_ref_rlev2_plain_decode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t* out,
                               uint64_t out_bytes, int32_t* status, uint64_t n_status, int threads) {
    if (in_bytes < 4 || chunk == 0 || chunk % 8) return -1;
    uint32_t np;
    memcpy(&np, in, 4);
    if (np == 0 || 4 + 8 * (uint64_t)np + 8 > in_bytes || (uint64_t)np - 1 > n_status) return -1;
    const uint64_t n = np - 1, hdr = 4 + 8 * (uint64_t)np + 8;
    uint64_t* ptr = (uint64_t*)malloc(8 * (uint64_t)np);
    memcpy(ptr, in + 4, 8 * (uint64_t)np);
    plain_dec_ctx x = {in + hdr, ptr, in_bytes - hdr, chunk, out, out_bytes, status};
    mdz_parallel_for(n, threads, plain_dec_task, &x);
    free(ptr);
    for (uint64_t c = 0; c < n; ++c)
        if (status[c] != MDZ_ST_OK) return status[c];
    return 0;
}
