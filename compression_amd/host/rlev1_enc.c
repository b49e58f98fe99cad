/* RLE v1 lane-transposed encoder (host side).
 *
 * Restates include/rle_v1/encoder.h: compression_warp :607-1018, the 4-byte
 * compacted round-robin interleave of write_byte_op :564-588, col_len/blk_off
 * :540-559 and :1188-1193. */
#include "mdz_host.h"

#include <stdlib.h>
#include <string.h>

static inline uint64_t ld_val(const uint8_t* p, uint64_t i, int eb) {
    uint64_t v = 0;
    memcpy(&v, p + i * (uint64_t)eb, (size_t)eb);   /* little-endian host */
    return v;
}
static inline void st_val(uint8_t* p, uint64_t i, int eb, uint64_t v) {
    memcpy(p + i * (uint64_t)eb, &v, (size_t)eb);
}

/* `int64_t temp_diff = read_data - prev_val` evaluated in DATA_TYPE
 * arithmetic (encoder.h:264,314,434,664,744,879).  u8/u16 promote to int
 * (signed difference); u32 subtracts modulo 2^32 and zero-extends, so a
 * "negative" delta is a large positive one; u64 subtracts modulo 2^64 and the
 * conversion to int64 wraps. */
static inline int64_t ref_diff(uint64_t rd, uint64_t pv, int eb) {
    switch (eb) {
    case 1: case 2: return (int64_t)((int32_t)rd - (int32_t)pv);
    case 4: return (int64_t)(uint32_t)((uint32_t)rd - (uint32_t)pv);
    default: return (int64_t)(rd - pv);
    }
}

/* `prev_val + cur_delta == read_data` (encoder.h:763) in DATA_TYPE arithmetic:
 * u8/u16 promote to int (no wrap), u32/u64 wrap. */
static inline int ref_match(uint64_t pv, int8_t d, uint64_t rd, int eb) {
    switch (eb) {
    case 1: case 2: return (int32_t)pv + (int32_t)d == (int32_t)rd;
    case 4: return (uint32_t)((uint32_t)pv + (uint32_t)(int32_t)d) == (uint32_t)rd;
    default: return (uint64_t)(pv + (uint64_t)(int64_t)d) == rd;
    }
}

typedef struct { uint8_t* out; uint64_t pos, cap; } wbuf;
static inline void put_byte(wbuf* w, uint8_t b) {
    if (w->pos < w->cap) w->out[w->pos] = b;
    w->pos++;
}
static inline void put_varint(wbuf* w, uint64_t v) {   /* unsigned LEB128 (write_varint_op :589-604) */
    do {
        uint8_t b = (uint8_t)(v & 0x7F);
        if (v & ~(uint64_t)0x7F) b |= 0x80;
        put_byte(w, b);
        v >>= 7;
    } while (v != 0);
}
static inline void patch(wbuf* w, uint64_t at, int lit) {
    if (at < w->cap) w->out[at] = (uint8_t)(int8_t)(-lit);
}

/* Lane encoder.  State mirrors compression_warp: a two-entry value ring
 * (buf/head/pend = data_buffer/_head/_count), the run being grown
 * (dcnt/delta/dok/first = delta_count/cur_delta/delta_flag/delta_first_val)
 * and the open literal group (lit/lit_at = lit_count/lit_idx, whose header
 * byte is written as a placeholder and back-patched to -lit). */
uint64_t mdz_rlev1_lane_encode(const void* vals, uint64_t n, int eb,
                                   uint8_t* out, uint64_t cap) {
    const uint8_t* v = (const uint8_t*)vals;
    wbuf w = {out, 0, cap};
    uint64_t buf[2] = {0, 0};
    int head = 0, pend = 0;
    int dcnt, dok = 0, lit = 0;
    int8_t delta = 0;
    uint64_t first = 0, lit_at = 0, prev;
    if (n < 2) return 0;

    /* first two values (:651-699) */
    uint64_t x = ld_val(v, 0, eb);
    dcnt = 1; prev = x;
    buf[head] = x; head ^= 1; pend++;
    x = ld_val(v, 1, eb);
    int64_t td = ref_diff(x, prev, eb);
    if (td > 127 || td < -128) {
        dok = 0;
        lit_at = w.pos; put_varint(&w, 1); lit = 1;
        put_varint(&w, buf[0]);
        pend--;
    } else {
        dok = 1; delta = (int8_t)td; dcnt++;
    }
    prev = x; buf[head] = x; head ^= 1; pend++;

    for (uint64_t i = 2; i < n; ++i) {                       /* main loop :704-953 */
        if (lit == 127) { patch(&w, lit_at, lit); lit = 0; }
        x = ld_val(v, i, eb);

        if (dcnt == 1) {                                     /* :725-757 */
            td = ref_diff(x, prev, eb);
            if (td > 127 || td < -128) {
                dok = 0;
                if (lit == 0) { lit_at = w.pos; put_varint(&w, 1); }
                put_varint(&w, buf[head ^ 1]);
                lit++; pend--;
            } else {
                dok = 1; delta = (int8_t)td; dcnt++;
            }
            prev = x; buf[head] = x; head ^= 1; pend++;
            continue;
        }

        if (dok && ref_match(prev, delta, x, eb)) {          /* matched :763-801 */
            dcnt++;
            if (dcnt == 3) {
                first = buf[head];
                if (lit != 0) { patch(&w, lit_at, lit); lit = 0; }
            } else if (dcnt == 130) {                        /* run cap (B2): 129 values */
                put_byte(&w, (uint8_t)(int8_t)(dcnt - 4));
                put_byte(&w, (uint8_t)delta);
                put_varint(&w, first);
                dcnt = 1; pend = 0; lit = 0;
            }
            buf[head] = x; head ^= 1; pend = pend + 1 < 2 ? pend + 1 : 2;
            prev = x;
        } else if (dcnt >= 3) {                              /* run ends :809-839 */
            put_byte(&w, (uint8_t)(int8_t)(dcnt - 3));
            put_byte(&w, (uint8_t)delta);
            put_varint(&w, first);
            dcnt = 1; pend = 0; lit = 0;
            buf[head] = x; head ^= 1; pend = 1;
            prev = x;
        } else {                                             /* pair broken :841-901 */
            if (lit == 0) { lit_at = w.pos; put_byte(&w, 3); }
            lit++;
            put_varint(&w, buf[head]);
            if (lit == 127) { patch(&w, lit_at, lit); lit = 0; }
            td = ref_diff(x, prev, eb);
            if (td > 127 || td < -128) {
                if (lit == 0) { lit_at = w.pos; put_byte(&w, 3); }
                dok = 0; pend = 0;
                put_varint(&w, buf[head ^ 1]);
                lit++; dcnt = 1;
            } else {
                pend = 1; dok = 1; delta = (int8_t)td; dcnt = 2;
            }
            prev = x; buf[head] = x; head ^= 1;
            pend = pend + 1 < 2 ? pend + 1 : 2;
        }
    }

    if (dcnt >= 3 && dok) {                                  /* tail run :957-972 */
        put_byte(&w, (uint8_t)(int8_t)(dcnt - 3));
        put_byte(&w, (uint8_t)delta);
        put_varint(&w, first);
    } else {                                                 /* tail literals :974-1007 */
        if (pend == 1) {
            if (lit == 127) { patch(&w, lit_at, lit); lit = 0; }
            if (lit == 0) { lit_at = w.pos; put_varint(&w, 1); }
            put_varint(&w, buf[head ^ 1]);
            lit++;
        }
        if (pend == 2) {
            if (lit == 127) { patch(&w, lit_at, lit); lit = 0; }
            if (lit == 0) { lit_at = w.pos; put_byte(&w, 1); }
            put_varint(&w, buf[head]);
            lit++;
            if (lit == 127) { patch(&w, lit_at, lit); lit = 0; }
            if (lit == 0) { lit_at = w.pos; put_byte(&w, 1); }
            head ^= 1;
            put_varint(&w, buf[head]);
            lit++;
        }
        patch(&w, lit_at, lit);                              /* B3: unconditional */
    }
    return w.pos;
}

typedef struct {
    const uint8_t* in; uint64_t in_bytes, chunk; int eb; uint32_t line;
    uint8_t** lane_bufs; uint64_t* col_len; uint64_t* chunk_len;
} enc_ctx;

/* Gather lane t's values: its lines t, t+32, ... of the zero-padded chunk
 * (compression_reader_warp, encoder.h:137-153). */
static void gather_lane(const uint8_t* in, uint64_t in_bytes, uint64_t cbase, uint64_t chunk,
                        uint32_t line, int t, uint8_t* dst) {
    uint64_t lane_bytes = chunk / 32, o = 0;
    for (uint64_t ln = (uint64_t)t; o < lane_bytes; ln += 32) {
        uint64_t src = cbase + ln * line;
        for (uint32_t j = 0; j < line; ++j, ++o) {
            uint64_t s = src + j;
            dst[o] = s < in_bytes ? in[s] : 0;
        }
    }
}

static void enc_task(void* vctx, uint64_t c) {
    enc_ctx* x = (enc_ctx*)vctx;
    uint64_t lane_bytes = x->chunk / 32, n = lane_bytes / (uint64_t)x->eb;
    uint8_t* vals = (uint8_t*)malloc(lane_bytes);
    uint64_t cap = n * 11 + 16, total = 0;
    for (int t = 0; t < 32; ++t) {
        gather_lane(x->in, x->in_bytes, c * x->chunk, x->chunk, x->line, t, vals);
        uint8_t* s = (uint8_t*)malloc(cap);
        uint64_t len = mdz_rlev1_lane_encode(vals, n, x->eb, s, cap);
        x->lane_bufs[c * 32 + t] = s;
        x->col_len[c * 32 + t] = len;
        total += (len + 3) / 4 * 4;
    }
    x->chunk_len[c] = (total + 127) / 128 * 128;
    free(vals);
}

/* Compacted round-robin interleave of 4-byte words in lane order
 * (write_byte_op, encoder.h:564-588; read side comp_read_data,
 * decoder.h:72-95).  Pad bytes (B10) are zero. */
static void interleave(uint8_t* dst, uint8_t* const* lanes, const uint64_t* len) {
    uint64_t words[32], maxw = 0, o = 0;
    for (int t = 0; t < 32; ++t) {
        words[t] = (len[t] + 3) / 4;
        if (words[t] > maxw) maxw = words[t];
    }
    for (uint64_t r = 0; r < maxw; ++r)
        for (int t = 0; t < 32; ++t)
            if (r < words[t]) {
                for (int b = 0; b < 4; ++b) {
                    uint64_t s = r * 4 + (uint64_t)b;
                    dst[o++] = s < len[t] ? lanes[t][s] : 0;
                }
            }
}

int mdz_rlev1_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk,
                         int eb, uint32_t line,
                         uint8_t** payload, uint64_t* payload_bytes,
                         uint64_t** col_len, uint64_t** blk_off, uint64_t* n_chunks,
                         int threads) {
    if (chunk % 32 || line == 0 || (chunk / 32) % line || line % (uint32_t)eb ||
        (chunk / 32) / (uint64_t)eb < 2 || in_bytes == 0)
        return -1;
    uint64_t n = (in_bytes + chunk - 1) / chunk;
    enc_ctx x = {in, in_bytes, chunk, eb, line,
                 (uint8_t**)calloc(n * 32, sizeof(uint8_t*)),
                 (uint64_t*)malloc(n * 32 * 8), (uint64_t*)malloc(n * 8)};
    mdz_parallel_for(n, threads, enc_task, &x);
    uint64_t* off = (uint64_t*)malloc((n + 1) * 8);
    off[0] = 0;
    for (uint64_t c = 0; c < n; ++c) off[c + 1] = off[c] + x.chunk_len[c];
    uint8_t* pay = (uint8_t*)calloc(off[n] ? off[n] : 1, 1);
    for (uint64_t c = 0; c < n; ++c) {
        interleave(pay + off[c], x.lane_bufs + c * 32, x.col_len + c * 32);
        for (int t = 0; t < 32; ++t) free(x.lane_bufs[c * 32 + t]);
    }
    free(x.lane_bufs);
    free(x.chunk_len);
    *payload = pay; *payload_bytes = off[n];
    *col_len = x.col_len; *blk_off = off; *n_chunks = n;
    return 0;
}

