/* mdz oracle: RLE v1 (lane-transposed).  TEST INFRASTRUCTURE ONLY.
 *
 * Restates include/rle_v1/encoder.h (compression_warp :607-1018, the 4-byte
 * compacted round-robin interleave of write_byte_op :564-588, col_len/blk_off
 * :540-559 and :1188-1193) and include/rle_v1/decoder.h (decoder_warp
 * :275-423, comp_read_data :72-95). */
#include "mdz_oracle.h"

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

/* Unsigned LEB128 of at most the canonical length for a DATA_TYPE of eb
 * bytes (2, 3, 5, 10 bytes for eb = 1, 2, 4, 8).  A stream that ends inside
 * the value is an input overrun; a longer value is a malformed symbol. */
static int leb128(const uint8_t* in, uint64_t in_len, uint64_t* pp, int eb, uint64_t* out) {
    const int maxb = eb == 1 ? 2 : eb == 2 ? 3 : eb == 4 ? 5 : 10;
    uint64_t v = 0, p = *pp;
    for (int i = 0;; ++i) {
        if (i >= maxb) return MDZ_ST_BAD_SYMBOL;
        if (p >= in_len) return MDZ_ST_INPUT_OVERRUN;
        uint8_t b = in[p++];
        if (7 * i < 64) v |= (uint64_t)(b & 0x7F) << (7 * i);
        if (!(b & 0x80)) break;
    }
    *pp = p;
    *out = v;
    return MDZ_ST_OK;
}

/* Lane decoder (decoder_warp, decoder.h:275-423).  Literal group: header
 * h < 0, -h LEB128 values.  Run: h >= 0, int8 delta, LEB128 base, h+3 values
 * (T)(base + i*delta) computed in 64-bit.  The reference stops after the
 * lane's CHUNK/32 bytes; this restatement also reports streams that end early
 * (status 2), tokens whose count overruns the lane (status 1, checked at the
 * header) and over-long LEB128 values (status -10). */
int mdz_ref_rlev1_lane_decode(const uint8_t* in, uint64_t in_len, int eb,
                              void* vals, uint64_t n_vals) {
    uint8_t* o = (uint8_t*)vals;
    uint64_t p = 0, k = 0;
    const uint64_t vmask = eb == 8 ? ~0ull : ((1ull << (8 * eb)) - 1);
    while (k < n_vals) {
        if (p >= in_len) return MDZ_ST_INPUT_OVERRUN;
        int8_t h = (int8_t)in[p++];
        if (h < 0) {
            uint64_t cnt = (uint64_t)(-(int)h);
            if (cnt > n_vals - k) return MDZ_ST_OUTPUT_OVERRUN;
            for (uint64_t i = 0; i < cnt; ++i) {
                uint64_t val = 0;
                int st = leb128(in, in_len, &p, eb, &val);
                if (st != MDZ_ST_OK) return st;
                st_val(o, k++, eb, val & vmask);
            }
        } else {
            uint64_t cnt = (uint64_t)h + 3;
            if (cnt > n_vals - k) return MDZ_ST_OUTPUT_OVERRUN;
            if (p >= in_len) return MDZ_ST_INPUT_OVERRUN;
            int8_t d = (int8_t)in[p++];
            uint64_t base = 0;
            int st = leb128(in, in_len, &p, eb, &base);
            if (st != MDZ_ST_OK) return st;
            base &= vmask;
            for (uint64_t i = 0; i < cnt; ++i)
                st_val(o, k++, eb, (base + i * (uint64_t)(int64_t)d) & vmask);
        }
    }
    return MDZ_ST_OK;
}

/* ---------------------------------------------------------- container ---- */
typedef struct {
    const uint8_t* pay; const uint64_t* off; const uint64_t* cl;
    uint64_t chunk; int eb; uint32_t line; uint8_t* out; uint64_t out_bytes;
    int32_t* status;
} dec_ctx;

static void dec_task(void* vctx, uint64_t c) {
    dec_ctx* x = (dec_ctx*)vctx;
    const uint64_t* cl = x->cl + c * 32;
    const uint8_t* src = x->pay + x->off[c];
    uint64_t avail = x->off[c + 1] - x->off[c];
    uint64_t lane_bytes = x->chunk / 32, words[32], maxw = 0, need = 0;
    for (int t = 0; t < 32; ++t) {
        words[t] = (cl[t] + 3) / 4;
        need += words[t] * 4;
        if (words[t] > maxw) maxw = words[t];
    }
    if (need > avail) { x->status[c] = MDZ_ST_INPUT_OVERRUN; return; }
    /* de-interleave: round r holds word r of every lane with > r words */
    uint8_t* lanes = (uint8_t*)malloc(need + 32 * 4);
    uint64_t lbase[32], acc = 0, o = 0;
    for (int t = 0; t < 32; ++t) { lbase[t] = acc; acc += words[t] * 4; }
    for (uint64_t r = 0; r < maxw; ++r)
        for (int t = 0; t < 32; ++t)
            if (r < words[t]) { memcpy(lanes + lbase[t] + r * 4, src + o, 4); o += 4; }
    uint8_t* vals = (uint8_t*)malloc(lane_bytes);
    int st = MDZ_ST_OK;
    for (int t = 0; t < 32; ++t) {
        int s = mdz_ref_rlev1_lane_decode(lanes + lbase[t], cl[t], x->eb, vals,
                                          lane_bytes / (uint64_t)x->eb);
        if (s != MDZ_ST_OK && st == MDZ_ST_OK) st = s;
        /* scatter lane t's lines (decoder.h:332-341,398-405) */
        for (uint64_t ln = (uint64_t)t, v = 0; v < lane_bytes; ln += 32, v += x->line) {
            uint64_t dst = c * x->chunk + ln * x->line;
            if (dst >= x->out_bytes) break;
            uint64_t m = x->line;
            if (dst + m > x->out_bytes) m = x->out_bytes - dst;
            memcpy(x->out + dst, vals + v, m);
        }
    }
    free(vals);
    free(lanes);
    x->status[c] = st;
}

int mdz_ref_rlev1_decode(const uint8_t* payload, const uint64_t* blk_off,
                         const uint64_t* col_len, uint64_t n_chunks, uint64_t chunk,
                         int eb, uint32_t line, uint8_t* out, uint64_t out_bytes,
                         int32_t* status, int threads) {
    if (chunk % 32 || line == 0 || (chunk / 32) % line || line % (uint32_t)eb)
        return MDZ_ST_BAD_HEADER;
    dec_ctx x = {payload, blk_off, col_len, chunk, eb, line, out, out_bytes, status};
    mdz_parallel_for(n_chunks, threads, dec_task, &x);
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (status[c] != MDZ_ST_OK) return status[c];
    return 0;
}
