/*
 * LZSS, lane-transposed container -- CPU restatement (TEST INFRASTRUCTURE ONLY).
 *
 * Token grammar (include/lzss/lzss_gpu2.h:20-67,209-297):
 *   window 2048 bytes, offset field 11 bits, length field 5 bits,
 *   match lengths 3..34.  A flag byte precedes up to 8 tokens; bit i (LSB
 *   first) is 1 for a literal byte, 0 for a match stored as 2 bytes
 *   little-endian v = (dist << 5) | (len - 3), dist in [1, 2047] back from
 *   the end of the lane's output so far.
 *
 * Container (lzss_gpu2.h:332-707,1063-1085,1952-2123):
 *   lane t of chunk c is the byte sequence chunk[128k + 4t + j] (j < 4);
 *   each lane is encoded independently; lanes are sorted by stream length,
 *   descending, ties to the higher lane id; col_len[32c+s] is the length of
 *   the lane at sorted position s and col_map[32c+s] its original lane id;
 *   the sorted streams are interleaved in 4-byte words, word k of every
 *   sorted lane with more than k words, in sorted order; each chunk is padded
 *   to 128 bytes; blk_off is the exclusive prefix of padded chunk sizes.
 */
#include "mdz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { HIST = 2048, MIN_LEN = 3, MAX_LEN = 34, LEN_BITS = 5 };

/* ---------------------------------------------------------------- match -- */

/* Restates lzss::find_match (lzss_gpu2.h:70-109) over a lane held whole in
 * memory: the history is the last min(cur, 2048) bytes, offsets 1..count-1
 * are tried oldest first (offset 0, the oldest byte, never is), a match may
 * not run past the end of the history or of the lane, the first longest
 * match wins and a 34-byte match stops the search.  Returns the length
 * (0 if below 3) and the distance in *dist.  O(window) per call: used to pin
 * the host encoder's hash-chain matcher. */
static uint32_t find_match_naive(const uint8_t* s, uint64_t n, uint64_t cur, uint32_t* dist) {
    const uint64_t count = cur < HIST ? cur : HIST;
    const uint64_t start = cur - count;
    const uint64_t look = n - cur;
    uint32_t best = 0, best_off = 0;
    for (uint64_t off = 1; off < count; ++off) {
        const uint64_t q = start + off;
        if (s[q] != s[cur]) continue;
        uint32_t f = 1;
        while (off + f < count && f < look && s[q + f] == s[cur + f]) {
            if (++f >= MAX_LEN) break;
        }
        if (f > best) { best = f; best_off = (uint32_t)off; }
        if (f >= MAX_LEN) { best = MAX_LEN; break; }
    }
    if (best < MIN_LEN) best = 0;
    *dist = (uint32_t)(count - best_off);
    return best;
}

/* One lane -> token stream with the literal find_match scan: the
 * reference's flag-byte layout (compress_init_func / compress_func_new,
 * lzss_gpu2.h:332-494,511-707: first flag byte reserved up front, the next
 * one right after every 8th token, so 8k tokens end in a reserved byte that
 * is counted and written here as 0).  O(window) per byte: the checker for
 * the host encoder's hash-chain matcher. */
uint64_t mdz_ref_lzss_lane_encode_naive(const uint8_t* s, uint64_t n, uint8_t* out, uint64_t cap) {
    uint64_t o = 1, flag_at = 0;
    uint32_t ntok = 0;
    uint8_t flags = 0;
    if (cap) out[0] = 0;
    for (uint64_t cur = 0; cur < n;) {
        uint32_t dist = 0;
        const uint32_t len = find_match_naive(s, n, cur, &dist);
        if (len == 0) {
            flags |= (uint8_t)(1u << ntok);
            if (o < cap) out[o] = s[cur];
            o += 1;
            cur += 1;
        } else {
            const uint32_t v = (dist << LEN_BITS) | (len - MIN_LEN);
            if (o < cap) out[o] = (uint8_t)v;
            if (o + 1 < cap) out[o + 1] = (uint8_t)(v >> 8);
            o += 2;
            cur += len;
        }
        if (++ntok == 8) {
            if (flag_at < cap) out[flag_at] = flags;
            flags = 0;
            ntok = 0;
            flag_at = o++;
            if (flag_at < cap) out[flag_at] = 0;
        }
    }
    if (ntok && flag_at < cap) out[flag_at] = flags;
    return o;
}

/* One token stream -> exactly out_len bytes.  Restates the transposed
 * decompress_func (lzss_gpu2.h:114-325) with these bounds: decoding stops
 * as soon as the lane's output is complete (the reference decodes the
 * reserved trailing flag byte as a token group and writes past the lane);
 * a token that needs bytes past the stream is INPUT_OVERRUN, as is a stream
 * that ends before the output is complete; a match with distance 0 or
 * reaching before the lane's first byte is BAD_DISTANCE; a match longer than
 * the output left is OUTPUT_OVERRUN (the bytes that fit are written).
 * Matches copy byte by byte, so an overlapping match repeats its source. */
int mdz_ref_lzss_lane_decode(const uint8_t* in, uint64_t in_len, uint8_t* out, uint64_t out_len) {
    uint64_t p = 0, o = 0;
    while (o < out_len) {
        if (p >= in_len) return MDZ_ST_INPUT_OVERRUN;
        const uint32_t flags = in[p++];
        for (uint32_t b = 0; b < 8 && o < out_len; ++b) {
            if (flags >> b & 1u) {
                if (p >= in_len) return MDZ_ST_INPUT_OVERRUN;
                out[o++] = in[p++];
            } else {
                if (p + 2 > in_len) return MDZ_ST_INPUT_OVERRUN;
                const uint32_t v = (uint32_t)in[p] | ((uint32_t)in[p + 1] << 8);
                p += 2;
                const uint32_t len = (v & ((1u << LEN_BITS) - 1)) + MIN_LEN;
                const uint32_t dist = v >> LEN_BITS;
                if (dist == 0 || dist > o) return MDZ_ST_BAD_DISTANCE;
                for (uint32_t i = 0; i < len; ++i) {
                    if (o >= out_len) return MDZ_ST_OUTPUT_OVERRUN;
                    out[o] = out[o - dist];
                    ++o;
                }
            }
        }
    }
    return MDZ_ST_OK;
}

typedef int (*lane_decoder)(const uint8_t* in, uint64_t in_len, uint8_t* out, uint64_t out_len);

typedef struct {
    const uint8_t* pay; const uint64_t* off; const uint64_t* cl; const uint8_t* cm;
    uint64_t chunk; uint64_t pay_bytes; uint8_t* out; uint64_t out_bytes; int32_t* status;
    lane_decoder dec;
} dec_ctx;

/* Container checks shared with the HIP decoder (status BAD_HEADER): col_map
 * a permutation of 0..31, col_len non-increasing and no longer than the
 * longest stream the encoder can emit (all literals, their flag bytes and the
 * reserved byte), blk_off[c] a multiple of 4, the interleaved words inside
 * the chunk's [blk_off[c], blk_off[c+1]) extent and inside the payload. */
static void dec_task(void* vctx, uint64_t c) {
    dec_ctx* x = (dec_ctx*)vctx;
    const uint64_t* L = x->cl + 32 * c;
    const uint8_t* cm = x->cm + 32 * c;
    const uint64_t lb = x->chunk / 32;
    const uint64_t cap_stream = lb + (lb + 7) / 8 + 1;   /* all literals + flags + reserved */
    uint32_t seen = 0;
    uint64_t words = 0;
    for (int s = 0; s < 32; ++s) {
        if (cm[s] > 31 || (seen >> cm[s] & 1u) || (s && L[s] > L[s - 1]) || L[s] > cap_stream) {
            x->status[c] = MDZ_ST_BAD_HEADER;
            return;
        }
        seen |= 1u << cm[s];
        words += (L[s] + 3) / 4;
    }
    if (x->off[c + 1] < x->off[c] || x->off[c + 1] > x->pay_bytes || (x->off[c] & 3) ||
        4 * words > x->off[c + 1] - x->off[c]) {
        x->status[c] = MDZ_ST_BAD_HEADER;
        return;
    }
    uint8_t* lanes = (uint8_t*)malloc(4 * words + 4);
    uint64_t base[32], acc = 0;
    for (int s = 0; s < 32; ++s) { base[s] = acc; acc += 4 * ((L[s] + 3) / 4); }
    const uint8_t* src = x->pay + x->off[c];
    for (uint64_t k = 0;; ++k) {
        int any = 0;
        for (int s = 0; s < 32 && (L[s] + 3) / 4 > k; ++s) {
            memcpy(lanes + base[s] + 4 * k, src, 4);
            src += 4;
            any = 1;
        }
        if (!any) break;
    }
    uint8_t* vals = (uint8_t*)calloc(lb, 1);
    int st = MDZ_ST_OK;
    for (int t = 0; t < 32; ++t) {              /* original lane order */
        int s = 0;
        while (cm[s] != t) ++s;
        memset(vals, 0, lb);
        const int r = x->dec(lanes + base[s], L[s], vals, lb);
        if (r != MDZ_ST_OK && st == MDZ_ST_OK) st = r;
        for (uint64_t i = 0; i < lb; i += 4) {
            const uint64_t d = c * x->chunk + 128 * (i / 4) + 4 * (uint64_t)t;
            for (uint64_t j = 0; j < 4 && d + j < x->out_bytes; ++j) x->out[d + j] = vals[i + j];
        }
    }
    free(vals);
    free(lanes);
    x->status[c] = st;
}

int mdz_ref_lzss_decode(const uint8_t* payload, uint64_t payload_bytes, const uint64_t* blk_off,
                        const uint64_t* col_len, const uint8_t* col_map, uint64_t n_chunks,
                        uint64_t chunk, uint8_t* out, uint64_t out_bytes, int32_t* status,
                        int threads) {
    if (chunk == 0 || chunk % 128) return MDZ_ST_BAD_HEADER;
    dec_ctx x = {payload, blk_off, col_len, col_map, chunk, payload_bytes, out, out_bytes, status,
                 mdz_ref_lzss_lane_decode};
    mdz_parallel_for(n_chunks, threads, dec_task, &x);
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (status[c] != MDZ_ST_OK) return status[c];
    return 0;
}


/* brle_trans::decompress_func (include/brle/brle_trans.h:72-202) for one
 * lane: flag byte, then 8 tokens, bit b (LSB first) 1 = literal byte, 0 =
 * run (count byte, value byte: `count` copies of value).  Output-driven with
 * mdz_ref_lzss_lane_decode's bounds and statuses; the reference reads its
 * col_len bytes and writes whatever they decode to. */
int mdz_ref_brle_trans_lane_decode(const uint8_t* in, uint64_t in_len, uint8_t* out, uint64_t out_len) {
    uint64_t p = 0, o = 0;
    while (o < out_len) {
        if (p >= in_len) return MDZ_ST_INPUT_OVERRUN;
        const uint32_t flags = in[p++];
        for (uint32_t b = 0; b < 8 && o < out_len; ++b) {
            if (flags >> b & 1u) {
                if (p >= in_len) return MDZ_ST_INPUT_OVERRUN;
                out[o++] = in[p++];
            } else {
                if (p + 2 > in_len) return MDZ_ST_INPUT_OVERRUN;
                const uint32_t cnt = in[p], v = in[p + 1];
                p += 2;
                for (uint32_t i = 0; i < cnt; ++i) {
                    if (o >= out_len) return MDZ_ST_OUTPUT_OVERRUN;
                    out[o++] = (uint8_t)v;
                }
            }
        }
    }
    return MDZ_ST_OK;
}

/* The same lane container as mdz_ref_lzss_decode, with the checks above. */
int mdz_ref_brle_trans_decode(const uint8_t* payload, uint64_t payload_bytes, const uint64_t* blk_off,
                              const uint64_t* col_len, const uint8_t* col_map, uint64_t n_chunks,
                              uint64_t chunk, uint8_t* out, uint64_t out_bytes, int32_t* status,
                              int threads) {
    if (chunk == 0 || chunk % 128) return MDZ_ST_BAD_HEADER;
    dec_ctx x = {payload, blk_off, col_len, col_map, chunk, payload_bytes, out, out_bytes, status,
                 mdz_ref_brle_trans_lane_decode};
    mdz_parallel_for(n_chunks, threads, dec_task, &x);
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (status[c] != MDZ_ST_OK) return status[c];
    return 0;
}

/* Plain (per-chunk) decompress_func (lzss_gpu2.h:718-792): chunk c's stream
 * is data[lens[c-1], lens[c]) and decodes to out[c*chunk, ...) (the last
 * chunk holds what is left of out_bytes), with mdz_ref_lzss_lane_decode's
 * bounds and statuses.  The reference strides its output by the compile-time
 * CHUNK_SIZE and stops when its input is used up. */
typedef struct {
    const uint8_t* data; const uint64_t* lens; uint64_t chunk; uint8_t* out; uint64_t out_bytes;
    int32_t* status;
} plain_dec_ctx;

static void plain_dec_task(void* vctx, uint64_t c) {
    plain_dec_ctx* x = (plain_dec_ctx*)vctx;
    const uint64_t b = c ? x->lens[c - 1] : 0, e = x->lens[c];
    const uint64_t o = c * x->chunk;
    const uint64_t n = o >= x->out_bytes ? 0 : (x->out_bytes - o < x->chunk ? x->out_bytes - o : x->chunk);
    x->status[c] = e < b ? MDZ_ST_BAD_HEADER : mdz_ref_lzss_lane_decode(x->data + b, e - b, x->out + o, n);
}

int mdz_ref_lzss_plain_decode(const uint8_t* data, const uint64_t* lens, uint64_t n_chunks, uint64_t chunk,
                              uint8_t* out, uint64_t out_bytes, int32_t* status, int threads) {
    plain_dec_ctx x = {data, lens, chunk, out, out_bytes, status};
    mdz_parallel_for(n_chunks, threads, plain_dec_task, &x);
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (status[c] != MDZ_ST_OK) return 1;
    return 0;
}
