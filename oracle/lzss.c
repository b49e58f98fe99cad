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
 * the fast matcher below. */
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

/* Fast equivalent: candidates share the first 3 bytes, found through hash
 * chains (newest first).  Length of candidate q is the common prefix capped
 * by cur - q (no overlap), the lane end and 34; the longest wins, ties to the
 * oldest -- the same choice as the oldest-first scan above. */
typedef struct {
    int32_t head[4096];
    int32_t* prev;
} matcher;

static inline uint32_t hash3(const uint8_t* p) {
    const uint32_t v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16);
    return (v * 2654435761u) >> 20;
}

static uint32_t find_match_fast(const matcher* m, const uint8_t* s, uint64_t n, uint64_t cur,
                                uint32_t* dist) {
    *dist = 0;
    if (n - cur < MIN_LEN || cur < 2) return 0;
    const uint64_t count = cur < HIST ? cur : HIST;
    const uint64_t lo = cur - count + 1;            /* oldest position tried */
    const uint64_t look = n - cur;
    uint32_t best = 0;
    uint64_t best_q = 0;
    for (int32_t q = m->head[hash3(s + cur)]; q >= 0 && (uint64_t)q >= lo; q = m->prev[q]) {
        const uint64_t room = cur - (uint64_t)q;
        uint64_t cap = room < look ? room : look;
        if (cap > MAX_LEN) cap = MAX_LEN;
        uint32_t f = 0;
        while (f < cap && s[q + f] == s[cur + f]) ++f;
        if (f >= best) { best = f; best_q = (uint64_t)q; }   /* >=: older wins ties */
    }
    if (best < MIN_LEN) return 0;
    *dist = (uint32_t)(cur - best_q);
    return best;
}

static inline void matcher_insert(matcher* m, const uint8_t* s, uint64_t n, uint64_t p) {
    if (p + MIN_LEN > n) return;
    const uint32_t h = hash3(s + p);
    m->prev[p] = m->head[h];
    m->head[h] = (int32_t)p;
}

/* ------------------------------------------------------------ lane codec -- */

/* One lane -> token stream.  Restates compress_init_func / compress_func_new
 * (lzss_gpu2.h:332-494,511-707): the first flag byte is reserved up front;
 * after every 8th token the next flag byte is reserved at once, so a lane of
 * 8k tokens ends in a reserved byte that is counted in its length and never
 * written by the reference -- here it is 0.  `naive` selects the literal
 * find_match restatement.  Returns the stream length; writes at most cap. */
static uint64_t lane_encode(const uint8_t* s, uint64_t n, uint8_t* out, uint64_t cap, int naive) {
    matcher* m = NULL;
    if (!naive) {
        m = (matcher*)malloc(sizeof(matcher));
        memset(m->head, 0xff, sizeof(m->head));
        m->prev = (int32_t*)malloc((n + 1) * sizeof(int32_t));
    }
    uint64_t o = 1, flag_at = 0;
    uint32_t ntok = 0;
    uint8_t flags = 0;
    if (cap) out[0] = 0;
    for (uint64_t cur = 0; cur < n;) {
        uint32_t dist = 0;
        const uint32_t len = naive ? find_match_naive(s, n, cur, &dist)
                                   : find_match_fast(m, s, n, cur, &dist);
        if (len == 0) {
            flags |= (uint8_t)(1u << ntok);
            if (o < cap) out[o] = s[cur];
            o += 1;
            if (!naive) matcher_insert(m, s, n, cur);
            cur += 1;
        } else {
            const uint32_t v = (dist << LEN_BITS) | (len - MIN_LEN);
            if (o < cap) out[o] = (uint8_t)v;
            if (o + 1 < cap) out[o + 1] = (uint8_t)(v >> 8);
            o += 2;
            if (!naive)
                for (uint32_t i = 0; i < len; ++i) matcher_insert(m, s, n, cur + i);
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
    if (m) { free(m->prev); free(m); }
    return o;
}

uint64_t mdz_ref_lzss_lane_encode(const uint8_t* lane, uint64_t n, uint8_t* out, uint64_t cap) {
    return lane_encode(lane, n, out, cap, 0);
}

uint64_t mdz_ref_lzss_lane_encode_naive(const uint8_t* lane, uint64_t n, uint8_t* out,
                                        uint64_t cap) {
    return lane_encode(lane, n, out, cap, 1);
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

/* ------------------------------------------------------------- container -- */

typedef struct {
    const uint8_t* in;
    uint64_t chunk;
    uint8_t** streams;       /* [32n] lane streams */
    uint64_t* raw_len;       /* [32n] unsorted lane lengths */
} enc_ctx;

static void enc_task(void* vctx, uint64_t c) {
    enc_ctx* x = (enc_ctx*)vctx;
    const uint64_t lb = x->chunk / 32;
    uint8_t* lane = (uint8_t*)malloc(lb);
    const uint64_t cap = lb + lb / 8 + 16;
    for (int t = 0; t < 32; ++t) {
        const uint8_t* src = x->in + c * x->chunk;
        for (uint64_t i = 0; i < lb; ++i) lane[i] = src[128 * (i / 4) + 4 * (uint64_t)t + (i & 3)];
        uint8_t* s = (uint8_t*)malloc(cap);
        x->raw_len[32 * c + (uint64_t)t] = mdz_ref_lzss_lane_encode(lane, lb, s, cap);
        x->streams[32 * c + (uint64_t)t] = s;
    }
    free(lane);
}

/* Sorted position s -> original lane: descending length, ties to the higher
 * lane id (a stable ascending sort by length, reversed; lzss_gpu2.h:1063-1085). */
static void sort_lanes(const uint64_t* len, uint8_t* map) {
    for (int s = 0; s < 32; ++s) map[s] = (uint8_t)(31 - s);
    for (int i = 1; i < 32; ++i) {          /* stable insertion sort, descending */
        const uint8_t v = map[i];
        int j = i - 1;
        while (j >= 0 && len[map[j]] < len[v]) { map[j + 1] = map[j]; --j; }
        map[j + 1] = v;
    }
}

int mdz_ref_lzss_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** payload,
                        uint64_t* payload_bytes, uint64_t** col_len, uint64_t** blk_off,
                        uint8_t** col_map, uint64_t* n_chunks, int threads) {
    if (chunk == 0 || chunk % 128) return MDZ_ST_BAD_HEADER;
    const uint64_t n = in_bytes / chunk;               /* trailing partial chunk dropped */
    enc_ctx x = {in, chunk, (uint8_t**)calloc(32 * n + 1, sizeof(uint8_t*)),
                 (uint64_t*)calloc(32 * n + 1, sizeof(uint64_t))};
    mdz_parallel_for(n, threads, enc_task, &x);
    uint64_t* cl = (uint64_t*)malloc((32 * n + 1) * sizeof(uint64_t));
    uint8_t* cm = (uint8_t*)malloc(32 * n + 1);
    uint64_t* off = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    off[0] = 0;
    for (uint64_t c = 0; c < n; ++c) {
        sort_lanes(x.raw_len + 32 * c, cm + 32 * c);
        uint64_t words = 0;
        for (int s = 0; s < 32; ++s) {
            cl[32 * c + s] = x.raw_len[32 * c + cm[32 * c + s]];
            words += (cl[32 * c + s] + 3) / 4;
        }
        off[c + 1] = off[c] + ((4 * words + 127) / 128) * 128;
    }
    uint8_t* pay = (uint8_t*)calloc(off[n] + 1, 1);
    for (uint64_t c = 0; c < n; ++c) {
        uint8_t* dst = pay + off[c];
        const uint64_t* L = cl + 32 * c;
        for (uint64_t k = 0;; ++k) {                   /* row k: word k of each lane with > k words */
            int any = 0;
            for (int s = 0; s < 32; ++s) {
                if ((L[s] + 3) / 4 <= k) break;        /* sorted: the rest are shorter */
                const uint8_t* src = x.streams[32 * c + cm[32 * c + s]];
                for (int j = 0; j < 4; ++j) dst[j] = 4 * k + j < L[s] ? src[4 * k + j] : 0;
                dst += 4;
                any = 1;
            }
            if (!any) break;
        }
    }
    for (uint64_t i = 0; i < 32 * n; ++i) free(x.streams[i]);
    free(x.streams);
    free(x.raw_len);
    *payload = pay;
    *payload_bytes = off[n];
    *col_len = cl;
    *blk_off = off;
    *col_map = cm;
    *n_chunks = n;
    return 0;
}

typedef struct {
    const uint8_t* pay; const uint64_t* off; const uint64_t* cl; const uint8_t* cm;
    uint64_t chunk; uint64_t pay_bytes; uint8_t* out; uint64_t out_bytes; int32_t* status;
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
        const int r = mdz_ref_lzss_lane_decode(lanes + base[s], L[s], vals, lb);
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
    dec_ctx x = {payload, blk_off, col_len, col_map, chunk, payload_bytes, out, out_bytes, status};
    mdz_parallel_for(n_chunks, threads, dec_task, &x);
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (status[c] != MDZ_ST_OK) return status[c];
    return 0;
}

/* ------------------------------------------------------------- generator -- */

static double std_normal(mdz_rng* r) {
    double u1 = mdz_rng_unit(r), u2 = mdz_rng_unit(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Gamma(a) for a < 1: Marsaglia-Tsang for Gamma(a+1), times U^(1/a). */
static double gamma_small(mdz_rng* r, double a) {
    const double d = a + 1.0 - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    double g;
    for (;;) {
        const double z = std_normal(r);
        double v = 1.0 + c * z;
        if (v <= 0) continue;
        v = v * v * v;
        const double u = mdz_rng_unit(r);
        if (u < 1.0 - 0.0331 * z * z * z * z || log(u) < 0.5 * z * z + d * (1.0 - v + log(v))) {
            g = d * v;
            break;
        }
    }
    double u = mdz_rng_unit(r);
    if (u < 1e-300) u = 1e-300;
    return g * pow(u, 1.0 / a);
}

/* 64-symbol order-2 Markov text with rows ~ Dirichlet(0.3), mixed with
 * copies: at each step, with probability copy_p (once there is history) copy
 * L ~ U{4..256} bytes from distance D ~ U{1..min(4096, pos)}, byte by byte
 * (overlap allowed); otherwise emit one Markov symbol. */
enum { SYM = 64, ROWS = SYM * SYM };

double* mdz_markov_table(uint64_t seed) {
    mdz_rng r;
    mdz_rng_seed(&r, seed);
    double* cdf = (double*)malloc(sizeof(double) * ROWS * SYM);
    for (int row = 0; row < ROWS; ++row) {
        double* c = cdf + row * SYM, tot = 0;
        for (int k = 0; k < SYM; ++k) { c[k] = gamma_small(&r, 0.3); tot += c[k]; }
        double acc = 0;
        for (int k = 0; k < SYM; ++k) { acc += c[k] / tot; c[k] = acc; }
        c[SYM - 1] = 2.0;                            /* the last symbol catches rounding */
    }
    return cdf;
}

void mdz_markov_gen(const double* cdf, uint8_t* out, uint64_t n, uint64_t seed, double copy_p) {
    mdz_rng r;
    mdz_rng_seed(&r, seed);
    uint64_t pos = 0;
    while (pos < n) {
        if (pos > 0 && mdz_rng_unit(&r) < copy_p) {
            uint64_t len = 4 + mdz_rng_below(&r, 253);
            const uint64_t dmax = pos < 4096 ? pos : 4096;
            const uint64_t d = 1 + mdz_rng_below(&r, dmax);
            if (len > n - pos) len = n - pos;
            for (uint64_t i = 0; i < len; ++i, ++pos) out[pos] = out[pos - d];
            continue;
        }
        const uint32_t a = pos >= 2 ? out[pos - 2] : 0, b = pos >= 1 ? out[pos - 1] : 0;
        const double* c = cdf + (a * SYM + b) * SYM;
        const double u = mdz_rng_unit(&r);
        int lo = 0, hi = SYM - 1;                    /* first k with c[k] > u */
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (c[mid] > u) hi = mid; else lo = mid + 1;
        }
        out[pos++] = (uint8_t)lo;
    }
}

/* Config 3: the Markov text with copy probability 0.4; the table and the
 * sequence use seeds derived from `seed`. */
void mdz_gen_lzss_markov(uint8_t* out, uint64_t n, uint64_t seed) {
    double* cdf = mdz_markov_table(seed);
    mdz_markov_gen(cdf, out, n, seed ^ 0x9e3779b97f4a7c15ull, 0.4);
    free(cdf);
}
