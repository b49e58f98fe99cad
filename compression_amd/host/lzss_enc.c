/* LZSS lane-transposed encoder (host side): find_match (include/lzss/
 * lzss_gpu2.h:70-109), the lane encoder of compress_init_func /
 * compress_func_new (:332-494,511-707), the lane sort (:1063-1085) and the
 * interleave (:497-509); see the decoder for the token grammar. */
#include "mdz_host.h"

#include <stdlib.h>
#include <string.h>

enum { HIST = 2048, MIN_LEN = 3, MAX_LEN = 34, LEN_BITS = 5 };

/* find_match's choice (lzss_gpu2.h:70-109: offsets 1..count-1 oldest first,
 * the oldest byte never tried, no match past the history or the lane end,
 * the first longest wins, 34 stops the search, < 3 is a literal), found
 * through hash chains: candidates share the first 3 bytes (newest first).  Length of candidate q is the common prefix capped
 * by cur - q (no overlap), the lane end and 34; the longest wins, ties to the
 * oldest -- the same choice as the oldest-first scan. */
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
 * written by the reference -- here it is 0.  Returns the stream length;
 * writes at most cap bytes. */
uint64_t mdz_lzss_lane_encode(const uint8_t* s, uint64_t n, uint8_t* out, uint64_t cap) {
    matcher* m = (matcher*)malloc(sizeof(matcher));
    memset(m->head, 0xff, sizeof(m->head));
    m->prev = (int32_t*)malloc((n + 1) * sizeof(int32_t));
    uint64_t o = 1, flag_at = 0;
    uint32_t ntok = 0;
    uint8_t flags = 0;
    if (cap) out[0] = 0;
    for (uint64_t cur = 0; cur < n;) {
        uint32_t dist = 0;
        const uint32_t len = find_match_fast(m, s, n, cur, &dist);
        if (len == 0) {
            flags |= (uint8_t)(1u << ntok);
            if (o < cap) out[o] = s[cur];
            o += 1;
            matcher_insert(m, s, n, cur);
            cur += 1;
        } else {
            const uint32_t v = (dist << LEN_BITS) | (len - MIN_LEN);
            if (o < cap) out[o] = (uint8_t)v;
            if (o + 1 < cap) out[o + 1] = (uint8_t)(v >> 8);
            o += 2;
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
    free(m->prev);
    free(m);
    return o;
}

/* ------------------------------------------------------------- container -- */

typedef uint64_t (*lane_encoder)(const uint8_t* lane, uint64_t n, uint8_t* out, uint64_t cap);

typedef struct {
    const uint8_t* in;
    uint64_t chunk;
    lane_encoder enc;
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
        uint8_t* s = (uint8_t*)calloc(cap, 1);
        x->raw_len[32 * c + (uint64_t)t] = x->enc(lane, lb, s, cap);
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

/* The lane container both lane-transposed flag-group codecs share
 * (lzss_gpu2.h:332-509,1063-1085; brle_trans.h:204-224,238-398): each lane encoded
 * alone, lanes sorted, rows of 4-byte words interleaved in sorted order,
 * chunk extents rounded up to 128 bytes. */
static int lanes_encode(lane_encoder enc, const uint8_t* in, uint64_t in_bytes, uint64_t chunk,
                        uint8_t** payload, uint64_t* payload_bytes, uint64_t** col_len, uint64_t** blk_off,
                        uint8_t** col_map, uint64_t* n_chunks, int threads) {
    if (chunk == 0 || chunk % 128) return -1;
    const uint64_t n = in_bytes / chunk;               /* trailing partial chunk dropped */
    enc_ctx x = {in, chunk, enc, (uint8_t**)calloc(32 * n + 1, sizeof(uint8_t*)),
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

int mdz_lzss_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** payload,
                    uint64_t* payload_bytes, uint64_t** col_len, uint64_t** blk_off,
                    uint8_t** col_map, uint64_t* n_chunks, int threads) {
    return lanes_encode(mdz_lzss_lane_encode, in, in_bytes, chunk, payload, payload_bytes, col_len, blk_off,
                        col_map, n_chunks, threads);
}

/* ------------------------------------------- byte RLE, lane-transposed -- */

/* brle_trans::compress_func (include/brle/brle_trans.h:401-562) for one lane,
 * with compress_init_func's sizes (:238-379).  A run of the same byte
 * counted in a uint8_t: at 3..255 it becomes a run token (count byte, value
 * byte; flag bit 0), shorter ones become literals (flag bit 1).  The counter
 * wraps at 256, so 256 bytes of a longer run vanish per wrap (the reference
 * encoder's own loss, kept for byte identity).  A token that completes a
 * group of 8 reserves the next flag byte, except the lane's final run,
 * which never does.  A flag byte that no token rewrites stays 0 here (the
 * reference leaves that device byte as it was). */
uint64_t mdz_brle_trans_lane_encode(const uint8_t* s, uint64_t n, uint8_t* out, uint64_t cap) {
    uint64_t o = 1, hpos = 0;
    uint8_t hdr = 0, counter = 0, prev = 0;
    uint32_t blocks = 0;
    if (cap < n + n / 8 + 2) return 0;
    out[0] = 0;
#define MDZ_BRLE_T_NEXT_TOKEN()                                    \
    do {                                                           \
        if (++blocks == 8) {                                       \
            out[hpos] = hdr;                                       \
            hdr = 0;                                               \
            hpos = o++;                                            \
            out[hpos] = 0;                                         \
            blocks = 0;                                            \
        }                                                          \
        if (blocks != 0) out[hpos] = hdr;                          \
    } while (0)
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t v = s[i];
        if (counter == 0) prev = v;
        if (prev == v) {
            ++counter;                                /* uint8_t: wraps at 256 */
            continue;
        }
        if (counter >= 3) {
            out[o++] = counter;
            out[o++] = prev;
            MDZ_BRLE_T_NEXT_TOKEN();
        } else {
            for (uint32_t j = 0; j < counter; ++j) {
                hdr = (uint8_t)(hdr | (1u << blocks));
                out[o++] = prev;
                MDZ_BRLE_T_NEXT_TOKEN();
            }
        }
        counter = 1;
        prev = v;
    }
    if (counter >= 3) {                               /* final run: no flag bookkeeping */
        out[o++] = counter;
        out[o++] = prev;
    } else {
        for (uint32_t j = 0; j < counter; ++j) {
            hdr = (uint8_t)(hdr | (1u << blocks));
            out[o++] = prev;
            MDZ_BRLE_T_NEXT_TOKEN();
        }
    }
#undef MDZ_BRLE_T_NEXT_TOKEN
    return o;
}

int mdz_brle_trans_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** payload,
                          uint64_t* payload_bytes, uint64_t** col_len, uint64_t** blk_off,
                          uint8_t** col_map, uint64_t* n_chunks, int threads) {
    return lanes_encode(mdz_brle_trans_lane_encode, in, in_bytes, chunk, payload, payload_bytes, col_len,
                        blk_off, col_map, n_chunks, threads);
}


/* ------------------------------------------------------ plain format -- */

typedef struct {
    const uint8_t* in;
    uint64_t in_bytes, chunk;
    uint8_t** streams;
    uint64_t* len;
} plain_ctx;

static void plain_task(void* vctx, uint64_t c) {
    plain_ctx* x = (plain_ctx*)vctx;
    const uint64_t beg = c * x->chunk;
    const uint64_t n = x->in_bytes - beg < x->chunk ? x->in_bytes - beg : x->chunk;
    const uint64_t cap = n + n / 8 + 16;
    uint8_t* s = (uint8_t*)malloc(cap);
    x->len[c] = mdz_lzss_lane_encode(x->in + beg, n, s, cap);
    x->streams[c] = s;
}

/* The plain (per-chunk) format of compress_func / kernel_shift_data
 * (lzss_gpu2.h:1450-1566,1926-1948): each chunk of `chunk` input bytes (the
 * last may be shorter) is one token stream with the lane grammar -- flag
 * byte, 8 tokens, the next flag byte reserved after every 8th token -- and
 * the streams are packed back to back; lens[c] is the inclusive prefix sum
 * of their lengths (stream c is data[lens[c-1], lens[c])). */
int mdz_lzss_plain_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** data,
                          uint64_t** lens, uint64_t* n_chunks, int threads) {
    if (chunk == 0) return -1;
    const uint64_t n = (in_bytes + chunk - 1) / chunk;
    plain_ctx x = {in, in_bytes, chunk, (uint8_t**)calloc(n + 1, sizeof(uint8_t*)),
                   (uint64_t*)calloc(n + 1, sizeof(uint64_t))};
    mdz_parallel_for(n, threads, plain_task, &x);
    uint64_t* L = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    uint64_t tot = 0;
    for (uint64_t c = 0; c < n; ++c) L[c] = (tot += x.len[c]);
    uint8_t* d = (uint8_t*)malloc(tot + 1);
    for (uint64_t c = 0; c < n; ++c) {
        memcpy(d + (c ? L[c - 1] : 0), x.streams[c], x.len[c]);
        free(x.streams[c]);
    }
    free(x.streams);
    free(x.len);
    *data = d;
    *lens = L;
    *n_chunks = n;
    return 0;
}
