/* mdz oracle: ORC-style byte RLE (include/brle/brle.h).  TEST INFRASTRUCTURE
 * ONLY.
 *
 * Control byte c: c <= 127 -> run of c+3 copies of the next byte; c >= 128 ->
 * c-127 literal bytes follow (decode :76-97, encode :129-230).  Chunks are
 * independent; chunk i decodes to output offset i*chunk. */
#include "mdz_oracle.h"

#include <stdlib.h>
#include <string.h>

#define BRLE_MIN_REPEAT 3
#define BRLE_MAX_REPEAT 130
#define BRLE_MAX_LITERAL 128

/* Restates brle::encode (brle.h:129-230, the non-lookahead branch that the
 * USE_LOOKAHEADd spelling at :23 leaves live).  `lits` holds the pending run
 * or literal group, `tail` the length of the equal-byte suffix. */
uint64_t mdz_ref_brle_chunk_encode(const uint8_t* in, uint64_t n, uint8_t* out) {
    uint8_t lits[BRLE_MAX_LITERAL];
    uint32_t nl = 0, tail = 0;
    int rep = 0;
    uint64_t pos = 0;
#define FLUSH()                                                       \
    do {                                                              \
        if (nl) {                                                     \
            if (rep) {                                                \
                out[pos++] = (uint8_t)(nl - BRLE_MIN_REPEAT);         \
                out[pos++] = lits[0];                                 \
            } else {                                                  \
                out[pos++] = (uint8_t)(nl + 127);                     \
                memcpy(out + pos, lits, nl);                          \
                pos += nl;                                            \
            }                                                         \
            rep = 0; tail = 0; nl = 0;                                \
        }                                                             \
    } while (0)
    for (uint64_t i = 0; i < n; ++i) {
        uint8_t c = in[i];
        if (nl == 0) {
            lits[nl++] = c; tail = 1;
        } else if (rep) {
            if (c == lits[0]) {
                if (++nl == BRLE_MAX_REPEAT) FLUSH();
            } else {
                FLUSH();
                lits[nl++] = c; tail = 1;
            }
        } else {
            tail = (c == lits[nl - 1]) ? tail + 1 : 1;
            if (tail == BRLE_MIN_REPEAT) {
                if (nl + 1 == BRLE_MIN_REPEAT) {
                    nl += 1; rep = 1;
                } else {
                    nl -= BRLE_MIN_REPEAT - 1;
                    FLUSH();
                    lits[0] = c; nl = BRLE_MIN_REPEAT; rep = 1;
                }
            } else {
                lits[nl++] = c;
                if (nl == BRLE_MAX_LITERAL) FLUSH();
            }
        }
    }
    FLUSH();
#undef FLUSH
    return pos;
}

/* brle::decode (brle.h:76-97) with bounds checks.  A token that does not
 * fit the output writes the bytes that do and reports OUTPUT_OVERRUN. */
int mdz_ref_brle_chunk_decode(const uint8_t* in, uint64_t in_len, uint8_t* out,
                              uint64_t out_cap, uint64_t* out_len) {
    uint64_t p = 0, o = 0;
    while (p < in_len) {
        uint8_t c = in[p++];
        if (c > 127) {
            uint64_t cnt = (uint64_t)c - 127;
            if (p + cnt > in_len) { *out_len = o; return MDZ_ST_INPUT_OVERRUN; }
            if (o + cnt > out_cap) {
                memcpy(out + o, in + p, out_cap - o);
                *out_len = out_cap;
                return MDZ_ST_OUTPUT_OVERRUN;
            }
            memcpy(out + o, in + p, cnt);
            p += cnt; o += cnt;
        } else {
            uint64_t cnt = (uint64_t)c + BRLE_MIN_REPEAT;
            if (p >= in_len) { *out_len = o; return MDZ_ST_INPUT_OVERRUN; }
            if (o + cnt > out_cap) {
                memset(out + o, in[p], out_cap - o);
                *out_len = out_cap;
                return MDZ_ST_OUTPUT_OVERRUN;
            }
            memset(out + o, in[p++], cnt);
            o += cnt;
        }
    }
    *out_len = o;
    return MDZ_ST_OK;
}

typedef struct {
    const uint8_t* in; uint64_t in_bytes, chunk;
    uint8_t** bufs; uint64_t* lens;
} benc_ctx;

static void benc_task(void* v, uint64_t c) {
    benc_ctx* x = (benc_ctx*)v;
    uint64_t b = c * x->chunk, n = x->in_bytes - b < x->chunk ? x->in_bytes - b : x->chunk;
    /* worst case: one header per 128 literals (OUTPUT_CHUNK_SIZE_, brle.h:13) */
    uint8_t* o = (uint8_t*)malloc(n + (n + BRLE_MAX_LITERAL - 1) / BRLE_MAX_LITERAL + 1);
    x->lens[c] = mdz_ref_brle_chunk_encode(x->in + b, n, o);
    x->bufs[c] = o;
}

/* brle::compress_gpu (brle.h:251-303) minus the container header: returns the
 * concatenated chunk data and ptr[n+1] (ptr[0] = 0, inclusive scan). */
int mdz_ref_brle_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk,
                        uint8_t** data, uint64_t** ptr, uint64_t* n_chunks, int threads) {
    if (in_bytes == 0 || chunk == 0) return MDZ_ST_BAD_HEADER;
    uint64_t n = (in_bytes - 1) / chunk + 1;
    benc_ctx x = {in, in_bytes, chunk, (uint8_t**)calloc(n, sizeof(uint8_t*)),
                  (uint64_t*)malloc(n * 8)};
    mdz_parallel_for(n, threads, benc_task, &x);
    uint64_t* p = (uint64_t*)malloc((n + 1) * 8);
    p[0] = 0;
    for (uint64_t c = 0; c < n; ++c) p[c + 1] = p[c] + x.lens[c];
    uint8_t* d = (uint8_t*)malloc(p[n] ? p[n] : 1);
    for (uint64_t c = 0; c < n; ++c) {
        memcpy(d + p[c], x.bufs[c], x.lens[c]);
        free(x.bufs[c]);
    }
    free(x.bufs);
    free(x.lens);
    *data = d; *ptr = p; *n_chunks = n;
    return 0;
}

typedef struct {
    const uint8_t* data; const uint64_t* ptr; uint64_t chunk;
    uint8_t* out; uint64_t out_bytes; int32_t* status;
} bdec_ctx;

static void bdec_task(void* v, uint64_t c) {
    bdec_ctx* x = (bdec_ctx*)v;
    uint64_t ob = c * x->chunk;
    uint64_t cap = ob >= x->out_bytes ? 0 : x->out_bytes - ob;
    if (cap > x->chunk) cap = x->chunk;
    uint64_t got = 0;
    x->status[c] = mdz_ref_brle_chunk_decode(x->data + x->ptr[c], x->ptr[c + 1] - x->ptr[c],
                                             x->out + ob, cap, &got);
}

int mdz_ref_brle_decode(const uint8_t* data, const uint64_t* ptr, uint64_t n_chunks,
                        uint64_t chunk, uint8_t* out, uint64_t out_bytes,
                        int32_t* status, int threads) {
    bdec_ctx x = {data, ptr, chunk, out, out_bytes, status};
    mdz_parallel_for(n_chunks, threads, bdec_task, &x);
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (status[c] != MDZ_ST_OK) return status[c];
    return 0;
}
