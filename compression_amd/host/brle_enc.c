/* Byte RLE encoder (host side): brle::encode (include/brle/brle.h:129-230)
 * and compress_gpu (:251-303). */
#include "mdz_host.h"

#include <stdlib.h>
#include <string.h>

#define BRLE_MIN_REPEAT 3
#define BRLE_MAX_REPEAT 130
#define BRLE_MAX_LITERAL 128

/* Restates brle::encode (brle.h:129-230, the non-lookahead branch that the
 * USE_LOOKAHEADd spelling at :23 leaves live).  `lits` holds the pending run
 * or literal group, `tail` the length of the equal-byte suffix. */
uint64_t mdz_brle_chunk_encode(const uint8_t* in, uint64_t n, uint8_t* out) {
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

typedef struct {
    const uint8_t* in; uint64_t in_bytes, chunk;
    uint8_t** bufs; uint64_t* lens;
} benc_ctx;

static void benc_task(void* v, uint64_t c) {
    benc_ctx* x = (benc_ctx*)v;
    uint64_t b = c * x->chunk, n = x->in_bytes - b < x->chunk ? x->in_bytes - b : x->chunk;
    /* worst case: one header per 128 literals (OUTPUT_CHUNK_SIZE_, brle.h:13) */
    uint8_t* o = (uint8_t*)malloc(n + (n + BRLE_MAX_LITERAL - 1) / BRLE_MAX_LITERAL + 1);
    x->lens[c] = mdz_brle_chunk_encode(x->in + b, n, o);
    x->bufs[c] = o;
}

/* brle::compress_gpu (brle.h:251-303) minus the container header: returns the
 * concatenated chunk data and ptr[n+1] (ptr[0] = 0, inclusive scan). */
int mdz_brle_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk,
                        uint8_t** data, uint64_t** ptr, uint64_t* n_chunks, int threads) {
    if (in_bytes == 0 || chunk == 0) return -1;
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

