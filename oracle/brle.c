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
