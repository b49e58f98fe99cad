/*
 * Deflate / zlib inflate -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * zlib itself is the oracle: RFC 1950/1951 fix every stream's output and the
 * system libz decides it.  The reference path this stands beside is
 * deflate::decompress_gpu (include/zlib/zlib.h:966-1082) and cuDF's
 * gpuinflate (include/cudf_deflate.cuh:1229-1345,1399-1407): per-stream
 * {bytes_written, status} with cuDF's status convention.
 *
 * The 2-byte zlib header is checked here (CM = 8, CINFO <= 7, no preset
 * dictionary, FCHECK) and the deflate data is then inflated raw, so the
 * Adler-32 trailer is not verified -- as cuDF's parse_hdr path does not
 * (cudf_deflate.cuh:1180-1227).
 */
#include "mdz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

/* zlib's error message -> status (inflate.c / inftrees.c messages). */
static int status_of_msg(const char* m) {
    static const struct { const char* msg; int st; } tab[] = {
        {"invalid block type", MDZ_ST_BAD_BLOCK_TYPE},
        {"invalid stored block lengths", MDZ_ST_STORED_LEN},
        {"too many length or distance symbols", MDZ_ST_BAD_COUNTS},
        {"invalid code lengths set", MDZ_ST_BAD_CODE_LENS},
        {"invalid bit length repeat", MDZ_ST_BAD_REPEAT},
        {"invalid code -- missing end-of-block", MDZ_ST_NO_EOB},
        {"invalid literal/lengths set", MDZ_ST_BAD_LITLEN},
        {"invalid distances set", MDZ_ST_BAD_DIST},
        {"invalid literal/length code", MDZ_ST_BAD_SYMBOL},
        {"invalid distance code", MDZ_ST_BAD_DISTANCE},
        {"invalid distance too far back", MDZ_ST_BAD_DISTANCE},
    };
    if (m)
        for (size_t i = 0; i < sizeof tab / sizeof tab[0]; ++i)
            if (!strcmp(m, tab[i].msg)) return tab[i].st;
    return MDZ_ST_BAD_SYMBOL;
}

int mdz_ref_inflate(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap, uint64_t* written,
                    int zlib_hdr) {
    *written = 0;
    if (zlib_hdr) {
        if (n < 2) return MDZ_ST_INPUT_OVERRUN;
        const unsigned cmf = src[0], flg = src[1];
        if ((cmf & 15) != 8 || (cmf >> 4) > 7 || (flg & 0x20) || ((cmf << 8) | flg) % 31)
            return MDZ_ST_BAD_HEADER;
        src += 2;
        n -= 2;
    }
    z_stream z;
    memset(&z, 0, sizeof z);
    if (inflateInit2(&z, -15) != Z_OK) return MDZ_ST_BAD_HEADER;
    uint8_t dummy = 0;
    z.next_in = (Bytef*)(n ? src : &dummy);
    z.avail_in = 0;
    z.next_out = cap ? dst : &dummy;
    z.avail_out = 0;
    uint64_t in_left = n, out_left = cap;
    int st = MDZ_ST_OK;
    for (;;) {
        if (z.avail_in == 0 && in_left) {
            const uInt k = in_left > (1u << 30) ? (1u << 30) : (uInt)in_left;
            z.avail_in = k;
            in_left -= k;
        }
        if (z.avail_out == 0 && out_left) {
            const uInt k = out_left > (1u << 30) ? (1u << 30) : (uInt)out_left;
            z.avail_out = k;
            out_left -= k;
        }
        const int r = inflate(&z, Z_NO_FLUSH);
        if (r == Z_STREAM_END) break;
        if (r == Z_DATA_ERROR) { st = status_of_msg(z.msg); break; }
        if (r == Z_MEM_ERROR) { st = MDZ_ST_BAD_HEADER; break; }
        if (z.avail_out == 0 && out_left == 0) {
            /* the output is full: the stream must end (or fail) without
             * producing another byte, else it overruns */
            uint8_t probe;
            for (;;) {
                if (z.avail_in == 0 && in_left) {
                    const uInt k = in_left > (1u << 30) ? (1u << 30) : (uInt)in_left;
                    z.avail_in = k;
                    in_left -= k;
                }
                z.next_out = &probe;
                z.avail_out = 1;
                const int r2 = inflate(&z, Z_NO_FLUSH);
                if (z.avail_out == 0) { st = MDZ_ST_OUTPUT_OVERRUN; z.total_out -= 1; break; }
                if (r2 == Z_STREAM_END) break;
                if (r2 == Z_DATA_ERROR) { st = status_of_msg(z.msg); break; }
                if (z.avail_in == 0 && in_left == 0) { st = MDZ_ST_INPUT_OVERRUN; break; }
            }
            break;
        }
        if (z.avail_in == 0 && in_left == 0 && (r == Z_BUF_ERROR || r == Z_OK)) {
            st = MDZ_ST_INPUT_OVERRUN;
            break;
        }
    }
    *written = z.total_out;
    inflateEnd(&z);
    return st;
}

typedef struct {
    const uint8_t* src; const uint64_t* src_off; const uint64_t* src_len;
    uint8_t* dst; const uint64_t* dst_off; const uint64_t* dst_cap;
    uint64_t* written; int32_t* status; int zlib_hdr;
} inf_ctx;

static void inf_task(void* v, uint64_t i) {
    inf_ctx* x = (inf_ctx*)v;
    x->status[i] = mdz_ref_inflate(x->src + x->src_off[i], x->src_len[i], x->dst + x->dst_off[i],
                                   x->dst_cap[i], &x->written[i], x->zlib_hdr);
}

int mdz_ref_inflate_batch(const uint8_t* src, const uint64_t* src_off, const uint64_t* src_len,
                          uint8_t* dst, const uint64_t* dst_off, const uint64_t* dst_cap,
                          uint64_t count, int zlib_hdr, uint64_t* written, int32_t* status,
                          int threads) {
    inf_ctx x = {src, src_off, src_len, dst, dst_off, dst_cap, written, status, zlib_hdr};
    mdz_parallel_for(count, threads, inf_task, &x);
    for (uint64_t i = 0; i < count; ++i)
        if (status[i] != MDZ_ST_OK) return status[i];
    return 0;
}

const char* mdz_zlib_version(void) { return zlibVersion(); }
