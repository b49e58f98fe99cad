/*
 * Deflate / zlib inflate -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * zlib itself is the oracle: RFC 1950/1951 fix every stream's output and the
 * system libz decides it.  The reference path this stands beside is
 * deflate::decompress_gpu (include/zlib/zlib.h:966-1082) and cuDF's
 * gpuinflate (include/cudf_deflate.cuh:1229-1345,1399-1407): per-stream
 * {bytes_written, status} with cuDF's status convention.
 *
 * Framing: the 2-byte zlib header is checked here (CM = 8, CINFO <= 7, no
 * preset dictionary, FCHECK); a gzip header is measured by cuDF's rules
 * (cudf_deflate.cuh:1179-1217, with RFC 1952's FEXTRA layout) and its 8-byte
 * footer cut off (:1244-1253).  The Deflate data is then inflated raw, and
 * with the verify flag the trailer is compared with zlib's own adler32_z /
 * crc32_z of the output (cuDF verifies neither).
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

/* RFC 1952 member header: 10 fixed bytes, then the optional fields FLG
 * announces.  cuDF's bounds: 18 bytes at least, every field inside the
 * stream, and one Deflate byte at least before the 8-byte footer. */
int64_t mdz_ref_gzip_header_len(const uint8_t* src, uint64_t n) {
    if (n < 18 || src[0] != 0x1f || src[1] != 0x8b || src[2] != 8) return -1;
    const unsigned flg = src[3];
    uint64_t at = 10;
    if (flg & 4) {                                   /* FEXTRA: XLEN (LE16) + XLEN bytes */
        at += 2 + (uint64_t)(src[10] | (src[11] << 8));
        if (at >= n) return -1;
    }
    if (flg & 8) {                                   /* FNAME: zero-terminated */
        while (at < n && src[at]) ++at;
        if (at >= n) return -1;
        ++at;
    }
    if (flg & 16) {                                  /* FCOMMENT: zero-terminated */
        while (at < n && src[at]) ++at;
        if (at >= n) return -1;
        ++at;
    }
    if (flg & 2) at += 2;                            /* FHCRC */
    return at + 8 >= n ? -1 : (int64_t)at;
}

static int inflate_raw(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap,
                       uint64_t* written, uint64_t* consumed);

int mdz_ref_inflate(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap, uint64_t* written,
                    int hdr) {
    *written = 0;
    const uint8_t* const s0 = src;
    const uint64_t n0 = n;
    const int framing = hdr & 3;
    if (framing == 1) {
        if (n < 2) return MDZ_ST_INPUT_OVERRUN;
        const unsigned cmf = src[0], flg = src[1];
        if ((cmf & 15) != 8 || (cmf >> 4) > 7 || (flg & 0x20) || ((cmf << 8) | flg) % 31)
            return MDZ_ST_BAD_HEADER;
        src += 2;
        n -= 2;
    } else if (framing == 2) {
        const int64_t h = mdz_ref_gzip_header_len(src, n);
        if (h < 0) return MDZ_ST_BAD_BLOCK_TYPE;
        src += h;
        n = n - 8 - (uint64_t)h;
    }
    uint64_t consumed = 0;
    int st = inflate_raw(src, n, dst, cap, written, &consumed);
    if (st != MDZ_ST_OK || !(hdr & 0x100) || framing == 0) return st;
    if (framing == 1) {
        const uint64_t tb = 2 + consumed;
        if (tb + 4 > n0) return MDZ_ST_INPUT_OVERRUN;
        const uint32_t want = ((uint32_t)s0[tb] << 24) | ((uint32_t)s0[tb + 1] << 16) |
                              ((uint32_t)s0[tb + 2] << 8) | s0[tb + 3];
        return adler32_z(1, dst, *written) == want ? MDZ_ST_OK : MDZ_ST_BAD_CHECKSUM;
    }
    const uint8_t* f = s0 + n0 - 8;
    const uint32_t crc = f[0] | (f[1] << 8) | (f[2] << 16) | ((uint32_t)f[3] << 24);
    const uint32_t isz = f[4] | (f[5] << 8) | (f[6] << 16) | ((uint32_t)f[7] << 24);
    return crc32_z(0, dst, *written) == crc && isz == (uint32_t)*written ? MDZ_ST_OK : MDZ_ST_BAD_CHECKSUM;
}

/* Raw Deflate into dst[0, cap); *consumed = input bytes up to the byte holding
 * the last block's end. */
static int inflate_raw(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap,
                       uint64_t* written, uint64_t* consumed) {
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
    *consumed = n - in_left - z.avail_in;
    inflateEnd(&z);
    return st;
}

typedef struct {
    const uint8_t* src; const uint64_t* src_off; const uint64_t* src_len;
    uint8_t* dst; const uint64_t* dst_off; const uint64_t* dst_cap;
    uint64_t* written; int32_t* status; int hdr;
} inf_ctx;

static void inf_task(void* v, uint64_t i) {
    inf_ctx* x = (inf_ctx*)v;
    x->status[i] = mdz_ref_inflate(x->src + x->src_off[i], x->src_len[i], x->dst + x->dst_off[i],
                                   x->dst_cap[i], &x->written[i], x->hdr);
}

int mdz_ref_inflate_batch(const uint8_t* src, const uint64_t* src_off, const uint64_t* src_len,
                          uint8_t* dst, const uint64_t* dst_off, const uint64_t* dst_cap,
                          uint64_t count, int hdr, uint64_t* written, int32_t* status,
                          int threads) {
    inf_ctx x = {src, src_off, src_len, dst, dst_off, dst_cap, written, status, hdr};
    mdz_parallel_for(count, threads, inf_task, &x);
    for (uint64_t i = 0; i < count; ++i)
        if (status[i] != MDZ_ST_OK) return status[i];
    return 0;
}

const char* mdz_zlib_version(void) { return zlibVersion(); }
