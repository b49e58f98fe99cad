/*
 * mdz oracle -- CPU restatement of the reference decoders.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP decoders
 * and the timed CPU baseline in bench.py.  It is never linked into, loaded by
 * or called from the product libraries (compression_amd/libmdz.so,
 * libmdz_host.so); the product fails loudly when its HIP build is missing
 * instead of falling back here.  It uses the host library's thread pool.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the ZaidQureshi/compression tree).
 */
#ifndef MDZ_ORACLE_H
#define MDZ_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#include "mdz_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status --
 * Per-chunk / per-stream status codes.  They follow the cuDF inflate
 * convention (include/cudf_deflate.cuh:440,527-574,1051,1273-1298) and are
 * shared with the HIP library (include/mdz.h). */
enum {
    MDZ_ST_OK = 0,
    MDZ_ST_OUTPUT_OVERRUN = 1,   /* decoded data would exceed the output slot */
    MDZ_ST_INPUT_OVERRUN = 2,    /* stream ended before the output was complete */
    MDZ_ST_BAD_BLOCK_TYPE = -1,  /* Deflate BTYPE == 3 / unknown RLE v2 header */
    MDZ_ST_STORED_LEN = -2,      /* Deflate stored LEN != ~NLEN */
    MDZ_ST_BAD_CODE_LENS = -3,   /* Deflate: code-length code set incomplete/over-subscribed */
    MDZ_ST_BAD_REPEAT = -4,      /* Deflate: repeat with no previous length / too many lengths */
    MDZ_ST_BAD_LITLEN = -5,      /* Deflate: literal/length code set invalid */
    MDZ_ST_BAD_DIST = -6,        /* Deflate: distance code set invalid */
    MDZ_ST_BAD_COUNTS = -7,      /* Deflate: HLIT/HDIST out of range */
    MDZ_ST_NO_EOB = -8,          /* Deflate: missing end-of-block code */
    MDZ_ST_BAD_DISTANCE = -9,    /* back-reference before the start of output / bad dist symbol */
    MDZ_ST_BAD_SYMBOL = -10,     /* undecodable symbol / invalid token */
    MDZ_ST_BAD_HEADER = -11,     /* zlib header check failed / bad container metadata */
    MDZ_ST_BAD_CHECKSUM = -13    /* Deflate trailer check: Adler-32 / CRC-32 / ISIZE differs */
};

/* ============================================================ RLE v1 ===== *
 * decoder_warp (include/rle_v1/decoder.h:275-423) with bounds checks; the
 * container decode reads the 4-byte compacted round-robin interleave
 * (comp_read_data :72-95) and writes lane t's lines t, t+32, ... */
int mdz_ref_rlev1_lane_decode(const uint8_t* in, uint64_t in_len, int elem_bytes,
                              void* vals, uint64_t n_vals);
int mdz_ref_rlev1_decode(const uint8_t* payload, const uint64_t* blk_off,
                         const uint64_t* col_len, uint64_t n_chunks, uint64_t chunk,
                         int elem_bytes, uint32_t line, uint8_t* out, uint64_t out_bytes,
                         int32_t* status, int threads);

/* ============================================================ RLE v2 ===== *
 * decoder_warp (include/rlev2/decoder_new.h:354-611), 64-bit patch mask. */
int mdz_ref_rlev2_lane_decode(const uint8_t* in, uint64_t in_len, int elem_bytes, uint64_t* vals,
                              uint64_t n_vals);
int mdz_ref_rlev2_decode(const uint8_t* payload, const uint64_t* blk_off, const uint64_t* col_len,
                         uint64_t n_chunks, uint64_t chunk, int elem_bytes, uint32_t read_unit,
                         uint8_t* out, uint64_t out_bytes, int32_t* status, int threads);

/* Plain RLE v2 (include/rlev2/decoder.h:190-262): one chunk's run stream
 * into at most cap int64 values (*n_out decoded), and a whole container
 * [u32 n+1][u64 ptr[n+1]][u64 raw_bytes][streams] into out. */
int mdz_ref_rlev2_plain_chunk_decode(const uint8_t* in, uint64_t in_len, uint64_t* vals, uint64_t cap,
                                     uint64_t* n_out);
int mdz_ref_rlev2_plain_decode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t* out,
                               uint64_t out_bytes, int32_t* status, uint64_t n_status, int threads);

/* ========================================================== byte RLE ===== *
 * brle::decode (include/brle/brle.h:76-97) with bounds checks. */
int mdz_ref_brle_chunk_decode(const uint8_t* in, uint64_t in_len, uint8_t* out,
                              uint64_t out_cap, uint64_t* out_len);
int mdz_ref_brle_decode(const uint8_t* data, const uint64_t* ptr, uint64_t n_chunks,
                        uint64_t chunk, uint8_t* out, uint64_t out_bytes,
                        int32_t* status, int threads);

/* ============================================================== LZSS ===== *
 * Transposed decompress_func (include/lzss/lzss_gpu2.h:114-325) with
 * bounds checks; the literal find_match scan (:70-109) drives a second lane
 * encoder that pins the host encoder's matcher. */
uint64_t mdz_ref_lzss_lane_encode_naive(const uint8_t* lane, uint64_t n, uint8_t* out,
                                        uint64_t cap);
int mdz_ref_lzss_lane_decode(const uint8_t* in, uint64_t in_len, uint8_t* out, uint64_t out_len);
int mdz_ref_lzss_decode(const uint8_t* payload, uint64_t payload_bytes, const uint64_t* blk_off,
                        const uint64_t* col_len, const uint8_t* col_map, uint64_t n_chunks,
                        uint64_t chunk, uint8_t* out, uint64_t out_bytes, int32_t* status,
                        int threads);
/* Lane-transposed byte RLE: brle_trans::decompress_func
 * (include/brle/brle_trans.h:72-202) per lane, over the LZSS lane container. */
int mdz_ref_brle_trans_lane_decode(const uint8_t* in, uint64_t in_len, uint8_t* out, uint64_t out_len);
int mdz_ref_brle_trans_decode(const uint8_t* payload, uint64_t payload_bytes, const uint64_t* blk_off,
                              const uint64_t* col_len, const uint8_t* col_map, uint64_t n_chunks,
                              uint64_t chunk, uint8_t* out, uint64_t out_bytes, int32_t* status,
                              int threads);
/* Plain (per-chunk) decompress_func (lzss_gpu2.h:718-792) over packed
 * streams with inclusive-prefix lens[]; per-chunk status as the lane
 * decoder's. */
int mdz_ref_lzss_plain_decode(const uint8_t* data, const uint64_t* lens, uint64_t n_chunks, uint64_t chunk,
                              uint8_t* out, uint64_t out_bytes, int32_t* status, int threads);

/* =========================================================== Deflate ===== *
 * zlib is the oracle (RFC 1950/1951/1952).  hdr = framing | verify:
 *   0 raw Deflate; 1 zlib: check the 2-byte header, then inflate raw;
 *   2 gzip: header length by cuDF's parse_gzip_header rules
 *     (include/cudf_deflate.cuh:1179-1217; FEXTRA per RFC 1952), rejected
 *     header -> -1, the last 8 bytes excluded from the Deflate input
 *     (:1244-1253);
 *   | 0x100 verify: after status 0, compare zlib's adler32() (zlib) or
 *     crc32() and the length (gzip) with the trailer -> -13 on a mismatch, 2
 *     for a cut-off zlib trailer.  Writes
 * at most cap bytes; *written = bytes produced.  Status per cuDF's codes:
 * 1 output full before the end, 2 input ended early, negatives per zlib's
 * error (invalid block type -1, stored lengths -2, code lengths set -3,
 * bit length repeat -4, literal/lengths set -5, distances set -6, too many
 * symbols -7, missing end-of-block -8, invalid distance code / too far back
 * -9, invalid literal/length code -10, header -11). */
int mdz_ref_inflate(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap, uint64_t* written,
                    int hdr);
int mdz_ref_inflate_batch(const uint8_t* src, const uint64_t* src_off, const uint64_t* src_len,
                          uint8_t* dst, const uint64_t* dst_off, const uint64_t* dst_cap,
                          uint64_t count, int hdr, uint64_t* written, int32_t* status,
                          int threads);
/* gzip member header length or -1 (rules above). */
int64_t mdz_ref_gzip_header_len(const uint8_t* src, uint64_t n);
const char* mdz_zlib_version(void);

#ifdef __cplusplus
}
#endif
#endif
