/*
 * mdz oracle -- CPU restatement of the reference codecs.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP decoders
 * and the timed CPU baseline in bench.py.  It is never linked into, loaded by
 * or called from the product library (compression_amd/libmdz.so); the product
 * fails loudly when its HIP build is missing instead of falling back here.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the ZaidQureshi/compression tree).
 */
#ifndef MDZ_ORACLE_H
#define MDZ_ORACLE_H

#include <stdint.h>
#include <stddef.h>

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
    MDZ_ST_BAD_HEADER = -11      /* zlib header check failed / bad container metadata */
};

/* ------------------------------------------------------------------ PRNG --
 * splitmix64-seeded xoshiro256** (the generator family named for the
 * synthetic workloads). */
typedef struct { uint64_t s[4]; } mdz_rng;
void     mdz_rng_seed(mdz_rng* r, uint64_t seed);
uint64_t mdz_rng_next(mdz_rng* r);
/* uniform integer in [0, n) by 64x64->128 multiply-high (n >= 1) */
uint64_t mdz_rng_below(mdz_rng* r, uint64_t n);
/* uniform double in [0, 1) from the top 53 bits */
double   mdz_rng_unit(mdz_rng* r);

/* --------------------------------------------------------- thread pool --- */
typedef void (*mdz_task_fn)(void* ctx, uint64_t index);
/* Runs fn(ctx, i) for i in [0, n) on `threads` pthreads with a shared atomic
 * cursor (dynamic scheduling).  threads <= 1 runs inline. */
void mdz_parallel_for(uint64_t n, int threads, mdz_task_fn fn, void* ctx);
int  mdz_host_cores(void);      /* sched_getaffinity count */
void mdz_free(void* p);         /* frees buffers the oracle allocated */

/* ============================================================ RLE v1 ===== *
 * Lane-transposed ORC-style RLE v1 (include/rle_v1/).  A chunk of `chunk`
 * bytes is cut into lines of `line` bytes (COMP_COL_LEN); line j belongs to
 * lane j % 32.  Each lane's values are encoded to a byte stream; the 32
 * streams are interleaved in 4-byte words, compacted round-robin in lane
 * order, and the chunk padded to 128 bytes. */

/* Encode one lane's value sequence (n >= 2 values of elem_bytes each,
 * little-endian) -- restates compression_warp, include/rle_v1/encoder.h:607-1018.
 * Returns the stream length; writes at most `cap` bytes to out. */
uint64_t mdz_ref_rlev1_lane_encode(const void* vals, uint64_t n, int elem_bytes,
                                   uint8_t* out, uint64_t cap);

/* Decode one lane stream into `n_vals` values -- restates decoder_warp,
 * include/rle_v1/decoder.h:275-423 (with bounds checks).  Returns status. */
int mdz_ref_rlev1_lane_decode(const uint8_t* in, uint64_t in_len, int elem_bytes,
                              void* vals, uint64_t n_vals);

/* Whole-container encode (include/rle_v1/rlev1.h:24-97): input is zero-padded
 * to n_chunks*chunk.  Allocates *payload (chunk data, no 8-byte size header),
 * *col_len (u64[32n]) and *blk_off (u64[n+1]); free with mdz_free. */
int mdz_ref_rlev1_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk,
                         int elem_bytes, uint32_t line,
                         uint8_t** payload, uint64_t* payload_bytes,
                         uint64_t** col_len, uint64_t** blk_off, uint64_t* n_chunks,
                         int threads);

/* Whole-container decode (include/rle_v1/rlev1.h:99-199 + decoder.h:426-470).
 * Writes min(out_bytes, n_chunks*chunk) bytes; status[n_chunks] per chunk.
 * Returns 0, or the first nonzero chunk status. */
int mdz_ref_rlev1_decode(const uint8_t* payload, const uint64_t* blk_off,
                         const uint64_t* col_len, uint64_t n_chunks, uint64_t chunk,
                         int elem_bytes, uint32_t line, uint8_t* out, uint64_t out_bytes,
                         int32_t* status, int threads);

/* ============================================================ RLE v2 ===== *
 * Lane-transposed ORC RLE v2 (include/rlev2/, theirs side).  Lane t owns
 * chunk elements k*32*RU + t*RU + j (j < RU, RU = READ_UNIT); values are the
 * unsigned eb-byte elements widened to 64 bits.  blk_off[n] holds the
 * uncompressed size, as the reference stores it. */
uint64_t mdz_ref_rlev2_lane_encode(const uint64_t* vals, uint64_t n, uint8_t* out, uint64_t cap);
int mdz_ref_rlev2_lane_decode(const uint8_t* in, uint64_t in_len, int elem_bytes, uint64_t* vals,
                              uint64_t n_vals);
int mdz_ref_rlev2_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, int elem_bytes,
                         uint32_t read_unit, uint8_t** payload, uint64_t* payload_bytes,
                         uint64_t** col_len, uint64_t** blk_off, uint64_t* n_chunks, int threads);
int mdz_ref_rlev2_decode(const uint8_t* payload, const uint64_t* blk_off, const uint64_t* col_len,
                         uint64_t n_chunks, uint64_t chunk, int elem_bytes, uint32_t read_unit,
                         uint8_t* out, uint64_t out_bytes, int32_t* status, int threads);

/* ========================================================== byte RLE ===== *
 * include/brle/brle.h.  Container: [u32 n_ptr=n+1][u64 ptr[n+1]][u64 out_bytes][data]. */
uint64_t mdz_ref_brle_chunk_encode(const uint8_t* in, uint64_t n, uint8_t* out);
int mdz_ref_brle_chunk_decode(const uint8_t* in, uint64_t in_len, uint8_t* out,
                              uint64_t out_cap, uint64_t* out_len);
int mdz_ref_brle_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk,
                        uint8_t** data, uint64_t** ptr, uint64_t* n_chunks, int threads);
int mdz_ref_brle_decode(const uint8_t* data, const uint64_t* ptr, uint64_t n_chunks,
                        uint64_t chunk, uint8_t* out, uint64_t out_bytes,
                        int32_t* status, int threads);

/* ============================================================== LZSS ===== *
 * Lane-transposed LZSS (include/lzss/lzss_gpu2.h).  Lane t of a chunk is the
 * bytes chunk[128k + 4t + j]; lanes are sorted by stream length (descending,
 * ties to the higher lane id) and col_map[32c+s] names the original lane at
 * sorted position s.  blk_off[n] is the payload size. */
uint64_t mdz_ref_lzss_lane_encode(const uint8_t* lane, uint64_t n, uint8_t* out, uint64_t cap);
/* same stream, matches found by the literal find_match scan (lzss_gpu2.h:70-109) */
uint64_t mdz_ref_lzss_lane_encode_naive(const uint8_t* lane, uint64_t n, uint8_t* out,
                                        uint64_t cap);
int mdz_ref_lzss_lane_decode(const uint8_t* in, uint64_t in_len, uint8_t* out, uint64_t out_len);
int mdz_ref_lzss_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** payload,
                        uint64_t* payload_bytes, uint64_t** col_len, uint64_t** blk_off,
                        uint8_t** col_map, uint64_t* n_chunks, int threads);
int mdz_ref_lzss_decode(const uint8_t* payload, uint64_t payload_bytes, const uint64_t* blk_off,
                        const uint64_t* col_len, const uint8_t* col_map, uint64_t n_chunks,
                        uint64_t chunk, uint8_t* out, uint64_t out_bytes, int32_t* status,
                        int threads);

/* =========================================================== Deflate ===== *
 * zlib is the oracle (RFC 1950/1951).  zlib_hdr: check the 2-byte zlib
 * header, then inflate raw (the Adler-32 trailer is not verified).  Writes
 * at most cap bytes; *written = bytes produced.  Status per cuDF's codes:
 * 1 output full before the end, 2 input ended early, negatives per zlib's
 * error (invalid block type -1, stored lengths -2, code lengths set -3,
 * bit length repeat -4, literal/lengths set -5, distances set -6, too many
 * symbols -7, missing end-of-block -8, invalid distance code / too far back
 * -9, invalid literal/length code -10, header -11). */
int mdz_ref_inflate(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap, uint64_t* written,
                    int zlib_hdr);
int mdz_ref_inflate_batch(const uint8_t* src, const uint64_t* src_off, const uint64_t* src_len,
                          uint8_t* dst, const uint64_t* dst_off, const uint64_t* dst_cap,
                          uint64_t count, int zlib_hdr, uint64_t* written, int32_t* status,
                          int threads);
/* Config 4 batch generator (see inflate.c). */
int mdz_gen_deflate_batch(uint64_t count, uint64_t seed, uint64_t min_size, uint64_t max_size,
                          uint8_t** src, uint64_t** src_off, uint64_t** src_len, uint8_t** raw,
                          uint64_t** raw_off, uint64_t** raw_len, uint8_t** mode, int threads);
const char* mdz_zlib_version(void);

/* ======================================================== generators ===== */
/* Config 1: RLE v1 u32 column (runs 70% / literal segments 30% by count). */
void mdz_gen_rlev1_u32(uint32_t* out, uint64_t n, uint64_t seed);
/* Config 2(a): int64 column of SHORT_REPEAT / DELTA / DIRECT / PATCHED_BASE
 * shaped segments weighted 15/35/35/15. */
void mdz_gen_rlev2_i64(uint64_t* out, uint64_t n, uint64_t seed);
/* Config 2(b): byte runs, fresh byte != previous, length Geom(1/12). */
void mdz_gen_brle_u8(uint8_t* out, uint64_t n, uint64_t seed);
/* Config 3: 64-symbol order-2 Markov bytes (Dirichlet(0.3) rows) with
 * 40% copy events of 4..256 bytes from up to 4096 back. */
void mdz_gen_lzss_markov(uint8_t* out, uint64_t n, uint64_t seed);
/* The Markov source behind config 3 and 4: a transition table (4096 x 64
 * CDF rows, free with mdz_free) and a sequence with copy probability copy_p. */
double* mdz_markov_table(uint64_t seed);
void mdz_markov_gen(const double* cdf, uint8_t* out, uint64_t n, uint64_t seed, double copy_p);

#ifdef __cplusplus
}
#endif
#endif
