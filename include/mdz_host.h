/*
 * mdz host library (compression_amd/libmdz_host.so): the CPU side of the
 * codecs -- the container encoders (compress paths) that produce the
 * reference's chunked formats, and the seeded synthetic workloads.
 *
 * Each encoder restates the reference encoder it names, so its output is the
 * byte stream the reference's decoders (and libmdz's) take.  Buffers these
 * functions allocate are released with mdz_free.  Return codes: 0, or -1 for
 * parameters the format cannot hold.
 */
#ifndef MDZ_HOST_H
#define MDZ_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- PRNG --
 * splitmix64-seeded xoshiro256** (the generator family named for the
 * synthetic workloads). */
typedef struct { uint64_t s[4]; } mdz_rng;
void     mdz_rng_seed(mdz_rng* r, uint64_t seed);
uint64_t mdz_rng_next(mdz_rng* r);
uint64_t mdz_rng_below(mdz_rng* r, uint64_t n);   /* [0, n), 64x64->128 multiply-high */
double   mdz_rng_unit(mdz_rng* r);                /* [0, 1), top 53 bits */

/* ------------------------------------------------------------ threads -- */
typedef void (*mdz_task_fn)(void* ctx, uint64_t index);
/* fn(ctx, i) for i in [0, n) on `threads` pthreads (dynamic scheduling). */
void mdz_parallel_for(uint64_t n, int threads, mdz_task_fn fn, void* ctx);
int  mdz_host_cores(void);      /* sched_getaffinity count */
void mdz_free(void* p);

/* ------------------------------------------------------------- RLE v1 --
 * include/rle_v1/encoder.h (compression_warp :607-1018, interleave
 * :564-588, col_len/blk_off :540-559,1188-1193): chunk of `chunk` bytes cut
 * into lines of `line` bytes (COMP_COL_LEN), line j to lane j % 32. */
uint64_t mdz_rlev1_lane_encode(const void* vals, uint64_t n, int elem_bytes, uint8_t* out,
                               uint64_t cap);
int mdz_rlev1_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, int elem_bytes,
                     uint32_t line, uint8_t** payload, uint64_t* payload_bytes,
                     uint64_t** col_len, uint64_t** blk_off, uint64_t* n_chunks, int threads);

/* ------------------------------------------------------------- RLE v2 --
 * include/rlev2/encoder.h + encoder_trans.h (theirs side): lane t owns chunk
 * elements k*32*RU + t*RU + j; blk_off[n] = the raw size. */
uint64_t mdz_rlev2_lane_encode(const uint64_t* vals, uint64_t n, uint8_t* out, uint64_t cap);
int mdz_rlev2_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, int elem_bytes,
                     uint32_t read_unit, uint8_t** payload, uint64_t* payload_bytes,
                     uint64_t** col_len, uint64_t** blk_off, uint64_t* n_chunks, int threads);

/* Plain RLE v2 (one run stream per chunk of int64 values): one chunk's
 * stream (block_encode, include/rlev2/encoder.h:808-958, HEAD side), and the
 * whole container as compress_gpu writes it (encoder.h:1003-1058):
 * [u32 n+1][u64 ptr[n+1]][u64 in_bytes][streams].  Values must lie in
 * [0, 2^63) (returns -2 otherwise); *out is freed with mdz_free. */
uint64_t mdz_rlev2_plain_chunk_encode(const int64_t* vals, uint64_t nv, uint8_t* out, uint64_t cap);
int mdz_rlev2_plain_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** out,
                           uint64_t* out_bytes, int threads);

/* ----------------------------------------------------------- byte RLE --
 * include/brle/brle.h (encode :129-230, compress_gpu :251-303). */
uint64_t mdz_brle_chunk_encode(const uint8_t* in, uint64_t n, uint8_t* out);
int mdz_brle_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** data,
                    uint64_t** ptr, uint64_t* n_chunks, int threads);

/* --------------------------------------------------------------- LZSS --
 * include/lzss/lzss_gpu2.h (find_match :70-109, lane encoder :332-707, lane
 * sort :1063-1085): lane t = chunk bytes 128k + 4t + j; lanes sorted by
 * length, descending, ties to the higher lane; trailing partial chunk
 * dropped. */
uint64_t mdz_lzss_lane_encode(const uint8_t* lane, uint64_t n, uint8_t* out, uint64_t cap);
int mdz_lzss_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** payload,
                    uint64_t* payload_bytes, uint64_t** col_len, uint64_t** blk_off,
                    uint8_t** col_map, uint64_t* n_chunks, int threads);
/* Plain (per-chunk) LZSS, compress_func + kernel_shift_data
 * (lzss_gpu2.h:1450-1566,1926-1948): chunk streams packed back to back,
 * lens[c] = inclusive prefix sum of their lengths; the last chunk may be
 * shorter than `chunk`. */
int mdz_lzss_plain_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** data,
                          uint64_t** lens, uint64_t* n_chunks, int threads);

/* ------------------------------------------ byte RLE, lane-transposed --
 * include/brle/brle_trans.h (compress_init_func :238-379, compress_func
 * :401-562, lane sort :204-224, compress_gpu :565-669): the LZSS lane
 * container (the same sort, interleave and 128-byte chunk extents) holding
 * flag-group streams whose 2-byte tokens are (count, value) runs. */
uint64_t mdz_brle_trans_lane_encode(const uint8_t* lane, uint64_t n, uint8_t* out, uint64_t cap);
int mdz_brle_trans_encode(const uint8_t* in, uint64_t in_bytes, uint64_t chunk, uint8_t** payload,
                          uint64_t* payload_bytes, uint64_t** col_len, uint64_t** blk_off,
                          uint8_t** col_map, uint64_t* n_chunks, int threads);

/* ----------------------------------------------------- synthetic data -- */
/* Config 1: u32 column, runs 70% / literal segments 30% by count. */
void mdz_gen_rlev1_u32(uint32_t* out, uint64_t n, uint64_t seed);
/* Config 2(a): int64 column of SHORT_REPEAT / DELTA / DIRECT / PATCHED_BASE
 * shaped segments weighted 15/35/35/15. */
void mdz_gen_rlev2_i64(uint64_t* out, uint64_t n, uint64_t seed);
/* Config 2(b): byte runs, fresh byte != previous, length Geom(1/12). */
void mdz_gen_brle_u8(uint8_t* out, uint64_t n, uint64_t seed);
/* Configs 3-5: 64-symbol order-2 Markov text (Dirichlet(0.3) rows) with
 * copies of 4..256 bytes from up to 4096 back with probability copy_p. */
double* mdz_markov_table(uint64_t seed);
void mdz_markov_gen(const double* cdf, uint8_t* out, uint64_t n, uint64_t seed, double copy_p);
void mdz_gen_lzss_markov(uint8_t* out, uint64_t n, uint64_t seed);   /* config 3, copy_p 0.4 */
/* Config 4: `count` zlib streams, sizes log-uniform [min, max], 80% level 6,
 * 10% level 1 + Z_FIXED, 10% stored; 16-byte aligned starts. */
int mdz_gen_deflate_batch(uint64_t count, uint64_t seed, uint64_t min_size, uint64_t max_size,
                          uint8_t** src, uint64_t** src_off, uint64_t** src_len, uint8_t** raw,
                          uint64_t** raw_off, uint64_t** raw_len, uint8_t** mode, int threads);

#ifdef __cplusplus
}
#endif
#endif
