/*
 * mdz -- MI355X-native (gfx950) decompression library: C ABI.
 *
 * Every decode entry point takes DEVICE pointers, is stream-ordered (it only
 * enqueues work on `stream`, a hipStream_t passed as void*; NULL = default
 * stream) and is reentrant.  It returns MDZ_OK or a negative MDZ_ERR_* API
 * error; per-chunk (per-stream) decode status lands in a device int32 array
 * using the MDZ_ST_* codes below, which follow the cuDF inflate status
 * convention (include/cudf_deflate.cuh:440,527-574 in the reference).
 *
 * The *_host entry points keep the reference's host-pointer signature shape
 * (callee allocates *out, caller frees it with mdz_host_free) so a reference
 * driver works by swapping the call.  They copy in, decode, synchronise and
 * copy out.
 */
#ifndef MDZ_H
#define MDZ_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ API errors */
#define MDZ_OK               0
#define MDZ_ERR_INVALID     -1   /* bad argument / container metadata */
#define MDZ_ERR_HIP         -2   /* a HIP runtime call failed */
#define MDZ_ERR_NOMEM       -3   /* device or host allocation failed */
#define MDZ_ERR_UNSUPPORTED -4   /* parameter combination not built */
#define MDZ_ERR_CORRUPT     -5   /* *_host wrappers: some chunk status != 0 */

/* ------------------------------------------------- per-chunk status codes */
#define MDZ_ST_OK               0
#define MDZ_ST_OUTPUT_OVERRUN   1   /* decoded data exceeds the output slot */
#define MDZ_ST_INPUT_OVERRUN    2   /* input ended before the output was complete */
#define MDZ_ST_BAD_BLOCK_TYPE  -1
#define MDZ_ST_STORED_LEN      -2
#define MDZ_ST_BAD_CODE_LENS   -3
#define MDZ_ST_BAD_REPEAT      -4
#define MDZ_ST_BAD_LITLEN      -5
#define MDZ_ST_BAD_DIST        -6
#define MDZ_ST_BAD_COUNTS      -7
#define MDZ_ST_NO_EOB          -8
#define MDZ_ST_BAD_DISTANCE    -9
#define MDZ_ST_BAD_SYMBOL     -10
#define MDZ_ST_BAD_HEADER     -11
#define MDZ_ST_INTERNAL       -12   /* a decoder's bounded internal wait expired (never expected) */
#define MDZ_ST_BAD_CHECKSUM   -13   /* Deflate with MDZ_HDR_VERIFY: trailer checksum or size differs */

/* --------------------------------------------------------------- version */
/* Returns "mdz <version> gfx950". */
const char* mdz_version(void);
/* Text of the last error an mdz call on this host thread recorded ("" if none). */
const char* mdz_last_error(void);

/* -------------------------------------------- device plumbing (thin HIP) */
/* For hosts that do not link HIP themselves (the Python tests and bench). */
int mdz_device_count(int* count);
int mdz_dev_malloc(void** ptr, uint64_t bytes);
int mdz_dev_free(void* ptr);
/* Pinned (page-locked) host memory: copies from it to the device are DMA
 * transfers that return at once (hipHostMalloc / hipHostFree). */
int mdz_pinned_alloc(void** ptr, uint64_t bytes);
int mdz_pinned_free(void* ptr);
int mdz_memcpy_htod(void* dst, const void* src, uint64_t bytes, void* stream);
int mdz_memcpy_dtoh(void* dst, const void* src, uint64_t bytes, void* stream);
int mdz_memset_d8(void* dst, int value, uint64_t bytes, void* stream);
/* Device-to-device copy by a streaming kernel (16-byte vector loads/stores);
 * also the known-size calibration kernel for HBM traffic counters. */
int mdz_dev_copy(void* dst, const void* src, uint64_t bytes, void* stream);
int mdz_stream_create(void** stream);
int mdz_stream_destroy(void* stream);
int mdz_stream_sync(void* stream);
int mdz_device_sync(void);
int mdz_event_create(void** event);
int mdz_event_destroy(void* event);
int mdz_event_record(void* event, void* stream);
int mdz_event_sync(void* event);
/* 1 if the event's work has completed, 0 if not, negative on error. */
int mdz_event_query(void* event);
int mdz_event_elapsed_ms(void* start, void* stop, float* ms);
void mdz_host_free(void* p);

/* ==================================================== lane-transposed ==== *
 * RLE v1, RLE v2 and LZSS share one container shape (reference
 * include/rle_v1/encoder.h:540-588, include/rlev2/encoder_trans.h:291-349,
 * include/lzss/lzss_gpu2.h:497-509): chunk c's payload is
 * [blk_off[c], blk_off[c+1]) (128-byte aligned); inside it the 32 lane
 * streams of col_len[32c+t] bytes are interleaved in 4-byte words, compacted
 * round-robin over the lanes that still have words. */
typedef struct {
    const void*     payload;     /* chunk data (no leading size word) */
    const uint64_t* blk_off;     /* n_chunks + 1 byte offsets into payload */
    const uint64_t* col_len;     /* 32 * n_chunks lane-stream byte lengths */
    const uint8_t*  col_map;     /* LZSS only: 32 * n_chunks sorted->original lane */
    uint64_t        n_chunks;
    uint64_t        chunk_bytes; /* uncompressed bytes per chunk */
    uint32_t        elem_bytes;  /* RLE v1 DATA_TYPE / RLE v2 COMP_TYPE size {1,2,4,8} */
    uint32_t        lane_line;   /* RLE v1: COMP_COL_LEN bytes per lane line;
                                    RLE v2: READ_UNIT elements per lane line */
    uint64_t        payload_bytes; /* payload extent for bounds checks (0 = unchecked);
                                    RLE v2 needs it for its last chunk, whose
                                    blk_off[n] holds the uncompressed size */
} mdz_lanes_t;

/* Tuning knobs for the lane-transposed decoders; zero fields pick defaults. */
typedef struct {
    uint32_t ring_words;         /* force the streaming (windowed) path with this LDS ring,
                                    in 4-byte words (power of 2); 0 = automatic */
    uint32_t max_chunk_bytes;    /* upper bound on one chunk's payload bytes, if known; it
                                    sizes the per-chunk LDS area (0 = use the maximum) */
    uint32_t max_part_bytes;     /* RLE v2: upper bound, over chunks and their 4 lane quarters
                                    (lanes 8q..8q+7), of sum(round_up(col_len, 16)); when set,
                                    each quarter is decoded by its own workgroup (4 per chunk,
                                    smaller LDS, more chunks in flight).  0 = whole chunks */
    uint32_t max_lane_tokens;    /* RLE v1: at most this many tokens listed per lane stream
                                    (0 = as many as fit in LDS); a chunk with a longer lane
                                    takes the checkpoint decode.  For testing that path */
    uint32_t rlev2_window;       /* RLE v2: run records a consumer wavefront claims at once
                                    from the header walk's ring: 8, 16, 32 or 64 (0 = 16) */
    uint32_t reserved[3];
} mdz_lanes_tune_t;

/* RLE v1 decode -- replaces rle_v1::decompress_gpu<READ_COL_TYPE, DATA_TYPE,
 * OUT_COL_TYPE, QUEUE_SIZE> (reference include/rle_v1/rlev1.h:99-199) and its
 * inflate kernel (include/rle_v1/decoder.h:426-470).  Chunk c is written to
 * dst[c*chunk_bytes, (c+1)*chunk_bytes) clipped to dst_bytes.  status: device
 * int32[n_chunks] or NULL. */
int mdz_rlev1_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes,
                         int32_t* status, void* stream);
int mdz_rlev1_decompress_tuned(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes,
                               int32_t* status, const mdz_lanes_tune_t* tune, void* stream);

/* Host-pointer RLE v1 decode in the reference's signature shape
 * (include/rle_v1/rlev1.h:99-100): `in` is the reference payload file
 * [u64 uncompressed_bytes][chunks...]; metadata that the reference reads
 * from ./input_data/rle_v1_{col_len,blk_offset}.bin is passed explicitly.
 * *out (uncompressed_bytes) is allocated here; free with mdz_host_free. */
int mdz_rlev1_decompress_host(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                              uint64_t* out_n_bytes, int comp_col_len, uint64_t chunk_size,
                              uint32_t elem_bytes, const uint64_t* col_len,
                              const uint64_t* blk_off, uint64_t n_chunks);

/* Exactly rle_v1::decompress_gpu's arguments (rlev1.h:99-100), reading the
 * side files ./input_data/rle_v1_col_len.bin and rle_v1_blk_offset.bin from
 * the current directory as the reference does (rlev1.h:110-147). */
int mdz_rle_v1_decompress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                              uint64_t* out_n_bytes, int comp_col_len, uint64_t chunk_size,
                              uint32_t elem_bytes);

/* RLE v1 encode -- replaces rle_v1::compress_gpu and its kernels (reference
 * include/rle_v1/rlev1.h:24-97; encoder include/rle_v1/encoder.h:607-1018).
 * payload, col_len and blk_off equal that encoder's byte for byte.  Device
 * form: in (device, in_bytes) in chunks of chunk_bytes, elem_bytes-wide
 * values in lines of comp_col_len bytes; work (work_bytes) and payload
 * (payload_cap) sized by mdz_rlev1_compress_bound; col_len (device u64[32n])
 * and blk_off (device u64[n + 1]) receive the metadata, blk_off[n] = payload
 * bytes. */
int mdz_rlev1_compress_bound(uint64_t in_bytes, uint64_t chunk_bytes, uint32_t elem_bytes, uint32_t comp_col_len,
                             uint64_t* n_chunks, uint64_t* work_bytes, uint64_t* payload_cap);
int mdz_rlev1_compress(const void* in, uint64_t in_bytes, uint64_t chunk_bytes, uint32_t elem_bytes,
                       uint32_t comp_col_len, void* work, void* payload, uint64_t* col_len, uint64_t* blk_off,
                       void* stream);
/* rle_v1::compress_gpu's host form: *out = [u64 in_n_bytes][payload]
 * (allocated here; mdz_host_free), and col_len / blk_off written to
 * ./input_data/rle_v1_col_len.bin and rle_v1_blk_offset.bin as the reference
 * writes them (rlev1.h:85-93), where mdz_rle_v1_decompress_gpu reads them. */
int mdz_rle_v1_compress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes, uint64_t* out_n_bytes,
                            int comp_col_len, uint64_t chunk_size, uint32_t elem_bytes);

/* RLE v2 decode -- replaces rlev2::decompress_gpu<READ_UNIT, queue_size,
 * COMP_TYPE> (reference include/rlev2/decoder_trans.h:615-681) and its
 * inflate kernel (include/rlev2/decoder_new.h:618-657).  elem_bytes =
 * sizeof(COMP_TYPE), lane_line = READ_UNIT; lane t's q-th value is element
 * (q/RU)*32*RU + t*RU + q%RU of its chunk.  blk_off[n_chunks] is not read
 * (the reference stores the uncompressed size there); set payload_bytes to
 * bound the last chunk.  PATCHED_BASE patches use a 64-bit mask (the
 * reference's 32-bit mask, decoder_new.h:533, drops bits >= 32). */
int mdz_rlev2_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes,
                         int32_t* status, void* stream);
int mdz_rlev2_decompress_tuned(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes,
                               int32_t* status, const mdz_lanes_tune_t* tune, void* stream);

/* Host-pointer RLE v2 decode in the reference's signature shape
 * (decoder_trans.h:615-618): in/in_n_bytes = payload, blk_off[n_chunks+1]
 * with blk_off[n_chunks] = uncompressed bytes, col_len[32*n_chunks].
 * *out (blk_off[n_chunks] bytes) is allocated here; free with mdz_host_free. */
int mdz_rlev2_decompress_host(const uint8_t* in, uint64_t in_n_bytes, uint64_t n_chunks,
                              const uint64_t* blk_off, const uint64_t* col_len, uint8_t** out,
                              uint64_t* out_n_bytes, uint64_t chunk_size, uint32_t read_unit,
                              uint32_t elem_bytes);

/* Plain RLE v2 decode -- replaces rlev2::kernel_decompress / block_decode
 * (reference include/rlev2/decoder.h:190-226): one run stream per chunk of
 * int64 values.  Chunk c's stream is data[ptr[c], ptr[c+1]) (device u64
 * ptr[n_chunks+1], the container's offset table) and decodes to
 * dst[c*chunk_bytes, ...) clipped to dst_bytes (8-byte aligned).  A chunk
 * ends where its stream does.  status[c] (optional) is the first failing
 * run's problem: INPUT_OVERRUN (the stream ends inside a run),
 * OUTPUT_OVERRUN (more than chunk_bytes/8 values), BAD_SYMBOL (DELTA length
 * 1, a varint over 10 bytes, a patch past its run) or BAD_HEADER (ptr out of
 * order or past data_bytes); values of the runs before it are written.
 * PATCHED_BASE uses the full patch width as its mask (decoder.h:176 builds
 * it in 16 bits), and a packed-delta run writes only its own values
 * (decoder.h:117-124 rewrites the two after it). */
int mdz_rlev2_plain_decompress(const void* data, uint64_t data_bytes, const uint64_t* ptr, uint64_t n_chunks,
                               uint64_t chunk_bytes, void* dst, uint64_t dst_bytes, int32_t* status,
                               void* stream);
/* rlev2::decompress_gpu for the plain container (decoder.h:227-262):
 * in = [u32 n+1][u64 ptr[n+1]][u64 raw_bytes][streams]; *out (raw_bytes) is
 * allocated here, free with mdz_host_free.  chunk_bytes is the reference's
 * compile-time CHUNK_SIZE (32 KiB on the HEAD side of utils.h:62-66, which
 * mdz_rlev2_plain_decompress_gpu uses).  MDZ_ERR_CORRUPT if a chunk fails. */
int mdz_rlev2_plain_decompress_host(const uint8_t* in, uint64_t in_n_bytes, uint64_t chunk_bytes, uint8_t** out,
                                    uint64_t* out_n_bytes);
int mdz_rlev2_plain_decompress_gpu(const uint8_t* in, uint64_t in_n_bytes, int64_t** out, uint64_t* out_n_bytes);

/* RLE v2 encode -- replaces rlev2::compress_gpu_transpose (reference
 * include/rlev2/encoder_trans.h:428-537; block_encode :14-289; writers
 * include/rlev2/encoder.h:124-785).  payload, col_len and blk_off equal that
 * encoder's byte for byte; blk_off[n] holds in_bytes (the uncompressed size,
 * as the reference leaves it) and *payload_bytes (device u64, may be NULL)
 * the payload's length.  work / payload sized by mdz_rlev2_compress_bound. */
int mdz_rlev2_compress_bound(uint64_t in_bytes, uint64_t chunk_bytes, uint32_t elem_bytes, uint32_t read_unit,
                             uint64_t* n_chunks, uint64_t* work_bytes, uint64_t* payload_cap);
int mdz_rlev2_compress(const void* in, uint64_t in_bytes, uint64_t chunk_bytes, uint32_t elem_bytes,
                       uint32_t read_unit, void* work, void* payload, uint64_t* col_len, uint64_t* blk_off,
                       uint64_t* payload_bytes, void* stream);

/* ========================================================== byte RLE ===== *
 * Byte RLE decode -- replaces brle::decompress_gpu / kernel_decompress
 * (reference include/brle/brle.h:26-109,305-342).  Chunk c's compressed bytes
 * are data[ptr[c], ptr[c+1]) and decode to dst[c*chunk_bytes ...] (at most
 * chunk_bytes, clipped to dst_bytes).  max_chunk_bytes bounds one chunk's
 * compressed size (0 = chunk_bytes + chunk_bytes/128 + 1, the format's worst
 * case); it must fit a workgroup's LDS (about 150 KiB). */
int mdz_brle_decompress(const uint64_t* ptr, uint64_t n_chunks, uint64_t chunk_bytes,
                        const void* data, void* dst, uint64_t dst_bytes, int32_t* status,
                        void* stream);
int mdz_brle_decompress_tuned(const uint64_t* ptr, uint64_t n_chunks, uint64_t chunk_bytes,
                              const void* data, void* dst, uint64_t dst_bytes, int32_t* status,
                              uint32_t max_chunk_bytes, void* stream);

/* brle::decompress_gpu's host form (brle.h:305): `in` is the container
 * [u32 n_ptr][u64 ptr[n_ptr]][u64 out_bytes][data]; chunk_bytes is the
 * reference's CHUNK_SIZE (256 KiB there). */
int mdz_brle_decompress_host(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                             uint64_t* out_n_bytes, uint64_t chunk_bytes);
/* Exactly brle::decompress_gpu(in, out, in_n_bytes, out_n_bytes), CHUNK_SIZE
 * = 256 KiB (brle.h:8). */
int mdz_brle_decompress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                            uint64_t* out_n_bytes);

/* Byte RLE encode -- replaces brle::compress_gpu and its kernels (reference
 * include/brle/brle.h:232-303; the encoder :129-230).  The bytes equal that
 * encoder's.  Device form: in (in_bytes, device) is cut into chunk_bytes
 * chunks; work (work_bytes from mdz_brle_compress_bound) holds per-chunk
 * slots and sizes; data (data_cap bytes) receives the packed chunk data and
 * ptr (device u64[n + 1]) its offsets, ptr[0] = 0 and ptr[n] = data bytes. */
int mdz_brle_compress_bound(uint64_t in_bytes, uint64_t chunk_bytes, uint64_t* n_chunks,
                            uint64_t* work_bytes, uint64_t* data_cap);
int mdz_brle_compress(const void* in, uint64_t in_bytes, uint64_t chunk_bytes, void* work, void* data,
                      uint64_t* ptr, void* stream);
/* brle::compress_gpu's host form: in (host) -> *out = the reference container
 * [u32 n + 1][u64 ptr[n + 1]][u64 in_n_bytes][data] with 256 KiB chunks
 * (brle.h:8,278-292), allocated here; free with mdz_host_free. */
int mdz_brle_compress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes, uint64_t* out_n_bytes);

/* ============================================================== LZSS ===== *
 * LZSS decode -- replaces lzss::kernel_decompress / decompress_func
 * (reference include/lzss/lzss_gpu2.h:1898-1907,114-325) and the device part
 * of lzss::decompress_gpu (:2129-2314).  in->col_len holds the lengths in
 * sorted order and in->col_map[32c+s] the original lane of sorted position s
 * (lzss_gpu2.h:1063-1085); blk_off[n_chunks] is the payload size.  Lane t's
 * byte i decodes to dst[c*chunk + 128*(i/4) + 4t + i%4], clipped to
 * dst_bytes.  chunk_bytes must be a multiple of 128 and at most 128 KiB
 * (lanes of up to 4 KiB).  Decoding stops when a lane's chunk_bytes/32 bytes
 * are complete: the reserved flag byte after an 8-token group that ends the
 * stream is not decoded (the reference decodes it and writes past the lane,
 * lzss_gpu2.h:200-321).  status[c] (optional): the error of the chunk's
 * lowest failing lane, BAD_HEADER for bad container metadata. */
int mdz_lzss_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                        void* stream);

/* lzss::decompress_gpu's signature (lzss_gpu2.h:2129): in = payload,
 * in_metadata = the four .meta blobs {u64 chunk_size, u64 blk_off[n+1],
 * u64 col_len[32n], u8 col_map[32n]} with their byte lengths.  Returns -1 when
 * fewer than 4 blobs are given, as the reference does; *out (n*chunk_size
 * bytes, allocated here, free with mdz_host_free).  The output stride is the
 * chunk_size blob (the reference uses its compile-time 128 KiB). */
int mdz_lzss_decompress_host(const uint8_t* in, uint64_t in_n_bytes,
                             const uint8_t* const* in_metadata, const uint64_t* in_metadata_lens,
                             uint64_t in_metadata_num, uint8_t** out, uint64_t* out_n_bytes);

/* Lane-transposed byte RLE decode -- replaces brle_trans::decompress_func
 * (reference include/brle/brle_trans.h:72-202) and the device part of
 * brle_trans::decompress_gpu (:672-790).  The container is the LZSS one
 * (in->col_len sorted, in->col_map, blk_off; lzss_gpu2.h:1063-1085 and
 * brle_trans.h:204-224), with the same flag groups: bit b (LSB first) of a
 * flag byte 1 = literal byte, 0 = run (count byte, value byte).  Lane t's
 * byte i decodes to dst[c*chunk + 128*(i/4) + 4t + i%4], clipped to
 * dst_bytes; chunk_bytes a multiple of 128, at most 128 KiB.  A lane stops
 * when its chunk_bytes/32 bytes are complete; status[c] (optional) as for
 * mdz_lzss_decompress (a run past the lane end is OUTPUT_OVERRUN; the
 * reference reads col_len bytes and writes wherever they lead). */
int mdz_brle_trans_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                              void* stream);

/* Plain LZSS decode -- replaces lzss::kernel_decompress(in, out, in_n_bytes,
 * out_n_bytes, out_chunk_size, n_chunks, lens) and the plain decompress_func
 * (reference include/lzss/lzss_gpu2.h:1908-1913,718-792).  Chunk c's stream
 * is data[lens[c-1], lens[c]) (lens = inclusive prefix sums of the stream
 * lengths, as compress_func :1450-1566 and its scan leave them; lens[-1] = 0)
 * with the same token grammar as the lane streams, and decodes to
 * dst[c*chunk_bytes, ...) clipped to dst_bytes, so the last chunk may be
 * short.  The reference strides its output by the compile-time CHUNK_SIZE
 * (128 KiB) and stops when a chunk's input runs out; here a chunk stops when
 * its output is complete, and status[c] (optional) gives the first problem in
 * stream order: INPUT_OVERRUN (the stream ends first), BAD_DISTANCE (a match
 * reaching before the chunk start), OUTPUT_OVERRUN (a match past the chunk's
 * end), or BAD_HEADER (lens[c] < lens[c-1] or lens[c] > data_bytes).  Bytes
 * of a failed chunk are unspecified.  chunk_bytes at most 256 MiB; any
 * alignment (16-byte aligned output chunks take the vector-store path). */
int mdz_lzss_plain_decompress(const void* data, uint64_t data_bytes, const uint64_t* lens, uint64_t n_chunks,
                              uint64_t chunk_bytes, void* dst, uint64_t dst_bytes, int32_t* status, void* stream);

/* LZSS encode -- replaces lzss::compress_gpu (reference include/lzss/
 * lzss_gpu2.h:1952-2123: find_match :70-109, lane encoder :332-494,511-707,
 * lane sort :1063-1085, interleave :497-509).  payload, col_len (sorted),
 * col_map and blk_off equal that encoder's byte for byte.  Whole chunks only
 * (in_bytes / chunk_bytes of them; a trailing partial chunk is dropped, as
 * the reference encoder drops it); chunk_bytes a multiple of 128, at most
 * 1 MiB; in 4-byte aligned.  work / payload sized by mdz_lzss_compress_bound. */
int mdz_lzss_compress_bound(uint64_t in_bytes, uint64_t chunk_bytes, uint64_t* n_chunks, uint64_t* work_bytes,
                            uint64_t* payload_cap);
int mdz_lzss_compress(const void* in, uint64_t in_bytes, uint64_t chunk_bytes, void* work, void* payload,
                      uint64_t* col_len, uint8_t* col_map, uint64_t* blk_off, void* stream);
/* lzss::compress_gpu's host form (128 KiB chunks): *out = payload, and the
 * metadata blobs [u64 chunk][blk_off][col_len][col_map] with their lengths,
 * the shape mdz_lzss_decompress_host takes.  Free *out, each blob, and the
 * two arrays with mdz_host_free. */
int mdz_lzss_compress_gpu(const uint8_t* in, uint64_t in_n_bytes, uint8_t** out, uint64_t* out_n_bytes,
                          uint8_t*** out_metadata, uint64_t** out_metadata_lens, uint64_t* out_metadata_num);

/* Splits a .meta side file image [u64 num][u64 len[num]][blob...] into blob
 * pointers (into `meta`) and lengths -- the reader in src/lzss/lzss_gpu.cu:
 * 109-133 (format: include/utils.h:46-118).  *num is in/out: capacity of
 * blobs/lens in, blob count out.  MDZ_ERR_CORRUPT if the lengths overrun. */
int mdz_meta_split(const uint8_t* meta, uint64_t meta_bytes, const uint8_t** blobs, uint64_t* lens,
                   uint64_t* num);

/* =========================================================== Deflate ===== *
 * Batched zlib/Deflate inflate -- replaces cudf::io::gpuinflate and its
 * structs (reference include/cudf_deflate.cuh:236-250,1399-1407) and the
 * repo's deflate::decompress_gpu kernels (include/zlib/zlib.h:832-962).
 * Layouts equal gpu_inflate_input_s / gpu_inflate_status_s. */
typedef struct {
    const void* src;       /* device pointer to the stream */
    uint64_t src_bytes;
    void* dst;             /* device pointer to its output slot */
    uint64_t dst_bytes;
} mdz_inflate_in_t;

typedef struct {
    uint64_t bytes_written;
    uint32_t status;       /* MDZ_ST_* as uint32 (negative codes wrap) */
    uint32_t reserved;     /* input bytes consumed through the last block (from the
                              stream's start; gzip: from the Deflate data's start),
                              saturating at 2^32 - 1.  cuDF leaves a debug count here */
} mdz_inflate_status_t;

/* parse_hdr: the framing around each stream's Deflate data.
 *   MDZ_HDR_RAW   raw Deflate;
 *   MDZ_HDR_ZLIB  2-byte zlib header (checked: MDZ_ST_BAD_HEADER), then
 *                 Deflate; the 4-byte Adler-32 trailer follows the last block;
 *   MDZ_HDR_GZIP  gzip member (cuDF's parse_hdr != 0, reference
 *                 include/cudf_deflate.cuh:1179-1217,1244-1253): header
 *                 skipped by cuDF's rules (a rejected header reports -1, as
 *                 cuDF does), the last 8 bytes (CRC-32, ISIZE) are the footer
 *                 and are not Deflate input.  FEXTRA skips XLEN's own 2 bytes
 *                 too (RFC 1952; cuDF omits them).
 * MDZ_HDR_VERIFY may be or'ed onto ZLIB or GZIP: a stream that decoded with
 * status 0 is then checked against its trailer (Adler-32; or CRC-32 and
 * ISIZE) and reports MDZ_ST_BAD_CHECKSUM on a mismatch, or
 * MDZ_ST_INPUT_OVERRUN for a zlib stream whose trailer is cut off.  cuDF
 * verifies neither; this is the check zlib's own inflate makes. */
#define MDZ_HDR_RAW    0
#define MDZ_HDR_ZLIB   1
#define MDZ_HDR_GZIP   2
#define MDZ_HDR_VERIFY 0x100

/* in/out: device arrays of `count` entries; parse_hdr as above (other values:
 * MDZ_ERR_INVALID).  One stream per workgroup, launched in index order (the
 * plain per-stream decoder). */
int mdz_inflate_batch(const mdz_inflate_in_t* in, mdz_inflate_status_t* out, int count,
                      int parse_hdr, void* stream);

/* Stored streams: copies the first min(src_bytes, dst_bytes) bytes of each
 * entry of in[count] (device array) to its dst.  Replaces cuDF's
 * gpu_copy_uncompressed_blocks / copy_uncompressed_kernel (reference
 * include/cudf_deflate.cuh:1349-1397,1409-1415); sizes stay 64-bit where
 * cuDF truncates them to 32 bits. */
int mdz_copy_uncompressed(const mdz_inflate_in_t* in, int count, void* stream);

/* Per-stream decode-cost sample the decoder logs for the learned scheduler:
 * shader-clock cycles from the stream's first to last instruction and what
 * the stream held. */
typedef struct {
    uint64_t cycles;
    uint32_t symbols;          /* literals + matches */
    uint32_t matches;
    uint32_t stored_blocks;
    uint32_t fixed_blocks;
    uint32_t dynamic_blocks;
    uint32_t reserved;
} mdz_inflate_cost_t;

/* Persistent form: `blocks` workgroups (0 = one per co-resident slot) pull
 * streams from a device work queue in the order order[0..count) (device u32
 * array; NULL = index order).  `cursor` is a device u32 the call resets
 * before the launch; `cost` (device array of count entries, or NULL) receives
 * each stream's decode-cost sample.  No counterpart in the reference: it is
 * the batched scheduler's launch. */
int mdz_inflate_batch_sched(const mdz_inflate_in_t* in, mdz_inflate_status_t* out, int count,
                            int parse_hdr, const uint32_t* order, uint32_t blocks,
                            uint32_t* cursor, mdz_inflate_cost_t* cost, void* stream);

/* mdz_inflate_batch_sched without the cursor argument, for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The work-queue cursor lives in the library's own device memory (a 256-byte
 * file-scope array per device, of which the cursor uses 4 bytes), so the call
 * allocates nothing; count is limited only as for mdz_inflate_batch_sched
 * (0..INT32_MAX).  The first call on a device looks up that array's address.
 * All calls on one device share the cursor: calls that overlap on one device
 * (on different streams) race and must be serialised by the caller. */
int mdz_inflate_batch_sched_noalloc(const mdz_inflate_in_t* in, mdz_inflate_status_t* out,
                                    int count, int parse_hdr, const uint32_t* order,
                                    uint32_t blocks, mdz_inflate_cost_t* cost, void* stream);

/* Live work queue -- no counterpart in the reference; for serving loops that
 * fold Deflate streams into a decode already running on the GPU.  open
 * launches one persistent inflate kernel (blocks workgroups; 0 = a quarter
 * of the co-resident slots, leaving the rest to other kernels) on its own
 * non-blocking stream, over `capacity` items whose descriptors and statuses
 * live in pinned host memory.  submit appends n descriptors (device src/dst
 * pointers, as for mdz_inflate_batch; the source bytes must be in device
 * memory before the call) and publishes them: the running kernel starts on
 * them at once; *first is the first one's index.  statuses() points at the
 * status array: an item's status reads MDZ_ST_LIVE_PENDING until its output
 * and byte count are in memory.  Framing: MDZ_HDR_RAW or MDZ_HDR_ZLIB.  The
 * workers end when the queue is closed and drained, or after idle_ms (0 =
 * 8000) without their next item: once one has, expired() returns 1, items
 * already taken still finish, and items published later may stay pending
 * (close the queue and publish them to a new one).  close ends the
 * queue, waits for the kernel, copies the first `published` statuses to
 * `statuses` (optional) and frees everything.  While a queue is open, calls
 * that wait for the whole device (hipDeviceSynchronize, hipFree) wait for
 * its kernel too.  One thread drives a queue. */
#define MDZ_ST_LIVE_PENDING 0x7fffffff
typedef struct mdz_inflate_live mdz_inflate_live_t;
int mdz_inflate_live_open(uint32_t capacity, int parse_hdr, uint32_t blocks, uint32_t idle_ms,
                          mdz_inflate_live_t** q);
int mdz_inflate_live_submit(mdz_inflate_live_t* q, const mdz_inflate_in_t* in, uint32_t n, uint32_t* first);
const mdz_inflate_status_t* mdz_inflate_live_statuses(const mdz_inflate_live_t* q);
int mdz_inflate_live_expired(mdz_inflate_live_t* q);
int mdz_inflate_live_close(mdz_inflate_live_t* q, mdz_inflate_status_t* statuses);

/* Serving engine -- no counterpart in the reference; the admission and
 * launch loop of a decode server, native.  A request is one container in
 * host memory (keep it alive until the request is released; pinned memory,
 * mdz_pinned_alloc, makes its upload asynchronous): the lane-transposed
 * codecs give payload, blk_off, col_len (and col_map for LZSS), n_chunks,
 * chunk_bytes, elem_bytes, lane_line as for mdz_lanes_t; Deflate gives
 * n_streams zlib streams at payload + src_off[i] (src_len[i] bytes), stream
 * i decoding to output byte i*chunk_bytes.  mdz_serve_poll admits waiting
 * requests in submission order under hbm_budget (payload + output +
 * metadata bytes), uploads and launches them on n_streams streams, publishes
 * Deflate streams to a live inflate queue (mdz_inflate_live_*, live_blocks
 * and idle_ms as there) and reports up to cap finished requests: tag, the
 * first failing chunk or stream status, and the device output, which stays
 * valid until mdz_serve_release(tag).  The live queue's kernel needs a
 * hardware queue of its own (GPU_MAX_HW_QUEUES above the process's stream
 * count).  One thread drives a server; close waits for all its work. */
#define MDZ_SERVE_RLEV1   1
#define MDZ_SERVE_RLEV2   2
#define MDZ_SERVE_LZSS    3
#define MDZ_SERVE_DEFLATE 4
typedef struct {
    int             codec;         /* MDZ_SERVE_* */
    uint32_t        elem_bytes;    /* lane codecs */
    uint32_t        lane_line;     /* lane codecs */
    uint32_t        reserved;
    const void*     payload;
    uint64_t        payload_bytes;
    const uint64_t* blk_off;       /* lane codecs: n_chunks + 1 */
    const uint64_t* col_len;       /* lane codecs: 32 * n_chunks */
    const uint8_t*  col_map;       /* LZSS: 32 * n_chunks */
    uint64_t        n_chunks;
    uint64_t        chunk_bytes;   /* lane codecs: chunk size; Deflate: output bytes per stream */
    const uint64_t* src_off;       /* Deflate: n_streams */
    const uint64_t* src_len;
    uint64_t        n_streams;
    uint64_t        out_bytes;
    uint64_t        tag;           /* caller's request id, unique among unreleased requests */
} mdz_serve_request_t;
typedef struct {
    uint64_t    tag;
    int32_t     status;            /* first failing chunk's or stream's status, 0 if none */
    uint32_t    reserved;
    void*       d_out;             /* device output, out_bytes */
    uint64_t    out_bytes;
} mdz_serve_done_t;
typedef struct mdz_server mdz_server_t;
int mdz_serve_open(uint64_t hbm_budget, uint32_t n_streams, uint32_t live_blocks, uint32_t idle_ms,
                   mdz_server_t** s);
int mdz_serve_submit(mdz_server_t* s, const mdz_serve_request_t* r);
int mdz_serve_poll(mdz_server_t* s, mdz_serve_done_t* done, uint32_t cap, uint32_t* n_done);
int mdz_serve_release(mdz_server_t* s, uint64_t tag);
int mdz_serve_pending(const mdz_server_t* s, uint64_t* waiting, uint64_t* active);
int mdz_serve_close(mdz_server_t* s);

/* Compute units of the current device (sizes persistent grids). */
int mdz_device_cus(int* cus);
/* Workgroups of the persistent inflate kernel that fit on the device at once. */
int mdz_inflate_slots(uint32_t* slots);

/* deflate::decompress_gpu's host form (include/zlib/zlib.h:966-1082) over the
 * repo container (orig_compressor.py:44-125): in = concatenated zlib streams,
 * col_len_f = u64 padded stream sizes [n], blk_offset_f = [u64 data_size]
 * [u64 0][u64 cumulative padded offsets...] (n+2 entries), chunk_size = output
 * bytes per stream.  The writer zero-pads the input to a chunk multiple, so
 * every stream (the last included) inflates to chunk_size bytes; each gets a
 * full chunk_size slot on the device, as the reference reader allocates
 * chunk_size * num_blk (zlib.h:992-996), and the first data_size bytes come
 * back.  A last stream that holds only the remainder is accepted too.
 * *out (data_size bytes) is allocated here; free with mdz_host_free.
 * MDZ_ERR_CORRUPT (with *out set) when a stream fails or falls short of its
 * share of data_size; mdz_last_error() names the stream.  Stored blocks are
 * written (the reference consumes them without writing, B15). */
int mdz_deflate_decompress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                               uint64_t* out_n_bytes, const uint64_t* col_len_f,
                               uint64_t col_n_bytes, const uint64_t* blk_offset_f,
                               uint64_t blk_n_bytes, uint64_t chunk_size);

/* The cuDF container's reader, include/base_deflate.h:20-130 (`uncompress`),
 * over a buffer: file = [u64 chunk_size][u64 n][u64 size[n]][zlib streams back
 * to back] as rapids_compressor.py:222-275 writes it.  Each stream's 2-byte
 * zlib header is skipped and the rest inflated raw (base_deflate.h:66-70,
 * gpuinflate with parse_hdr 0) into a chunk_size slot.  The output is
 * (n-1)*chunk_size + the last stream's bytes_written (:97-98).  *out is
 * allocated here (free with mdz_host_free).  Returns MDZ_ERR_INVALID for a
 * header or size table that does not fit the buffer, MDZ_ERR_CORRUPT (with
 * *out set) when a stream fails or a stream before the last is short. */
int mdz_cudf_uncompress_host(const uint8_t* file, uint64_t file_bytes, uint8_t** out,
                             uint64_t* out_n_bytes);
/* The same reader from a file to a file, as `uncompress(in_file, out_file)`
 * (base_deflate.h:20); the output file is created or truncated. */
int mdz_cudf_uncompress(const char* in_file, const char* out_file);

/* ====================================================== container file ===== *
 * A self-describing file (MDZF) holding one container of any codec with all
 * its metadata: header (codec, parameters, counts, section offsets, CRC-32),
 * blk_off, col_len, col_map, Deflate stream sizes and the payload, each on a
 * 4 KiB boundary.  It replaces the reference's side files in the working
 * directory (include/rle_v1/rlev1.h:81-87,110-173; include/brle/
 * brle_trans.h:646-657) and the LZSS .meta blob (include/utils.h:46-118).
 * No counterpart in the reference otherwise. */
#define MDZ_CODEC_RLEV1   1
#define MDZ_CODEC_RLEV2   2
#define MDZ_CODEC_BRLE    3
#define MDZ_CODEC_LZSS    4
#define MDZ_CODEC_DEFLATE 5

typedef struct {
    uint32_t codec;          /* MDZ_CODEC_* */
    uint32_t elem_bytes;     /* RLE v1 DATA_TYPE / RLE v2 COMP_TYPE size */
    uint32_t lane_line;      /* RLE v1 COMP_COL_LEN / RLE v2 READ_UNIT */
    uint32_t framing;        /* Deflate: parse_hdr (MDZ_HDR_*) */
    uint64_t n_chunks;       /* chunks, or Deflate streams (< 2^32) */
    uint64_t chunk_bytes;    /* uncompressed bytes per chunk (unused for Deflate) */
    uint64_t raw_bytes;      /* uncompressed bytes in all */
    uint64_t payload_bytes;
} mdz_file_info_t;

typedef struct mdz_file_s* mdz_file_t;

/* Writes a file from host arrays: blk_off[n+1] (payload offsets, the last
 * entry as the codec defines it: RLE v2 holds the uncompressed size there),
 * col_len[32n] (RLE v1, RLE v2, LZSS), col_map[32n] (LZSS), raw_len[n]
 * (Deflate: output bytes per stream).  Tables are checked first. */
int mdz_file_write(const char* path, const mdz_file_info_t* info, const void* payload,
                   const uint64_t* blk_off, const uint64_t* col_len, const uint8_t* col_map,
                   const uint64_t* raw_len);
/* Maps a file read-only and checks header CRC, parameters and tables
 * (MDZ_ERR_INVALID, or MDZ_ERR_CORRUPT for a header CRC mismatch).  No
 * device call. */
int mdz_file_open(const char* path, mdz_file_t* file);
int mdz_file_info(mdz_file_t file, mdz_file_info_t* info);
/* Deflate files: raw_len[n] (host) = each stream's output bytes. */
int mdz_file_raw_len(mdz_file_t file, uint64_t* raw_len);
/* Decodes the whole file into dst (device, dst_bytes).  Chunked codecs write
 * chunk c at dst + c * chunk_bytes and status (device int32[n], or NULL);
 * Deflate writes stream i after streams 0..i-1 and needs status (device
 * mdz_inflate_status_t[n]) and dst_bytes >= raw_bytes.  The payload moves to
 * HBM in pieces of about stage_bytes (0 = 64 MiB) through two pinned buffers;
 * the host reads one from the file while the other is copied on a separate
 * stream (the call returns once the last piece is read), and each
 * piece's chunks decode on `stream` once their bytes have landed.  The first
 * call allocates the device copies of payload and tables and the staging
 * buffers; they stay with the handle for later calls. */
int mdz_file_decode(mdz_file_t file, void* dst, uint64_t dst_bytes, void* status,
                    uint64_t stage_bytes, void* stream);
/* Waits for the handle's copies, frees its device and host memory, unmaps.
 * A handle's staging buffers and device copies are its own: calls on one
 * handle must not overlap (use one handle per thread). */
int mdz_file_close(mdz_file_t file);

/* ================================================ learned cost model ===== *
 * Scheduler-side MLP (features -> 32 -> 32 -> 1, ReLU, fp32) predicting
 * log2(decode cycles) of a Deflate stream, trained on the decoder's own cost
 * log.  It only orders work.  No counterpart in the reference.
 * Parameters (floats): W1[F][H] b1[H] W2[H][H] b2[H] w3[H] b3. */
#define MDZ_MLP_FEATURES 11
#define MDZ_MLP_HIDDEN 32
#define MDZ_MLP_PARAMS (MDZ_MLP_FEATURES * MDZ_MLP_HIDDEN + MDZ_MLP_HIDDEN + \
                        MDZ_MLP_HIDDEN * MDZ_MLP_HIDDEN + MDZ_MLP_HIDDEN + MDZ_MLP_HIDDEN + 1)

/* x[count][F] (device): log2(1+src), log2(1+dst), their difference and the
 * first block type one-hot (stored, fixed, dynamic), read from the stream. */
int mdz_cost_features(const mdz_inflate_in_t* in, int count, int parse_hdr, float* x,
                      void* stream);
/* y[n] = model(x[n][F]) */
int mdz_mlp_forward(const float* params, const float* x, int n, float* y, void* stream);
/* grad[P] = d/dparams mean((y - t)^2); *loss_sum (device, optional) = sum (y - t)^2 */
int mdz_mlp_grad(const float* params, const float* x, const float* t, int n, float* grad,
                 float* loss_sum, void* stream);
/* One Adam step (beta 0.9/0.999) with bias correction for step >= 1; clears grad. */
int mdz_mlp_adam(float* params, float* grad, float* m, float* v, float lr, uint32_t step,
                 void* stream);

/* ============================================ batched persistent decode ===== *
 * One launch decodes the chunks of an RLE v2 container and a byte-RLE
 * container (config 2's pair) through a device work queue: a grid of
 * `blocks` persistent workgroups (0 = the co-resident maximum) claims items
 * from `cursor` (device u32, reset by the call).  items (device u32 array,
 * n_items entries, codec << 31 | chunk, codec 0 = RLE v2, 1 = byte RLE) is
 * the scheduler's order; NULL = every RLE v2 chunk, then every byte-RLE
 * chunk.  Results and statuses equal the per-codec entries'.  No counterpart
 * in the reference (its decoders launch one block per chunk). */
typedef struct {
    mdz_lanes_t rlev2;              /* n_chunks = 0: no RLE v2 part */
    void* rlev2_dst;
    uint64_t rlev2_dst_bytes;
    int32_t* rlev2_status;
    uint32_t rlev2_max_chunk;       /* largest chunk payload (LDS sizing), 0 = LDS maximum */
    uint32_t brle_max_chunk;        /* 0 = the byte-RLE worst case */
    uint32_t rlev2_max_part;        /* as mdz_lanes_tune_t.max_part_bytes: nonzero decodes RLE v2
                                       in lane quarters (256-thread workgroups throughout) */
    uint32_t rlev2_window;          /* as mdz_lanes_tune_t.rlev2_window (0 = 16) */
    const uint64_t* brle_ptr;
    const void* brle_data;
    uint64_t brle_n_chunks;         /* 0: no byte-RLE part */
    uint64_t brle_chunk_bytes;
    void* brle_dst;
    uint64_t brle_dst_bytes;
    int32_t* brle_status;
    const uint32_t* items;
    uint32_t n_items;
    uint32_t blocks;
    uint32_t* cursor;
} mdz_batch_t;

int mdz_batch_submit(const mdz_batch_t* batch, void* stream);

/* ==================================================== captured pipelines ===== *
 * The mdz calls a caller issues on one stream between mdz_graph_begin and
 * mdz_graph_end are captured into a hipGraph and instantiated; each
 * mdz_graph_launch replays the whole pipeline (e.g. a batch's status fill,
 * cursor reset, persistent kernel and status finish) with one launch.  No
 * counterpart in the reference, which issues each launch from the host; it
 * is for serving loops that replay the same batch shape.
 *  - `stream` must be a created stream (the legacy default stream cannot be
 *    captured); capture uses hipStreamCaptureModeRelaxed.
 *  - The graph keeps the pointers and sizes the captured calls were given:
 *    every buffer they name must stay allocated while the graph is used.
 *  - mdz decode pipelines record only kernel nodes (no memset nodes), so a
 *    graph may be replayed any number of times.
 *  - If a call fails during capture, end the capture anyway; mdz_graph_end
 *    then returns the error and no graph. */
typedef struct mdz_graph_s* mdz_graph_t;
int mdz_graph_begin(void* stream);
int mdz_graph_end(void* stream, mdz_graph_t* graph);
int mdz_graph_launch(mdz_graph_t graph, void* stream);
int mdz_graph_destroy(mdz_graph_t graph);
/* Node counts of a captured graph: all nodes, and those that are kernels. */
int mdz_graph_nodes(mdz_graph_t graph, uint64_t* nodes, uint64_t* kernels);

#ifdef __cplusplus
}
#endif
#endif
