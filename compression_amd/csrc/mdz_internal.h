// Internal (C++) interfaces between the C-ABI layer (api.hip) and the
// per-codec launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mdz.h"

namespace mdz {

// Records a failed HIP call for mdz_last_error(); returns MDZ_OK/MDZ_ERR_HIP.
int record_hip(hipError_t e, const char* what);
// Formats a message for mdz_last_error() and returns rc.
int set_error(int rc, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Raises a kernel's dynamic-LDS limit to at least `bytes` (gfx950 allows
// 160 KiB per workgroup, static LDS included).  Thread-safe; each
// kernel is raised once to the largest size requested so far.
hipError_t set_dynamic_lds(const void* kernel, size_t bytes);

int dev_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s);

int rlev1_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                     const mdz_lanes_tune_t* tune, hipStream_t s);

int rlev2_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                     const mdz_lanes_tune_t* tune, hipStream_t s);

int brle_decompress(const uint64_t* ptr, uint64_t n_chunks, uint64_t chunk_bytes, const void* data,
                    void* dst, uint64_t dst_bytes, int32_t* status, uint32_t max_chunk_bytes,
                    hipStream_t s);

int rlev1_compress_bound(uint64_t in_bytes, uint64_t chunk, uint32_t eb, uint32_t line, uint64_t* n_chunks,
                         uint64_t* work_bytes, uint64_t* payload_cap);
int rlev1_compress(const void* in, uint64_t in_bytes, uint64_t chunk, uint32_t eb, uint32_t line, void* work,
                   void* payload, uint64_t* col_len, uint64_t* blk_off, hipStream_t s);
int rlev2_compress_bound(uint64_t in_bytes, uint64_t chunk, uint32_t eb, uint32_t ru, uint64_t* n_chunks,
                         uint64_t* work_bytes, uint64_t* payload_cap);
int rlev2_compress(const void* in, uint64_t in_bytes, uint64_t chunk, uint32_t eb, uint32_t ru, void* work,
                   void* payload, uint64_t* col_len, uint64_t* blk_off, uint64_t* payload_bytes, hipStream_t s);
int lzss_compress_bound(uint64_t in_bytes, uint64_t chunk, uint64_t* n_chunks, uint64_t* work_bytes,
                        uint64_t* payload_cap);
int lzss_compress(const void* in, uint64_t in_bytes, uint64_t chunk, void* work, void* payload, uint64_t* col_len,
                  uint8_t* col_map, uint64_t* blk_off, hipStream_t s);
int brle_compress_bound(uint64_t in_bytes, uint64_t chunk, uint64_t* n_chunks, uint64_t* work_bytes,
                        uint64_t* data_cap);
int brle_compress(const void* in, uint64_t in_bytes, uint64_t chunk, void* work, void* data, uint64_t* ptr,
                  hipStream_t s);

int inflate_batch(const mdz_inflate_in_t* in, mdz_inflate_status_t* out, uint32_t count,
                  int parse_hdr, const uint32_t* order, uint32_t persistent_blocks,
                  uint32_t* cursor, mdz_inflate_cost_t* cost, hipStream_t s);
int inflate_slots(uint32_t* slots);
int inflate_live_launch(const mdz_inflate_in_t* in, mdz_inflate_status_t* out, uint32_t cap, int parse_hdr,
                        uint32_t blocks, uint32_t* cursor, uint32_t* live, hipStream_t s);
int copy_uncompressed(const mdz_inflate_in_t* in, uint32_t count, hipStream_t s);

// parse_hdr: MDZ_HDR_RAW / ZLIB / GZIP, optionally | MDZ_HDR_VERIFY (framed streams only)
inline bool inflate_hdr_ok(int h) {
    const int f = h & 3;
    return (h & ~(3 | MDZ_HDR_VERIFY)) == 0 && f != 3 && (f != 0 || !(h & MDZ_HDR_VERIFY));
}

int mlp_features(const mdz_inflate_in_t* in, uint32_t n, int parse_hdr, float* x, hipStream_t s);
int mlp_forward(const float* params, const float* x, uint32_t n, float* y, hipStream_t s);
int mlp_grad(const float* params, const float* x, const float* t, uint32_t n, float* grad,
             float* loss_sum, hipStream_t s);
int mlp_adam(float* params, float* grad, float* m, float* v, float lr, uint32_t step, hipStream_t s);

int batch_submit(const mdz_batch_t* b, hipStream_t s);

int lzss_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                    hipStream_t s);
int brle_trans_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                          hipStream_t s);
int lzss_plain_decompress(const void* data, uint64_t data_bytes, const uint64_t* lens, uint64_t n_chunks,
                          uint64_t chunk, void* dst, uint64_t dst_bytes, int32_t* status, hipStream_t s);
int rlev2_plain_decompress(const void* data, uint64_t data_bytes, const uint64_t* ptr, uint64_t n_chunks,
                           uint64_t chunk, void* dst, uint64_t dst_bytes, int32_t* status, hipStream_t s);

}  // namespace mdz
