// mdz C ABI: device plumbing, per-codec entry points and the host-pointer
// wrappers that keep the reference's signature shapes.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <stdarg.h>
#include <new>
#include <vector>
#include <unordered_map>

#include "mdz_internal.h"

namespace mdz {
thread_local char g_last_error[256] = "";
int record_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return MDZ_OK;
    snprintf(g_last_error, sizeof g_last_error, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    return MDZ_ERR_HIP;
}

int set_error(int rc, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
    return rc;
}

hipError_t set_dynamic_lds(const void* kernel, size_t bytes) {
    static std::mutex mu;
    static std::unordered_map<const void*, size_t> done;
    std::lock_guard<std::mutex> g(mu);
    size_t& have = done[kernel];
    if (have >= bytes) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    have = bytes;
    return hipSuccess;
}
}  // namespace mdz

namespace {

inline int herr(hipError_t e) { return mdz::record_hip(e, "hip"); }
inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Scoped device buffer for the host wrappers.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(uint64_t n) { return hipMalloc(&p, n ? n : 1) == hipSuccess ? MDZ_OK : MDZ_ERR_NOMEM; }
};

int read_file(const char* path, uint8_t** data, uint64_t* n) {
    FILE* f = fopen(path, "rb");
    if (!f) return MDZ_ERR_INVALID;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0) { fclose(f); return MDZ_ERR_INVALID; }
    *data = (uint8_t*)malloc(sz ? (size_t)sz : 1);
    if (!*data) { fclose(f); return MDZ_ERR_NOMEM; }
    size_t got = fread(*data, 1, (size_t)sz, f);
    fclose(f);
    if (got != (size_t)sz) { free(*data); *data = nullptr; return MDZ_ERR_INVALID; }
    *n = (uint64_t)sz;
    return MDZ_OK;
}

// Synchronises, copies `raw` output bytes and the per-chunk status to the
// host; *out is malloc'ed.  MDZ_ERR_CORRUPT if any chunk status is nonzero.
int finish_host(const void* d_out, uint64_t raw, const int32_t* d_st, uint64_t n, uint8_t** out,
                uint64_t* out_n_bytes) {
    if (hipDeviceSynchronize() != hipSuccess) return MDZ_ERR_HIP;
    int32_t* st = (int32_t*)malloc(n * 4 + 4);
    uint8_t* o = (uint8_t*)malloc(raw ? raw : 1);
    if (!st || !o) { free(st); free(o); return MDZ_ERR_NOMEM; }
    if (hipMemcpy(st, d_st, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        (raw && hipMemcpy(o, d_out, raw, hipMemcpyDeviceToHost) != hipSuccess)) {
        free(st); free(o);
        return MDZ_ERR_HIP;
    }
    int rc = MDZ_OK;
    for (uint64_t c = 0; c < n; ++c)
        if (st[c] != MDZ_ST_OK) {
            rc = mdz::set_error(MDZ_ERR_CORRUPT, "chunk %llu: status %d", (unsigned long long)c,
                                (int)st[c]);
            break;
        }
    free(st);
    *out = o;
    *out_n_bytes = raw;
    return rc;
}

}  // namespace

extern "C" {

const char* mdz_version(void) { return "mdz 0.1.0 gfx950"; }
const char* mdz_last_error(void) { return mdz::g_last_error; }

int mdz_device_count(int* count) { return herr(hipGetDeviceCount(count)); }
int mdz_device_cus(int* cus) {
    if (!cus) return MDZ_ERR_INVALID;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return herr(e);
    return herr(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
}
int mdz_dev_malloc(void** ptr, uint64_t bytes) {
    return hipMalloc(ptr, bytes ? bytes : 1) == hipSuccess ? MDZ_OK : MDZ_ERR_NOMEM;
}
int mdz_dev_free(void* ptr) { return herr(hipFree(ptr)); }
int mdz_pinned_alloc(void** ptr, uint64_t bytes) {
    return hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) == hipSuccess ? MDZ_OK : MDZ_ERR_NOMEM;
}
int mdz_pinned_free(void* ptr) { return herr(hipHostFree(ptr)); }
int mdz_memcpy_htod(void* dst, const void* src, uint64_t bytes, void* stream) {
    return herr(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
}
int mdz_memcpy_dtoh(void* dst, const void* src, uint64_t bytes, void* stream) {
    return herr(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
}
int mdz_memset_d8(void* dst, int value, uint64_t bytes, void* stream) {
    return herr(hipMemsetAsync(dst, value, bytes, S(stream)));
}
int mdz_dev_copy(void* dst, const void* src, uint64_t bytes, void* stream) {
    return mdz::dev_copy(dst, src, bytes, S(stream));
}
int mdz_stream_create(void** stream) {
    hipStream_t s;
    int rc = herr(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (rc == MDZ_OK) *stream = s;
    return rc;
}
int mdz_stream_destroy(void* stream) { return herr(hipStreamDestroy(S(stream))); }
int mdz_stream_sync(void* stream) { return herr(hipStreamSynchronize(S(stream))); }
int mdz_device_sync(void) { return herr(hipDeviceSynchronize()); }
int mdz_event_create(void** event) {
    hipEvent_t e;
    int rc = herr(hipEventCreate(&e));
    if (rc == MDZ_OK) *event = e;
    return rc;
}
int mdz_event_destroy(void* event) { return herr(hipEventDestroy((hipEvent_t)event)); }
int mdz_event_record(void* event, void* stream) { return herr(hipEventRecord((hipEvent_t)event, S(stream))); }
int mdz_event_sync(void* event) { return herr(hipEventSynchronize((hipEvent_t)event)); }
int mdz_event_query(void* event) {
    const hipError_t e = hipEventQuery((hipEvent_t)event);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) { (void)hipGetLastError(); return 0; }
    return herr(e);
}
int mdz_event_elapsed_ms(void* start, void* stop, float* ms) {
    return herr(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
}
void mdz_host_free(void* p) { free(p); }

// ------------------------------------------------------------------ RLE v1
int mdz_rlev1_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                         void* stream) {
    return mdz::rlev1_decompress(in, dst, dst_bytes, status, nullptr, S(stream));
}

int mdz_rlev1_decompress_tuned(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes,
                               int32_t* status, const mdz_lanes_tune_t* tune, void* stream) {
    return mdz::rlev1_decompress(in, dst, dst_bytes, status, tune, S(stream));
}

int mdz_rlev1_decompress_host(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                              uint64_t* out_n_bytes, int comp_col_len, uint64_t chunk_size,
                              uint32_t elem_bytes, const uint64_t* col_len,
                              const uint64_t* blk_off, uint64_t n_chunks) {
    if (!in || !out || !out_n_bytes || !col_len || !blk_off || in_n_bytes < 8 || comp_col_len <= 0)
        return MDZ_ERR_INVALID;
    uint64_t raw;
    memcpy(&raw, in, 8);
    const uint64_t pay_bytes = in_n_bytes - 8;
    if (blk_off[n_chunks] > pay_bytes || raw > n_chunks * chunk_size) return MDZ_ERR_INVALID;
    DevBuf d_pay, d_col, d_off, d_out, d_st;
    int rc;
    if ((rc = d_pay.alloc(pay_bytes)) || (rc = d_col.alloc(32 * n_chunks * 8)) ||
        (rc = d_off.alloc((n_chunks + 1) * 8)) || (rc = d_out.alloc(raw)) ||
        (rc = d_st.alloc(n_chunks * 4)))
        return rc;
    if (hipMemcpy(d_pay.p, in + 8, pay_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_col.p, col_len, 32 * n_chunks * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_off.p, blk_off, (n_chunks + 1) * 8, hipMemcpyHostToDevice) != hipSuccess)
        return MDZ_ERR_HIP;
    mdz_lanes_t L = {d_pay.p, (const uint64_t*)d_off.p, (const uint64_t*)d_col.p, nullptr,
                     n_chunks, chunk_size, elem_bytes, (uint32_t)comp_col_len, pay_bytes};
    rc = mdz::rlev1_decompress(&L, d_out.p, raw, (int32_t*)d_st.p, nullptr, nullptr);
    if (rc) return rc;
    return finish_host(d_out.p, raw, (const int32_t*)d_st.p, n_chunks, out, out_n_bytes);
}

int mdz_rle_v1_decompress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                              uint64_t* out_n_bytes, int comp_col_len, uint64_t chunk_size,
                              uint32_t elem_bytes) {
    uint8_t *col = nullptr, *off = nullptr;
    uint64_t ncol = 0, noff = 0;
    int rc = read_file("./input_data/rle_v1_col_len.bin", &col, &ncol);
    if (rc == MDZ_OK) rc = read_file("./input_data/rle_v1_blk_offset.bin", &off, &noff);
    if (rc == MDZ_OK) {
        uint64_t n = noff / 8 - 1;
        if (noff < 16 || ncol < 32 * n * 8) rc = MDZ_ERR_INVALID;
        else
            rc = mdz_rlev1_decompress_host(in, out, in_n_bytes, out_n_bytes, comp_col_len,
                                           chunk_size, elem_bytes, (const uint64_t*)col,
                                           (const uint64_t*)off, n);
    }
    free(col);
    free(off);
    return rc;
}

int mdz_rlev1_compress_bound(uint64_t in_bytes, uint64_t chunk_bytes, uint32_t elem_bytes, uint32_t comp_col_len,
                             uint64_t* n_chunks, uint64_t* work_bytes, uint64_t* payload_cap) {
    return mdz::rlev1_compress_bound(in_bytes, chunk_bytes, elem_bytes, comp_col_len, n_chunks, work_bytes,
                                     payload_cap);
}

int mdz_rlev1_compress(const void* in, uint64_t in_bytes, uint64_t chunk_bytes, uint32_t elem_bytes,
                       uint32_t comp_col_len, void* work, void* payload, uint64_t* col_len, uint64_t* blk_off,
                       void* stream) {
    return mdz::rlev1_compress(in, in_bytes, chunk_bytes, elem_bytes, comp_col_len, work, payload, col_len,
                               blk_off, S(stream));
}

int mdz_rle_v1_compress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes, uint64_t* out_n_bytes,
                            int comp_col_len, uint64_t chunk_size, uint32_t elem_bytes) {
    if (!in || !out || !out_n_bytes || comp_col_len <= 0) return MDZ_ERR_INVALID;
    uint64_t n = 0, wb = 0, cap = 0;
    int rc = mdz::rlev1_compress_bound(in_n_bytes, chunk_size, elem_bytes, (uint32_t)comp_col_len, &n, &wb, &cap);
    if (rc) return rc;
    DevBuf d_in, d_work, d_pay, d_col, d_off;
    if ((rc = d_in.alloc(in_n_bytes)) || (rc = d_work.alloc(wb)) || (rc = d_pay.alloc(cap)) ||
        (rc = d_col.alloc(32 * 8 * n)) || (rc = d_off.alloc(8 * (n + 1))))
        return rc;
    if (hipMemcpy(d_in.p, in, in_n_bytes, hipMemcpyHostToDevice) != hipSuccess) return MDZ_ERR_HIP;
    if ((rc = mdz::rlev1_compress(d_in.p, in_n_bytes, chunk_size, elem_bytes, (uint32_t)comp_col_len, d_work.p,
                                  d_pay.p, (uint64_t*)d_col.p, (uint64_t*)d_off.p, nullptr)))
        return rc;
    uint64_t* col = (uint64_t*)malloc(32 * 8 * n);
    uint64_t* off = (uint64_t*)malloc(8 * (n + 1));
    uint8_t* o = nullptr;
    rc = (col && off) ? MDZ_OK : MDZ_ERR_NOMEM;
    if (rc == MDZ_OK && (hipMemcpy(col, d_col.p, 32 * 8 * n, hipMemcpyDeviceToHost) != hipSuccess ||
                         hipMemcpy(off, d_off.p, 8 * (n + 1), hipMemcpyDeviceToHost) != hipSuccess))
        rc = MDZ_ERR_HIP;
    // [u64 in_n_bytes][payload]; col_len and blk_off go to the side files the
    // reference writes (rlev1.h:85-93) and mdz_rle_v1_decompress_gpu reads
    if (rc == MDZ_OK && !(o = (uint8_t*)malloc(off[n] + 8))) rc = MDZ_ERR_NOMEM;
    if (rc == MDZ_OK) {
        memcpy(o, &in_n_bytes, 8);
        if (off[n] && hipMemcpy(o + 8, d_pay.p, off[n], hipMemcpyDeviceToHost) != hipSuccess) rc = MDZ_ERR_HIP;
    }
    if (rc == MDZ_OK) {
        FILE* fc = fopen("./input_data/rle_v1_col_len.bin", "wb");
        FILE* fo = fopen("./input_data/rle_v1_blk_offset.bin", "wb");
        if (!fc || !fo || fwrite(col, 8, 32 * n, fc) != 32 * n || fwrite(off, 8, n + 1, fo) != n + 1)
            rc = MDZ_ERR_INVALID;
        if (fc) fclose(fc);
        if (fo) fclose(fo);
    }
    const uint64_t total = rc == MDZ_OK ? off[n] + 8 : 0;
    free(col);
    free(off);
    if (rc != MDZ_OK) {
        free(o);
        return rc;
    }
    *out = o;
    *out_n_bytes = total;
    return MDZ_OK;
}

// ------------------------------------------------------------------ RLE v2
int mdz_rlev2_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                         void* stream) {
    return mdz::rlev2_decompress(in, dst, dst_bytes, status, nullptr, S(stream));
}

int mdz_rlev2_decompress_tuned(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes,
                               int32_t* status, const mdz_lanes_tune_t* tune, void* stream) {
    return mdz::rlev2_decompress(in, dst, dst_bytes, status, tune, S(stream));
}

int mdz_rlev2_decompress_host(const uint8_t* in, uint64_t in_n_bytes, uint64_t n_chunks,
                              const uint64_t* blk_off, const uint64_t* col_len, uint8_t** out,
                              uint64_t* out_n_bytes, uint64_t chunk_size, uint32_t read_unit,
                              uint32_t elem_bytes) {
    if (!in || !out || !out_n_bytes || !col_len || !blk_off || !n_chunks) return MDZ_ERR_INVALID;
    const uint64_t raw = blk_off[n_chunks];
    if (raw > n_chunks * chunk_size) return MDZ_ERR_INVALID;
    DevBuf d_pay, d_col, d_off, d_out, d_st;
    int rc;
    if ((rc = d_pay.alloc(in_n_bytes + 16)) || (rc = d_col.alloc(32 * n_chunks * 8)) ||
        (rc = d_off.alloc((n_chunks + 1) * 8)) || (rc = d_out.alloc(raw)) ||
        (rc = d_st.alloc(n_chunks * 4)))
        return rc;
    if (hipMemcpy(d_pay.p, in, in_n_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_col.p, col_len, 32 * n_chunks * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_off.p, blk_off, (n_chunks + 1) * 8, hipMemcpyHostToDevice) != hipSuccess)
        return MDZ_ERR_HIP;
    mdz_lanes_t L = {d_pay.p, (const uint64_t*)d_off.p, (const uint64_t*)d_col.p, nullptr,
                     n_chunks, chunk_size, elem_bytes, read_unit, in_n_bytes};
    rc = mdz::rlev2_decompress(&L, d_out.p, raw, (int32_t*)d_st.p, nullptr, nullptr);
    if (rc) return rc;
    return finish_host(d_out.p, raw, (const int32_t*)d_st.p, n_chunks, out, out_n_bytes);
}

int mdz_rlev2_compress_bound(uint64_t in_bytes, uint64_t chunk_bytes, uint32_t elem_bytes, uint32_t read_unit,
                             uint64_t* n_chunks, uint64_t* work_bytes, uint64_t* payload_cap) {
    return mdz::rlev2_compress_bound(in_bytes, chunk_bytes, elem_bytes, read_unit, n_chunks, work_bytes,
                                     payload_cap);
}

int mdz_rlev2_compress(const void* in, uint64_t in_bytes, uint64_t chunk_bytes, uint32_t elem_bytes,
                       uint32_t read_unit, void* work, void* payload, uint64_t* col_len, uint64_t* blk_off,
                       uint64_t* payload_bytes, void* stream) {
    return mdz::rlev2_compress(in, in_bytes, chunk_bytes, elem_bytes, read_unit, work, payload, col_len, blk_off,
                               payload_bytes, S(stream));
}

// ---------------------------------------------------------------- byte RLE
int mdz_brle_decompress(const uint64_t* ptr, uint64_t n_chunks, uint64_t chunk_bytes,
                        const void* data, void* dst, uint64_t dst_bytes, int32_t* status,
                        void* stream) {
    return mdz::brle_decompress(ptr, n_chunks, chunk_bytes, data, dst, dst_bytes, status, 0,
                                S(stream));
}

int mdz_brle_decompress_tuned(const uint64_t* ptr, uint64_t n_chunks, uint64_t chunk_bytes,
                              const void* data, void* dst, uint64_t dst_bytes, int32_t* status,
                              uint32_t max_chunk_bytes, void* stream) {
    return mdz::brle_decompress(ptr, n_chunks, chunk_bytes, data, dst, dst_bytes, status,
                                max_chunk_bytes, S(stream));
}

int mdz_brle_decompress_host(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                             uint64_t* out_n_bytes, uint64_t chunk_bytes) {
    if (!in || !out || !out_n_bytes || in_n_bytes < 4 || chunk_bytes == 0) return MDZ_ERR_INVALID;
    uint32_t n_ptr;
    memcpy(&n_ptr, in, 4);
    if (n_ptr < 2) return MDZ_ERR_INVALID;
    const uint64_t n = n_ptr - 1, hdr = 4 + 8ull * n_ptr;
    if (in_n_bytes < hdr + 8) return MDZ_ERR_INVALID;
    uint64_t* ptr = (uint64_t*)malloc(8ull * n_ptr);
    if (!ptr) return MDZ_ERR_NOMEM;
    memcpy(ptr, in + 4, 8ull * n_ptr);
    uint64_t raw;
    memcpy(&raw, in + hdr, 8);
    const uint8_t* data = in + hdr + 8;
    const uint64_t data_bytes = in_n_bytes - hdr - 8;
    uint64_t max_chunk = 0;
    int rc = MDZ_OK;
    for (uint64_t c = 0; c < n; ++c) {
        if (ptr[c + 1] < ptr[c] || ptr[c + 1] > data_bytes) { rc = MDZ_ERR_INVALID; break; }
        if (ptr[c + 1] - ptr[c] > max_chunk) max_chunk = ptr[c + 1] - ptr[c];
    }
    if (rc == MDZ_OK && (raw > n * chunk_bytes || max_chunk > 0xffffffffull)) rc = MDZ_ERR_INVALID;
    DevBuf d_ptr, d_data, d_out, d_st;
    if (rc == MDZ_OK &&
        ((rc = d_ptr.alloc(8ull * n_ptr)) || (rc = d_data.alloc(data_bytes + 16)) ||
         (rc = d_out.alloc(raw)) || (rc = d_st.alloc(4 * n)))) {
    }
    if (rc == MDZ_OK &&
        (hipMemcpy(d_ptr.p, ptr, 8ull * n_ptr, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(d_data.p, data, data_bytes, hipMemcpyHostToDevice) != hipSuccess))
        rc = MDZ_ERR_HIP;
    free(ptr);
    if (rc) return rc;
    rc = mdz::brle_decompress((const uint64_t*)d_ptr.p, n, chunk_bytes, d_data.p, d_out.p, raw,
                              (int32_t*)d_st.p, (uint32_t)max_chunk, nullptr);
    if (rc) return rc;
    return finish_host(d_out.p, raw, (const int32_t*)d_st.p, n, out, out_n_bytes);
}

int mdz_brle_decompress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                            uint64_t* out_n_bytes) {
    return mdz_brle_decompress_host(in, out, in_n_bytes, out_n_bytes, 256 * 1024);
}

int mdz_brle_compress_bound(uint64_t in_bytes, uint64_t chunk_bytes, uint64_t* n_chunks,
                            uint64_t* work_bytes, uint64_t* data_cap) {
    return mdz::brle_compress_bound(in_bytes, chunk_bytes, n_chunks, work_bytes, data_cap);
}

int mdz_brle_compress(const void* in, uint64_t in_bytes, uint64_t chunk_bytes, void* work, void* data,
                      uint64_t* ptr, void* stream) {
    return mdz::brle_compress(in, in_bytes, chunk_bytes, work, data, ptr, S(stream));
}

int mdz_brle_compress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes, uint64_t* out_n_bytes) {
    if (!in || !out || !out_n_bytes) return MDZ_ERR_INVALID;
    const uint64_t chunk = 256 * 1024;
    uint64_t n = 0, wb = 0, cap = 0;
    int rc = mdz::brle_compress_bound(in_n_bytes, chunk, &n, &wb, &cap);
    if (rc) return rc;
    if (n + 1 > 0xFFFFFFFFull) return MDZ_ERR_INVALID;
    DevBuf d_in, d_work, d_data, d_ptr;
    if ((rc = d_in.alloc(in_n_bytes)) || (rc = d_work.alloc(wb)) || (rc = d_data.alloc(cap)) ||
        (rc = d_ptr.alloc(8 * (n + 1))))
        return rc;
    if (hipMemcpy(d_in.p, in, in_n_bytes, hipMemcpyHostToDevice) != hipSuccess) return MDZ_ERR_HIP;
    if ((rc = mdz::brle_compress(d_in.p, in_n_bytes, chunk, d_work.p, d_data.p, (uint64_t*)d_ptr.p, nullptr)))
        return rc;
    // [u32 n + 1][u64 ptr[n + 1]][u64 in_n_bytes][data], as brle.h:278-292 lays it out
    uint64_t* ptr = (uint64_t*)malloc(8 * (n + 1));
    if (!ptr) return MDZ_ERR_NOMEM;
    if (hipMemcpy(ptr, d_ptr.p, 8 * (n + 1), hipMemcpyDeviceToHost) != hipSuccess) {
        free(ptr);
        return MDZ_ERR_HIP;
    }
    const uint64_t hdr = 4 + 8 * (n + 1) + 8, total = hdr + ptr[n];
    uint8_t* o = (uint8_t*)malloc(total);
    if (!o) {
        free(ptr);
        return MDZ_ERR_NOMEM;
    }
    const uint32_t n_ptr = (uint32_t)(n + 1);
    memcpy(o, &n_ptr, 4);
    memcpy(o + 4, ptr, 8 * (n + 1));
    memcpy(o + 4 + 8 * (n + 1), &in_n_bytes, 8);
    const hipError_t e = ptr[n] ? hipMemcpy(o + hdr, d_data.p, ptr[n], hipMemcpyDeviceToHost) : hipSuccess;
    free(ptr);
    if (e != hipSuccess) {
        free(o);
        return MDZ_ERR_HIP;
    }
    *out = o;
    *out_n_bytes = total;
    return MDZ_OK;
}

// -------------------------------------------------------------------- LZSS
int mdz_lzss_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                        void* stream) {
    return mdz::lzss_decompress(in, dst, dst_bytes, status, S(stream));
}

int mdz_lzss_decompress_host(const uint8_t* in, uint64_t in_n_bytes,
                             const uint8_t* const* in_metadata, const uint64_t* in_metadata_lens,
                             uint64_t in_metadata_num, uint8_t** out, uint64_t* out_n_bytes) {
    if (in_metadata_num < 4 || !in_metadata || !in_metadata_lens) return -1;
    if (!in || !out || !out_n_bytes) return MDZ_ERR_INVALID;
    const uint64_t* lens = in_metadata_lens;
    if (lens[0] != 8 || lens[1] < 16 || lens[1] % 8) return MDZ_ERR_INVALID;
    const uint64_t n = lens[1] / 8 - 1;
    if (lens[2] != 32 * n * 8 || lens[3] != 32 * n) return MDZ_ERR_INVALID;
    uint64_t chunk;
    memcpy(&chunk, in_metadata[0], 8);
    const uint64_t raw = n * chunk;
    DevBuf d_pay, d_col, d_off, d_map, d_out, d_st;
    int rc;
    if ((rc = d_pay.alloc(in_n_bytes + 16)) || (rc = d_col.alloc(lens[2])) ||
        (rc = d_off.alloc(lens[1])) || (rc = d_map.alloc(lens[3])) || (rc = d_out.alloc(raw)) ||
        (rc = d_st.alloc(n * 4)))
        return rc;
    if (hipMemcpy(d_pay.p, in, in_n_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_off.p, in_metadata[1], lens[1], hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_col.p, in_metadata[2], lens[2], hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_map.p, in_metadata[3], lens[3], hipMemcpyHostToDevice) != hipSuccess)
        return MDZ_ERR_HIP;
    mdz_lanes_t L = {d_pay.p, (const uint64_t*)d_off.p, (const uint64_t*)d_col.p,
                     (const uint8_t*)d_map.p, n, chunk, 1, 4, in_n_bytes};
    rc = mdz::lzss_decompress(&L, d_out.p, raw, (int32_t*)d_st.p, nullptr);
    if (rc) return rc;
    return finish_host(d_out.p, raw, (const int32_t*)d_st.p, n, out, out_n_bytes);
}

int mdz_brle_trans_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                              void* stream) {
    return mdz::brle_trans_decompress(in, dst, dst_bytes, status, S(stream));
}

int mdz_lzss_plain_decompress(const void* data, uint64_t data_bytes, const uint64_t* lens, uint64_t n_chunks,
                              uint64_t chunk_bytes, void* dst, uint64_t dst_bytes, int32_t* status, void* stream) {
    return mdz::lzss_plain_decompress(data, data_bytes, lens, n_chunks, chunk_bytes, dst, dst_bytes, status,
                                      S(stream));
}

int mdz_rlev2_plain_decompress(const void* data, uint64_t data_bytes, const uint64_t* ptr, uint64_t n_chunks,
                               uint64_t chunk_bytes, void* dst, uint64_t dst_bytes, int32_t* status,
                               void* stream) {
    return mdz::rlev2_plain_decompress(data, data_bytes, ptr, n_chunks, chunk_bytes, dst, dst_bytes, status,
                                       S(stream));
}

int mdz_rlev2_plain_decompress_host(const uint8_t* in, uint64_t in_n_bytes, uint64_t chunk_bytes, uint8_t** out,
                                    uint64_t* out_n_bytes) {
    if (!in || !out || !out_n_bytes || in_n_bytes < 4 || chunk_bytes == 0 || chunk_bytes % 8)
        return MDZ_ERR_INVALID;
    uint32_t n_ptr;
    memcpy(&n_ptr, in, 4);
    if (n_ptr < 2) return MDZ_ERR_INVALID;
    const uint64_t n = n_ptr - 1, hdr = 4 + 8ull * n_ptr;
    if (in_n_bytes < hdr + 8) return MDZ_ERR_INVALID;
    uint64_t* ptr = (uint64_t*)malloc(8ull * n_ptr);
    if (!ptr) return MDZ_ERR_NOMEM;
    memcpy(ptr, in + 4, 8ull * n_ptr);
    uint64_t raw;
    memcpy(&raw, in + hdr, 8);
    const uint8_t* data = in + hdr + 8;
    const uint64_t data_bytes = in_n_bytes - hdr - 8;
    int rc = MDZ_OK;
    for (uint64_t c = 0; c < n && rc == MDZ_OK; ++c)
        if (ptr[c + 1] < ptr[c] || ptr[c + 1] > data_bytes) rc = MDZ_ERR_INVALID;
    if (rc == MDZ_OK && raw > n * chunk_bytes) rc = MDZ_ERR_INVALID;
    DevBuf d_ptr, d_data, d_out, d_st;
    if (rc == MDZ_OK &&
        ((rc = d_ptr.alloc(8ull * n_ptr)) || (rc = d_data.alloc(data_bytes + 16)) ||
         (rc = d_out.alloc(raw + 8)) || (rc = d_st.alloc(4 * n)))) {
    }
    if (rc == MDZ_OK &&
        (hipMemcpy(d_ptr.p, ptr, 8ull * n_ptr, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(d_data.p, data, data_bytes, hipMemcpyHostToDevice) != hipSuccess))
        rc = MDZ_ERR_HIP;
    free(ptr);
    if (rc) return rc;
    rc = mdz::rlev2_plain_decompress(d_data.p, data_bytes, (const uint64_t*)d_ptr.p, n, chunk_bytes, d_out.p, raw,
                                     (int32_t*)d_st.p, nullptr);
    if (rc) return rc;
    return finish_host(d_out.p, raw, (const int32_t*)d_st.p, n, out, out_n_bytes);
}

int mdz_rlev2_plain_decompress_gpu(const uint8_t* in, uint64_t in_n_bytes, int64_t** out, uint64_t* out_n_bytes) {
    return mdz_rlev2_plain_decompress_host(in, in_n_bytes, 32 * 1024, reinterpret_cast<uint8_t**>(out),
                                           out_n_bytes);
}

int mdz_lzss_compress_bound(uint64_t in_bytes, uint64_t chunk_bytes, uint64_t* n_chunks, uint64_t* work_bytes,
                            uint64_t* payload_cap) {
    return mdz::lzss_compress_bound(in_bytes, chunk_bytes, n_chunks, work_bytes, payload_cap);
}

int mdz_lzss_compress(const void* in, uint64_t in_bytes, uint64_t chunk_bytes, void* work, void* payload,
                      uint64_t* col_len, uint8_t* col_map, uint64_t* blk_off, void* stream) {
    return mdz::lzss_compress(in, in_bytes, chunk_bytes, work, payload, col_len, col_map, blk_off, S(stream));
}

int mdz_lzss_compress_gpu(const uint8_t* in, uint64_t in_n_bytes, uint8_t** out, uint64_t* out_n_bytes,
                          uint8_t*** out_metadata, uint64_t** out_metadata_lens, uint64_t* out_metadata_num) {
    if (!in || !out || !out_n_bytes || !out_metadata || !out_metadata_lens || !out_metadata_num)
        return MDZ_ERR_INVALID;
    const uint64_t chunk = 128 * 1024;
    uint64_t n = 0, wb = 0, cap = 0;
    int rc = mdz::lzss_compress_bound(in_n_bytes, chunk, &n, &wb, &cap);
    if (rc) return rc;
    DevBuf d_in, d_work, d_pay, d_col, d_map, d_off;
    if ((rc = d_in.alloc(in_n_bytes)) || (rc = d_work.alloc(wb)) || (rc = d_pay.alloc(cap)) ||
        (rc = d_col.alloc(256 * n)) || (rc = d_map.alloc(32 * n)) || (rc = d_off.alloc(8 * (n + 1))))
        return rc;
    if (hipMemcpy(d_in.p, in, in_n_bytes, hipMemcpyHostToDevice) != hipSuccess) return MDZ_ERR_HIP;
    if ((rc = mdz::lzss_compress(d_in.p, in_n_bytes, chunk, d_work.p, d_pay.p, (uint64_t*)d_col.p,
                                 (uint8_t*)d_map.p, (uint64_t*)d_off.p, nullptr)))
        return rc;
    // metadata blobs as the reference returns them: [u64 chunk][blk_off][col_len][col_map]
    uint8_t** md = (uint8_t**)calloc(4, sizeof(uint8_t*));
    uint64_t* lens = (uint64_t*)malloc(4 * sizeof(uint64_t));
    if (!md || !lens) {
        free(md);
        free(lens);
        return MDZ_ERR_NOMEM;
    }
    lens[0] = 8;
    lens[1] = 8 * (n + 1);
    lens[2] = 256 * n;
    lens[3] = 32 * n;
    for (int k = 0; k < 4 && rc == MDZ_OK; ++k)
        if (!(md[k] = (uint8_t*)malloc(lens[k]))) rc = MDZ_ERR_NOMEM;
    if (rc == MDZ_OK) {
        memcpy(md[0], &chunk, 8);
        if (hipMemcpy(md[1], d_off.p, lens[1], hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(md[2], d_col.p, lens[2], hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(md[3], d_map.p, lens[3], hipMemcpyDeviceToHost) != hipSuccess)
            rc = MDZ_ERR_HIP;
    }
    uint8_t* o = nullptr;
    uint64_t total = 0;
    if (rc == MDZ_OK) {
        memcpy(&total, md[1] + 8 * n, 8);
        if (!(o = (uint8_t*)malloc(total ? total : 1))) rc = MDZ_ERR_NOMEM;
        else if (total && hipMemcpy(o, d_pay.p, total, hipMemcpyDeviceToHost) != hipSuccess) rc = MDZ_ERR_HIP;
    }
    if (rc != MDZ_OK) {
        for (int k = 0; k < 4; ++k) free(md[k]);
        free(md);
        free(lens);
        free(o);
        return rc;
    }
    *out = o;
    *out_n_bytes = total;
    *out_metadata = md;
    *out_metadata_lens = lens;
    *out_metadata_num = 4;
    return MDZ_OK;
}

int mdz_meta_split(const uint8_t* meta, uint64_t meta_bytes, const uint8_t** blobs, uint64_t* lens,
                   uint64_t* num) {
    if (!meta || !num || meta_bytes < 8) return MDZ_ERR_INVALID;
    uint64_t k;
    memcpy(&k, meta, 8);
    if (k > (meta_bytes - 8) / 8) return MDZ_ERR_CORRUPT;
    if (k > *num || (k && (!blobs || !lens))) { *num = k; return MDZ_ERR_INVALID; }
    uint64_t off = 8 + 8 * k;
    for (uint64_t i = 0; i < k; ++i) {
        uint64_t len;
        memcpy(&len, meta + 8 + 8 * i, 8);
        if (len > meta_bytes - off) return MDZ_ERR_CORRUPT;
        blobs[i] = meta + off;
        lens[i] = len;
        off += len;
    }
    *num = k;
    return MDZ_OK;
}

// ----------------------------------------------------------------- Deflate
int mdz_inflate_batch(const mdz_inflate_in_t* in, mdz_inflate_status_t* out, int count,
                      int parse_hdr, void* stream) {
    if (count < 0) return MDZ_ERR_INVALID;
    return mdz::inflate_batch(in, out, (uint32_t)count, parse_hdr, nullptr, 0, nullptr, nullptr,
                              S(stream));
}

int mdz_inflate_batch_sched(const mdz_inflate_in_t* in, mdz_inflate_status_t* out, int count,
                            int parse_hdr, const uint32_t* order, uint32_t blocks,
                            uint32_t* cursor, mdz_inflate_cost_t* cost, void* stream) {
    if (count < 0 || !cursor) return MDZ_ERR_INVALID;
    return mdz::inflate_batch(in, out, (uint32_t)count, parse_hdr, order, blocks, cursor, cost,
                              S(stream));
}

int mdz_copy_uncompressed(const mdz_inflate_in_t* in, int count, void* stream) {
    if (count < 0) return MDZ_ERR_INVALID;
    return mdz::copy_uncompressed(in, (uint32_t)count, S(stream));
}

int mdz_inflate_slots(uint32_t* slots) {
    if (!slots) return MDZ_ERR_INVALID;
    return mdz::inflate_slots(slots);
}

int mdz_deflate_decompress_gpu(const uint8_t* in, uint8_t** out, uint64_t in_n_bytes,
                               uint64_t* out_n_bytes, const uint64_t* col_len_f,
                               uint64_t col_n_bytes, const uint64_t* blk_offset_f,
                               uint64_t blk_n_bytes, uint64_t chunk_size) {
    if (!in || !out || !out_n_bytes || !col_len_f || !blk_offset_f || blk_n_bytes < 24 ||
        !chunk_size)
        return MDZ_ERR_INVALID;
    const uint64_t n = blk_n_bytes / 8 - 2;
    if (col_n_bytes / 8 < n) return MDZ_ERR_INVALID;
    const uint64_t data_size = blk_offset_f[0];
    const uint64_t* off = blk_offset_f + 1;
    if (data_size > n * chunk_size) return MDZ_ERR_INVALID;
    for (uint64_t i = 0; i < n; ++i)
        if (off[i] > in_n_bytes || col_len_f[i] > in_n_bytes - off[i]) return MDZ_ERR_INVALID;
    // Every stream gets a whole chunk_size slot, as the reference's reader
    // allocates chunk_size * num_blk (zlib.h:992-996): its writer zero-pads the
    // input to a chunk multiple before compressing (orig_compressor.py:62-66),
    // so the last stream inflates to a full chunk and only data_size bytes of
    // the slots are returned.
    DevBuf d_in, d_out, d_desc, d_st;
    int rc;
    if ((rc = d_in.alloc(in_n_bytes)) || (rc = d_out.alloc(n * chunk_size)) ||
        (rc = d_desc.alloc(n * sizeof(mdz_inflate_in_t))) ||
        (rc = d_st.alloc(n * sizeof(mdz_inflate_status_t))))
        return rc;
    mdz_inflate_in_t* desc = (mdz_inflate_in_t*)malloc(n * sizeof(mdz_inflate_in_t) + 1);
    if (!desc) return MDZ_ERR_NOMEM;
    for (uint64_t i = 0; i < n; ++i) {
        desc[i].src = (const uint8_t*)d_in.p + off[i];
        desc[i].src_bytes = col_len_f[i];
        desc[i].dst = (uint8_t*)d_out.p + i * chunk_size;
        desc[i].dst_bytes = chunk_size;
    }
    const bool ok = hipMemcpy(d_in.p, in, in_n_bytes, hipMemcpyHostToDevice) == hipSuccess &&
                    hipMemcpy(d_desc.p, desc, n * sizeof(mdz_inflate_in_t), hipMemcpyHostToDevice) ==
                        hipSuccess;
    free(desc);
    if (!ok) return MDZ_ERR_HIP;
    rc = mdz::inflate_batch((const mdz_inflate_in_t*)d_desc.p, (mdz_inflate_status_t*)d_st.p,
                            (uint32_t)n, 1, nullptr, 0, nullptr, nullptr, nullptr);
    if (rc) return rc;
    if (hipDeviceSynchronize() != hipSuccess) return MDZ_ERR_HIP;
    mdz_inflate_status_t* st = (mdz_inflate_status_t*)malloc(n * sizeof(mdz_inflate_status_t) + 1);
    uint8_t* o = (uint8_t*)malloc(data_size ? data_size : 1);
    if (!st || !o) { free(st); free(o); return MDZ_ERR_NOMEM; }
    if (hipMemcpy(st, d_st.p, n * sizeof(mdz_inflate_status_t), hipMemcpyDeviceToHost) != hipSuccess ||
        (data_size && hipMemcpy(o, d_out.p, data_size, hipMemcpyDeviceToHost) != hipSuccess)) {
        free(st); free(o);
        return MDZ_ERR_HIP;
    }
    // A stream is good when it decodes cleanly and covers its share of the
    // first data_size bytes: a full chunk, or the remainder for the last one.
    // The reference's writer warp waits for exactly chunk_size bytes per
    // stream (zlib.h:741-743); a short stream inside data_size is corrupt.
    rc = MDZ_OK;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t o = i * chunk_size;
        const uint64_t need = o < data_size ? (data_size - o < chunk_size ? data_size - o : chunk_size) : 0;
        if (st[i].status != 0 || st[i].bytes_written < need) {
            rc = mdz::set_error(MDZ_ERR_CORRUPT, "deflate stream %llu: status %d, %llu of %llu bytes",
                                (unsigned long long)i, (int)st[i].status,
                                (unsigned long long)st[i].bytes_written, (unsigned long long)need);
            break;
        }
    }
    free(st);
    *out = o;
    *out_n_bytes = data_size;
    return rc;
}

int mdz_cudf_uncompress_host(const uint8_t* file, uint64_t file_bytes, uint8_t** out,
                             uint64_t* out_n_bytes) {
    if (!file || !out || !out_n_bytes || file_bytes < 16) return MDZ_ERR_INVALID;
    uint64_t chunk, n;
    memcpy(&chunk, file, 8);
    memcpy(&n, file + 8, 8);
    // [u64 chunk][u64 n][u64 size[n]] must fit, and so must the streams.
    if (!chunk || !n || n > (file_bytes - 16) / 8 || n > 0x7fffffffull) return MDZ_ERR_INVALID;
    const uint64_t data_off = 16 + 8 * n, data_size = file_bytes - data_off;
    std::vector<uint64_t> sz(n);
    memcpy(sz.data(), file + 16, 8 * n);
    uint64_t cur = 0;
    for (uint64_t i = 0; i < n; ++i) {
        // each stream carries a 2-byte zlib header that the reader skips
        if (sz[i] < 2 || sz[i] > data_size - cur) return MDZ_ERR_INVALID;
        cur += sz[i];
    }
    if (n > UINT64_MAX / chunk) return MDZ_ERR_INVALID;
    DevBuf d_in, d_out, d_desc, d_st;
    int rc;
    if ((rc = d_in.alloc(data_size)) || (rc = d_out.alloc(n * chunk)) ||
        (rc = d_desc.alloc(n * sizeof(mdz_inflate_in_t))) ||
        (rc = d_st.alloc(n * sizeof(mdz_inflate_status_t))))
        return rc;
    std::vector<mdz_inflate_in_t> desc(n);
    cur = 0;
    for (uint64_t i = 0; i < n; ++i) {
        desc[i].src = (const uint8_t*)d_in.p + cur + 2;
        desc[i].src_bytes = sz[i] - 2;
        desc[i].dst = (uint8_t*)d_out.p + i * chunk;
        desc[i].dst_bytes = chunk;
        cur += sz[i];
    }
    if (hipMemcpy(d_in.p, file + data_off, data_size, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_desc.p, desc.data(), n * sizeof(mdz_inflate_in_t), hipMemcpyHostToDevice) !=
            hipSuccess)
        return MDZ_ERR_HIP;
    rc = mdz::inflate_batch((const mdz_inflate_in_t*)d_desc.p, (mdz_inflate_status_t*)d_st.p,
                            (uint32_t)n, MDZ_HDR_RAW, nullptr, 0, nullptr, nullptr, nullptr);
    if (rc) return rc;
    if (hipDeviceSynchronize() != hipSuccess) return MDZ_ERR_HIP;
    std::vector<mdz_inflate_status_t> st(n);
    if (hipMemcpy(st.data(), d_st.p, n * sizeof(mdz_inflate_status_t), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return MDZ_ERR_HIP;
    const uint64_t out_size = (n - 1) * chunk + st[n - 1].bytes_written;
    uint8_t* o = (uint8_t*)malloc(out_size ? out_size : 1);
    if (!o) return MDZ_ERR_NOMEM;
    if (out_size && hipMemcpy(o, d_out.p, out_size, hipMemcpyDeviceToHost) != hipSuccess) {
        free(o);
        return MDZ_ERR_HIP;
    }
    rc = MDZ_OK;
    for (uint64_t i = 0; i < n; ++i)
        if (st[i].status != 0 || (i + 1 < n && st[i].bytes_written != chunk)) {
            rc = mdz::set_error(MDZ_ERR_CORRUPT, "cuDF container stream %llu: status %d, %llu bytes",
                                (unsigned long long)i, (int)st[i].status,
                                (unsigned long long)st[i].bytes_written);
            break;
        }
    *out = o;
    *out_n_bytes = out_size;
    return rc;
}

int mdz_cudf_uncompress(const char* in_file, const char* out_file) {
    if (!in_file || !out_file) return MDZ_ERR_INVALID;
    uint8_t* in = nullptr;
    uint64_t n = 0;
    int rc = read_file(in_file, &in, &n);
    if (rc) return mdz::set_error(rc, "cannot read %s", in_file);
    uint8_t* o = nullptr;
    uint64_t on = 0;
    rc = mdz_cudf_uncompress_host(in, n, &o, &on);
    free(in);
    if (rc != MDZ_OK && rc != MDZ_ERR_CORRUPT) return rc;
    FILE* f = fopen(out_file, "wb");
    const bool ok = f && (on == 0 || fwrite(o, 1, on, f) == on);
    if (f && fclose(f) != 0) rc = mdz::set_error(MDZ_ERR_INVALID, "cannot write %s", out_file);
    free(o);
    if (!ok) return mdz::set_error(MDZ_ERR_INVALID, "cannot write %s", out_file);
    return rc;
}

// ------------------------------------------------------- cost model (MLP)
int mdz_cost_features(const mdz_inflate_in_t* in, int count, int parse_hdr, float* x,
                      void* stream) {
    if (count < 0) return MDZ_ERR_INVALID;
    return mdz::mlp_features(in, (uint32_t)count, parse_hdr, x, S(stream));
}
int mdz_mlp_forward(const float* params, const float* x, int n, float* y, void* stream) {
    if (n < 0) return MDZ_ERR_INVALID;
    return mdz::mlp_forward(params, x, (uint32_t)n, y, S(stream));
}
int mdz_mlp_grad(const float* params, const float* x, const float* t, int n, float* grad,
                 float* loss_sum, void* stream) {
    if (n < 0) return MDZ_ERR_INVALID;
    return mdz::mlp_grad(params, x, t, (uint32_t)n, grad, loss_sum, S(stream));
}
int mdz_mlp_adam(float* params, float* grad, float* m, float* v, float lr, uint32_t step,
                 void* stream) {
    return mdz::mlp_adam(params, grad, m, v, lr, step, S(stream));
}

// ------------------------------------------------ batched persistent decode
int mdz_batch_submit(const mdz_batch_t* batch, void* stream) {
    return mdz::batch_submit(batch, S(stream));
}

// ------------------------------------------------------ captured pipelines
}  // extern "C"

struct mdz_graph_s {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" {

int mdz_graph_begin(void* stream) {
    if (!stream) return MDZ_ERR_INVALID;
    return herr(hipStreamBeginCapture(S(stream), hipStreamCaptureModeRelaxed));
}

int mdz_graph_end(void* stream, mdz_graph_t* graph) {
    if (!stream || !graph) return MDZ_ERR_INVALID;
    *graph = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(S(stream), &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return mdz::record_hip(e, "hipStreamEndCapture");
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return mdz::record_hip(e, "hipGraphInstantiate");
    }
    mdz_graph_s* h = new (std::nothrow) mdz_graph_s{g, x};
    if (!h) {
        (void)hipGraphExecDestroy(x);
        (void)hipGraphDestroy(g);
        return MDZ_ERR_NOMEM;
    }
    *graph = h;
    return MDZ_OK;
}

int mdz_graph_launch(mdz_graph_t graph, void* stream) {
    if (!graph) return MDZ_ERR_INVALID;
    return mdz::record_hip(hipGraphLaunch(graph->exec, S(stream)), "hipGraphLaunch");
}

int mdz_graph_destroy(mdz_graph_t graph) {
    if (!graph) return MDZ_OK;
    const hipError_t e1 = hipGraphExecDestroy(graph->exec);
    const hipError_t e2 = hipGraphDestroy(graph->graph);
    delete graph;
    return herr(e1 != hipSuccess ? e1 : e2);
}

int mdz_graph_nodes(mdz_graph_t graph, uint64_t* nodes, uint64_t* kernels) {
    if (!graph || !nodes || !kernels) return MDZ_ERR_INVALID;
    size_t n = 0;
    hipError_t e = hipGraphGetNodes(graph->graph, nullptr, &n);
    if (e != hipSuccess) return herr(e);
    std::vector<hipGraphNode_t> v(n);
    if (n && (e = hipGraphGetNodes(graph->graph, v.data(), &n)) != hipSuccess) return herr(e);
    uint64_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        hipGraphNodeType t;
        if ((e = hipGraphNodeGetType(v[i], &t)) != hipSuccess) return herr(e);
        k += t == hipGraphNodeTypeKernel;
    }
    *nodes = n;
    *kernels = k;
    return MDZ_OK;
}

}  // extern "C"
