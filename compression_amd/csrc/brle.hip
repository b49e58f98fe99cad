// Byte RLE (ORC byte RLE) decoder for gfx950.
//
// Replaces brle::kernel_decompress / brle::decode (reference
// include/brle/brle.h:26-109, one thread per 256 KiB chunk) and
// brle::decompress_gpu (:305-342).  Control byte c: c <= 127 -> a run of
// c+3 copies of the next byte; c >= 128 -> c-127 literal bytes follow.
//
// One workgroup of 256 threads per chunk:
//   1. the chunk's compressed bytes are staged into LDS with 16-byte loads;
//   2. token discovery in parallel: the bytes are cut into 256 windows;
//      thread i walks the token chain speculatively from kOverlap bytes
//      before its window and notes the first token start in its window and
//      the first one past it.  A wrong starting point usually falls into
//      step with the true chain, so adjacent threads agree where their
//      chains meet; a window that disagrees re-walks from its predecessor's
//      exit, in parallel rounds until every window agrees;
//   3. a block-wide prefix sum of the windows' output lengths gives every
//      thread its output offset;
//   4. output is cut at 16-byte boundaries of the destination: each thread
//      writes the aligned blocks from the boundary at or after its window's
//      output start to the one at or after its end, assembling each block in
//      registers from token pieces (runs splat, literals read from LDS at
//      the block offset), so all stores but the chunk's two end blocks are
//      aligned 16-byte stores.
#include "brle_dev.h"

namespace mdz {
namespace {

__global__ __launch_bounds__(kBrleThreads) void brle_decode_kernel(BrleArgs a) {
    extern __shared__ __align__(16) uint32_t smem[];
    brle_chunk<kBrleThreads>(a, blockIdx.x, smem);
}

}  // namespace

int brle_decompress(const uint64_t* ptr, uint64_t n_chunks, uint64_t chunk_bytes, const void* data,
                    void* dst, uint64_t dst_bytes, int32_t* status, uint32_t max_chunk_bytes,
                    hipStream_t s) {
    if (!ptr || !data || (!dst && dst_bytes) || chunk_bytes == 0 || chunk_bytes > (1u << 30))
        return MDZ_ERR_INVALID;
    if (n_chunks == 0) return MDZ_OK;
    if (n_chunks > 0x7fffffffull) return MDZ_ERR_UNSUPPORTED;
    // worst case one control byte per 128 literals (brle.h:13)
    uint64_t need = max_chunk_bytes ? max_chunk_bytes : chunk_bytes + chunk_bytes / 128 + 1;
    need = (need + kBrlePad + 255) & ~255ull;
    const uint64_t lds_max = 160 * 1024 - sizeof(BrleTables<kBrleThreads>) - 256;
    if (need > lds_max) return MDZ_ERR_UNSUPPORTED;
    const size_t smem = sizeof(BrleTables<kBrleThreads>) + need;
    hipError_t e = set_dynamic_lds((const void*)brle_decode_kernel, smem);
    if (e == hipSuccess) {
        const BrleArgs a{(const uint8_t*)data, ptr, n_chunks, chunk_bytes, (uint8_t*)dst, dst_bytes,
                         status, (uint32_t)need};
        hipLaunchKernelGGL(brle_decode_kernel, dim3((uint32_t)n_chunks), dim3(kBrleThreads), smem, s, a);
        e = hipGetLastError();
    }
    return record_hip(e, "brle_decode_kernel launch");
}

}  // namespace mdz
