// RLE v2 (ORC integer RLEv2, lane-transposed) decoder for gfx950.
//
// Replaces the reference's reader-warp / decoder-warp kernel
// (include/rlev2/decoder_new.h:618-657, decoder_warp :354-611) and host entry
// rlev2::decompress_gpu (include/rlev2/decoder_trans.h:615-681).
//
// One workgroup (4 wave64s) per chunk.  The chunk's interleaved payload is
// staged into LDS once and de-interleaved into 32 contiguous lane streams.
// Each wave then decodes whole lane streams cooperatively: the run headers
// are walked once per wave (uniform control flow), and each run's values are
// produced 64 at a time, one per SIMD lane:
//   SHORT_REPEAT  lane i writes the repeated value;
//   DIRECT        lane i unpacks the i-th MSB-first bit field;
//   DELTA         fixed stride: base + i*delta; packed deltas: lane i unpacks
//                 delta i and a wave prefix sum turns deltas into values;
//   PATCHED_BASE  lane k reads patch entry k, a prefix sum of the gaps places
//                 it, and every lane ORs in the patches that land on its value
//                 (64-bit mask: decoder_new.h:533's 32-bit mask drops patch
//                 bits at positions >= 32).
// Values are written straight to their final place: lane t's q-th value is
// element (q / RU) * 32 * RU + t * RU + q % RU of the chunk, so consecutive
// SIMD lanes store consecutive elements.
#include "rlev2_dev.h"

namespace mdz {
namespace {

template <int EB>
__global__ __launch_bounds__(kThreads) void rlev2_decode_kernel(Rlev2Args a) {
    extern __shared__ __align__(16) uint32_t smem[];
    rlev2_chunk<EB>(a, blockIdx.x, smem);
}

template <int EB>
hipError_t launch(const mdz_lanes_t& in, void* dst, uint64_t dst_bytes, int32_t* status,
                  uint32_t data_cap, hipStream_t s) {
    const size_t smem = (size_t)data_cap + 16 + sizeof(Rlev2Tables);
    auto k = rlev2_decode_kernel<EB>;
    hipError_t e = set_dynamic_lds((const void*)k, smem);
    if (e != hipSuccess) return e;
    const Rlev2Args a{(const uint8_t*)in.payload, in.blk_off, in.col_len, in.n_chunks,
                      in.chunk_bytes, in.lane_line, data_cap, in.payload_bytes, (uint8_t*)dst,
                      dst_bytes, status};
    hipLaunchKernelGGL(k, dim3((uint32_t)in.n_chunks), dim3(kThreads), smem, s, a);
    return hipGetLastError();
}

}  // namespace

int rlev2_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                     const mdz_lanes_tune_t* tune, hipStream_t s) {
    if (!in || !in->payload || !in->blk_off || !in->col_len || (!dst && dst_bytes))
        return MDZ_ERR_INVALID;
    const uint32_t eb = in->elem_bytes, ru = in->lane_line;
    const uint64_t chunk = in->chunk_bytes;
    if (!(eb == 1 || eb == 2 || eb == 4 || eb == 8) || ru == 0 || chunk % (32ull * eb) ||
        (chunk / eb / 32) % ru || chunk / eb / 32 > 0xffffffu)
        return MDZ_ERR_INVALID;
    if (in->n_chunks == 0) return MDZ_OK;
    if (in->n_chunks > 0x7fffffffull) return MDZ_ERR_UNSUPPORTED;
    if (((uintptr_t)in->payload & 15) != 0 || ((uintptr_t)dst % eb) != 0) return MDZ_ERR_INVALID;
    const uint32_t lds_max = (160 * 1024 - 16 - (uint32_t)sizeof(Rlev2Tables) - 256) & ~255u;
    uint32_t cap = lds_max;
    if (tune && tune->max_chunk_bytes) {
        const uint64_t want = ((uint64_t)tune->max_chunk_bytes + 32 * 12 + 255) & ~255ull;
        if (want < cap) cap = (uint32_t)want;
    }
    hipError_t e;
    switch (eb) {
    case 1: e = launch<1>(*in, dst, dst_bytes, status, cap, s); break;
    case 2: e = launch<2>(*in, dst, dst_bytes, status, cap, s); break;
    case 4: e = launch<4>(*in, dst, dst_bytes, status, cap, s); break;
    default: e = launch<8>(*in, dst, dst_bytes, status, cap, s); break;
    }
    return record_hip(e, "rlev2_decode_kernel launch");
}

}  // namespace mdz
