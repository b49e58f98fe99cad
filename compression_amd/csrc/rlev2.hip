// RLE v2 (ORC integer RLEv2, lane-transposed) decoder for gfx950.
//
// Replaces the reference's reader-warp / decoder-warp kernel
// (include/rlev2/decoder_new.h:618-657, decoder_warp :354-611) and host entry
// rlev2::decompress_gpu (include/rlev2/decoder_trans.h:615-681).
//
// The decode itself is rlev2_part (rlev2_dev.h): stage, parallel header walk,
// overlapped check and decode.  Two launch shapes:
//   * whole chunks: one 512-thread workgroup per chunk (the LDS area sized
//     by the largest chunk);
//   * lane quarters: when the caller bounds the largest quarter (8 lanes) of
//     any chunk, one 256-thread workgroup per quarter.  The LDS area shrinks
//     to the quarter's bytes, so several times more header walks -- the
//     serial part -- run per CU.  The four quarters of a chunk get block ids
//     on the same XCD, so the chunk's payload rows they all read come from
//     one L2; statuses combine through status_dev.h.
#include "rlev2_dev.h"
#include "status_dev.h"

namespace mdz {
namespace {

constexpr uint32_t kQuarterLanes = 8, kQuarterThreads = 256;

template <int EB>
__global__ __launch_bounds__(kThreads) void rlev2_decode_kernel(Rlev2Args a) {
    extern __shared__ __align__(16) uint32_t smem[];
    rlev2_part<EB, 32, kThreads>(a, blockIdx.x, 0, smem);
}

// Block id -> (chunk, part) with the parts of a chunk on one XCD (block ids
// go to XCDs round-robin by id mod 8): groups of 8 chunks take 8*P
// consecutive ids, part p of chunk 8g + r at 8*P*g + 8*p + r; the last
// partial group is chunk-major.
MDZ_DEV void part_of_block(uint64_t bid, uint64_t n, uint32_t P, uint64_t& c, uint32_t& part) {
    const uint64_t full = n / 8 * 8 * P;
    if (bid < full) {
        const uint64_t g = bid / (8ull * P), rem = bid % (8ull * P);
        part = (uint32_t)(rem / 8);
        c = 8 * g + rem % 8;
    } else {
        const uint64_t i = bid - full;
        c = n / 8 * 8 + i / P;
        part = (uint32_t)(i % P);
    }
}

template <int EB>
__global__ __launch_bounds__(kQuarterThreads) void rlev2_quarter_kernel(Rlev2Args a) {
    extern __shared__ __align__(16) uint32_t smem[];
    uint64_t c;
    uint32_t part;
    part_of_block(blockIdx.x, a.n_chunks, 32 / kQuarterLanes, c, part);
    rlev2_part<EB, kQuarterLanes, kQuarterThreads>(a, c, part, smem);
}

template <int EB>
hipError_t launch(const mdz_lanes_t& in, void* dst, uint64_t dst_bytes, int32_t* status,
                  uint32_t data_cap, bool quarters, uint32_t win, hipStream_t s) {
    const Rlev2Args a{(const uint8_t*)in.payload, in.blk_off, in.col_len, in.n_chunks,
                      in.chunk_bytes, in.lane_line, data_cap, in.payload_bytes, (uint8_t*)dst,
                      dst_bytes, status, win};
    hipError_t e;
    if (!quarters) {
        const size_t smem = (size_t)data_cap + 16 + sizeof(Rlev2Tables<PartPool<32>::value>);
        auto k = rlev2_decode_kernel<EB>;
        if ((e = set_dynamic_lds((const void*)k, smem)) != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3((uint32_t)in.n_chunks), dim3(kThreads), smem, s, a);
        return hipGetLastError();
    }
    const size_t smem = (size_t)data_cap + 16 + sizeof(Rlev2Tables<PartPool<kQuarterLanes>::value>);
    auto k = rlev2_quarter_kernel<EB>;
    if ((e = set_dynamic_lds((const void*)k, smem)) != hipSuccess) return e;
    const uint64_t n = in.n_chunks;
    if (status) {
        hipLaunchKernelGGL(status_fill_kernel, status_grid(n), dim3(256), 0, s, status, n);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3((uint32_t)(n * (32 / kQuarterLanes))), dim3(kQuarterThreads), smem, s, a);
    if ((e = hipGetLastError()) != hipSuccess || !status) return e;
    hipLaunchKernelGGL(status_finish_kernel, status_grid(n), dim3(256), 0, s, status, n);
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
    if (in->n_chunks > 0x1fffffffull) return MDZ_ERR_UNSUPPORTED;
    if (((uintptr_t)in->payload & 15) != 0 || ((uintptr_t)dst % eb) != 0) return MDZ_ERR_INVALID;
    const bool quarters = tune && tune->max_part_bytes;
    const uint32_t win = tune && tune->rlev2_window ? tune->rlev2_window : kWinDefault;
    if (!rlev2_window_ok(win)) return MDZ_ERR_INVALID;
    const uint32_t tables = quarters ? (uint32_t)sizeof(Rlev2Tables<PartPool<kQuarterLanes>::value>)
                                     : (uint32_t)sizeof(Rlev2Tables<PartPool<32>::value>);
    const uint32_t lds_max = (160 * 1024 - 16 - tables - 256) & ~255u;
    uint32_t cap = lds_max;
    const uint64_t want = quarters ? ((uint64_t)tune->max_part_bytes + 64 + 255) & ~255ull
                          : tune && tune->max_chunk_bytes
                                ? ((uint64_t)tune->max_chunk_bytes + 32 * 12 + 255) & ~255ull
                                : (uint64_t)lds_max;
    if (want < cap) cap = (uint32_t)want;
    hipError_t e;
    switch (eb) {
    case 1: e = launch<1>(*in, dst, dst_bytes, status, cap, quarters, win, s); break;
    case 2: e = launch<2>(*in, dst, dst_bytes, status, cap, quarters, win, s); break;
    case 4: e = launch<4>(*in, dst, dst_bytes, status, cap, quarters, win, s); break;
    default: e = launch<8>(*in, dst, dst_bytes, status, cap, quarters, win, s); break;
    }
    return record_hip(e, "rlev2_decode_kernel launch");
}

}  // namespace mdz
