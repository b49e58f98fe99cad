// Batched persistent decode: RLE v2 and byte-RLE chunks of two containers
// decoded by one launch.
//
// A grid sized to the co-resident workgroup slots loops over a device work
// queue: thread 0 takes the next item index from a global cursor, the item
// (codec << 31 | chunk) comes from the scheduler's order array (or index
// order: every RLE v2 chunk, then every byte-RLE chunk), and the whole
// workgroup decodes that chunk with the same device code as the per-chunk
// kernels (rlev2_chunk, brle_chunk with 512 windows).  Uneven chunks
// therefore balance across CUs as the queue drains, instead of by the
// hardware's one-time block dispatch.  The queue has no counterpart in the
// reference, whose decoders launch one block per chunk.
#include "brle_dev.h"
#include "rlev2_dev.h"

namespace mdz {
namespace {

constexpr uint32_t kBatchThreads = 512;
static_assert(kBatchThreads == kThreads, "RLE v2 chunk code assumes its workgroup size");

template <int EB>
__global__ __launch_bounds__(kBatchThreads) void batch_kernel(Rlev2Args r, BrleArgs b,
                                                              const uint32_t* __restrict__ items,
                                                              uint32_t n_items,
                                                              uint32_t* __restrict__ cursor) {
    extern __shared__ __align__(16) uint32_t smem[];
    uint32_t* const slot = smem;              // the claimed item index
    uint32_t* const region = smem + 4;        // codec LDS, 16-byte aligned
    for (;;) {
        if (threadIdx.x == 0) *slot = atomicAdd(cursor, 1u);
        __syncthreads();
        const uint32_t k = *slot;
        __syncthreads();
        if (k >= n_items) return;
        const uint32_t it = items ? items[k] : (k < r.n_chunks ? k : (1u << 31) | (k - (uint32_t)r.n_chunks));
        const uint32_t c = it & 0x7fffffffu;
        if ((it >> 31) == 0) {
            if (c < r.n_chunks) rlev2_chunk<EB>(r, c, region);
        } else {
            if (c < b.n_chunks) brle_chunk<kBatchThreads>(b, c, region);
        }
        __syncthreads();
    }
}

__global__ void batch_reset_kernel(uint32_t* c) { *c = 0; }

template <int EB>
int launch_batch(const Rlev2Args& r, const BrleArgs& b, const uint32_t* items, uint32_t n_items,
                 uint32_t blocks, uint32_t* cursor, size_t region, hipStream_t s) {
    auto k = batch_kernel<EB>;
    const size_t smem = 16 + region;
    hipError_t e = set_dynamic_lds((const void*)k, smem);
    if (e != hipSuccess) return record_hip(e, "batch_kernel LDS");
    if (!blocks) {
        int dev = 0, per_cu = 0, cus = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess ||
            (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess ||
            (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, kBatchThreads,
                                                              smem)) != hipSuccess)
            return record_hip(e, "batch occupancy");
        blocks = (uint32_t)(per_cu * cus);
    }
    blocks = min(blocks, n_items);
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, s, cursor);
    if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "batch cursor reset");
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kBatchThreads), smem, s, r, b, items, n_items, cursor);
    return record_hip(hipGetLastError(), "batch_kernel launch");
}

}  // namespace

int batch_submit(const mdz_batch_t* bt, hipStream_t s) {
    if (!bt || !bt->cursor) return MDZ_ERR_INVALID;
    const mdz_lanes_t& L = bt->rlev2;
    Rlev2Args r{};
    BrleArgs b{};
    size_t region = 0;
    int eb = 8;
    if (L.n_chunks) {
        eb = (int)L.elem_bytes;
        const uint32_t ru = L.lane_line;
        const uint64_t chunk = L.chunk_bytes;
        if (!L.payload || !L.blk_off || !L.col_len || (!bt->rlev2_dst && bt->rlev2_dst_bytes) ||
            !(eb == 1 || eb == 2 || eb == 4 || eb == 8) || ru == 0 || chunk % (32ull * eb) ||
            (chunk / eb / 32) % ru || chunk / eb / 32 > 0xffffffu || ((uintptr_t)L.payload & 15) ||
            ((uintptr_t)bt->rlev2_dst % eb))
            return MDZ_ERR_INVALID;
        if (L.n_chunks > 0x7fffffffull) return MDZ_ERR_UNSUPPORTED;
        const uint32_t lds_max = (160 * 1024 - 16 - 16 - (uint32_t)sizeof(Rlev2Tables) - 256) & ~255u;
        uint32_t cap = lds_max;
        if (bt->rlev2_max_chunk) {
            const uint64_t want = ((uint64_t)bt->rlev2_max_chunk + 32 * 12 + 255) & ~255ull;
            if (want < cap) cap = (uint32_t)want;
        }
        r = Rlev2Args{(const uint8_t*)L.payload, L.blk_off, L.col_len, L.n_chunks, chunk, ru, cap,
                      L.payload_bytes, (uint8_t*)bt->rlev2_dst, bt->rlev2_dst_bytes, bt->rlev2_status};
        region = (size_t)cap + 16 + sizeof(Rlev2Tables);
    }
    if (bt->brle_n_chunks) {
        if (!bt->brle_ptr || !bt->brle_data || (!bt->brle_dst && bt->brle_dst_bytes) ||
            bt->brle_chunk_bytes == 0 || bt->brle_chunk_bytes > (1u << 30))
            return MDZ_ERR_INVALID;
        if (bt->brle_n_chunks > 0x7fffffffull) return MDZ_ERR_UNSUPPORTED;
        const uint64_t c = bt->brle_chunk_bytes;
        uint64_t need = bt->brle_max_chunk ? bt->brle_max_chunk : c + c / 128 + 1;
        need = (need + kBrlePad + 255) & ~255ull;
        const uint64_t lds_max = 160 * 1024 - 16 - sizeof(BrleTables<kBatchThreads>) - 256;
        if (need > lds_max) return MDZ_ERR_UNSUPPORTED;
        b = BrleArgs{(const uint8_t*)bt->brle_data, bt->brle_ptr, bt->brle_n_chunks, c,
                     (uint8_t*)bt->brle_dst, bt->brle_dst_bytes, bt->brle_status, (uint32_t)need};
        region = max(region, sizeof(BrleTables<kBatchThreads>) + (size_t)need);
    }
    const uint64_t total = L.n_chunks + bt->brle_n_chunks;
    const uint32_t n_items = bt->items ? bt->n_items : (uint32_t)total;
    if (bt->items && bt->n_items > total) return MDZ_ERR_INVALID;
    if (n_items == 0) return MDZ_OK;
    region = (region + 15) & ~(size_t)15;
    switch (eb) {
    case 1: return launch_batch<1>(r, b, bt->items, n_items, bt->blocks, bt->cursor, region, s);
    case 2: return launch_batch<2>(r, b, bt->items, n_items, bt->blocks, bt->cursor, region, s);
    case 4: return launch_batch<4>(r, b, bt->items, n_items, bt->blocks, bt->cursor, region, s);
    default: return launch_batch<8>(r, b, bt->items, n_items, bt->blocks, bt->cursor, region, s);
    }
}

}  // namespace mdz
