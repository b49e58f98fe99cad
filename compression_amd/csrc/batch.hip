// Batched persistent decode: RLE v2 and byte-RLE chunks of two containers
// decoded by one launch.
//
// A grid sized to the co-resident workgroup slots loops over a device work
// queue: thread 0 takes the next item index from a global cursor, the item
// comes from the scheduler's order array (or index order: every RLE v2 unit,
// then every byte-RLE chunk), and the whole workgroup decodes it with the
// same device code as the per-codec kernels (rlev2_part, brle_chunk).  An RLE
// v2 unit is a whole chunk (512-thread workgroups) or, when the caller
// bounds the largest lane quarter, one quarter of a chunk (256-thread
// workgroups, four units per chunk).  Uneven units therefore balance across
// CUs as the queue drains, instead of by the hardware's one-time block
// dispatch.  The queue has no counterpart in the reference, whose decoders
// launch one block per chunk.
#include "brle_dev.h"
#include "rlev2_dev.h"
#include "status_dev.h"

namespace mdz {
namespace {

// Item: codec << 31 | part << 29 | chunk (codec 0 = RLE v2, 1 = byte RLE;
// part 0..3 is the RLE v2 lane quarter when quarters are on).
template <int EB, uint32_t NL, uint32_t NT>
__global__ __launch_bounds__(NT) void batch_kernel(Rlev2Args r, BrleArgs b,
                                                   const uint32_t* __restrict__ items,
                                                   uint32_t n_items, uint32_t* __restrict__ cursor) {
    constexpr uint32_t P = 32 / NL;
    extern __shared__ __align__(16) uint32_t smem[];
    uint32_t* const slot = smem;              // the claimed item index
    uint32_t* const region = smem + 4;        // codec LDS, 16-byte aligned
    const uint64_t nr = r.n_chunks * P;       // RLE v2 work units
    for (;;) {
        if (threadIdx.x == 0) *slot = atomicAdd(cursor, 1u);
        __syncthreads();
        const uint32_t k = *slot;
        __syncthreads();
        if (k >= n_items) return;
        uint32_t it;
        if (items) it = items[k];
        else if (k < nr) it = (uint32_t)(k % P) << 29 | (uint32_t)(k / P);
        else it = (1u << 31) | (uint32_t)(k - nr);
        if ((it >> 31) == 0) {
            const uint32_t c = it & 0x1fffffffu, part = (it >> 29) & 3u;
            if (c < r.n_chunks && part < P) rlev2_part<EB, NL, NT>(r, c, part, region);
        } else {
            const uint32_t c = it & 0x7fffffffu;
            if (c < b.n_chunks) brle_chunk<NT>(b, c, region);
        }
        __syncthreads();
    }
}

__global__ void batch_reset_kernel(uint32_t* c) { *c = 0; }

template <int EB, uint32_t NL, uint32_t NT>
int launch_batch(const Rlev2Args& r, const BrleArgs& b, const uint32_t* items, uint32_t n_items,
                 uint32_t blocks, uint32_t* cursor, size_t region, hipStream_t s) {
    auto k = batch_kernel<EB, NL, NT>;
    const size_t smem = 16 + region;
    hipError_t e = set_dynamic_lds((const void*)k, smem);
    if (e != hipSuccess) return record_hip(e, "batch_kernel LDS");
    if (!blocks) {
        int dev = 0, per_cu = 0, cus = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess ||
            (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess ||
            (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, NT, smem)) != hipSuccess)
            return record_hip(e, "batch occupancy");
        blocks = (uint32_t)(per_cu * cus);
    }
    blocks = min(blocks, n_items);
    if (NL < 32 && r.status && r.n_chunks) {
        hipLaunchKernelGGL(status_fill_kernel, status_grid(r.n_chunks), dim3(256), 0, s, r.status, r.n_chunks);
        if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "batch status init");
    }
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, s, cursor);
    if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "batch cursor reset");
    hipLaunchKernelGGL(k, dim3(blocks), dim3(NT), smem, s, r, b, items, n_items, cursor);
    if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "batch_kernel launch");
    if (NL < 32 && r.status && r.n_chunks) {
        hipLaunchKernelGGL(status_finish_kernel, status_grid(r.n_chunks), dim3(256), 0, s, r.status,
                           r.n_chunks);
        return record_hip(hipGetLastError(), "batch status finish");
    }
    return MDZ_OK;
}

template <int EB>
int launch_shape(bool quarters, const Rlev2Args& r, const BrleArgs& b, const uint32_t* items,
                 uint32_t n_items, uint32_t blocks, uint32_t* cursor, size_t region, hipStream_t s) {
    return quarters ? launch_batch<EB, 8, 256>(r, b, items, n_items, blocks, cursor, region, s)
                    : launch_batch<EB, 32, 512>(r, b, items, n_items, blocks, cursor, region, s);
}

}  // namespace

int batch_submit(const mdz_batch_t* bt, hipStream_t s) {
    if (!bt || !bt->cursor) return MDZ_ERR_INVALID;
    const mdz_lanes_t& L = bt->rlev2;
    const bool quarters = bt->rlev2_max_part != 0;
    const uint32_t P = quarters ? 4 : 1, NT = quarters ? 256 : 512;
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
        if (L.n_chunks > 0x1fffffffull) return MDZ_ERR_UNSUPPORTED;
        const uint32_t tables = quarters ? (uint32_t)sizeof(Rlev2Tables<PartPool<8>::value>)
                                         : (uint32_t)sizeof(Rlev2Tables<PartPool<32>::value>);
        const uint32_t lds_max = (160 * 1024 - 16 - 16 - tables - 256) & ~255u;
        uint32_t cap = lds_max;
        const uint64_t want = quarters ? ((uint64_t)bt->rlev2_max_part + 64 + 255) & ~255ull
                              : bt->rlev2_max_chunk
                                    ? ((uint64_t)bt->rlev2_max_chunk + 32 * 12 + 255) & ~255ull
                                    : (uint64_t)lds_max;
        if (want < cap) cap = (uint32_t)want;
        const uint32_t win = bt->rlev2_window ? bt->rlev2_window : kWinDefault;
        if (!rlev2_window_ok(win)) return MDZ_ERR_INVALID;
        r = Rlev2Args{(const uint8_t*)L.payload, L.blk_off, L.col_len, L.n_chunks, chunk, ru, cap,
                      L.payload_bytes, (uint8_t*)bt->rlev2_dst, bt->rlev2_dst_bytes, bt->rlev2_status,
                      win};
        region = (size_t)cap + 16 + tables;
    }
    if (bt->brle_n_chunks) {
        if (!bt->brle_ptr || !bt->brle_data || (!bt->brle_dst && bt->brle_dst_bytes) ||
            bt->brle_chunk_bytes == 0 || bt->brle_chunk_bytes > (1u << 30))
            return MDZ_ERR_INVALID;
        if (bt->brle_n_chunks > 0x7fffffffull) return MDZ_ERR_UNSUPPORTED;
        const uint64_t c = bt->brle_chunk_bytes;
        uint64_t need = bt->brle_max_chunk ? bt->brle_max_chunk : c + c / 128 + 1;
        need = (need + kBrlePad + 255) & ~255ull;
        const size_t tables = quarters ? sizeof(BrleTables<256>) : sizeof(BrleTables<512>);
        const uint64_t lds_max = 160 * 1024 - 16 - tables - 256;
        if (need > lds_max) return MDZ_ERR_UNSUPPORTED;
        b = BrleArgs{(const uint8_t*)bt->brle_data, bt->brle_ptr, bt->brle_n_chunks, c,
                     (uint8_t*)bt->brle_dst, bt->brle_dst_bytes, bt->brle_status, (uint32_t)need};
        region = max(region, tables + (size_t)need);
    }
    const uint64_t total = L.n_chunks * P + bt->brle_n_chunks;
    if (total > 0xffffffffull) return MDZ_ERR_UNSUPPORTED;
    const uint32_t n_items = bt->items ? bt->n_items : (uint32_t)total;
    if (bt->items && bt->n_items > total) return MDZ_ERR_INVALID;
    if (n_items == 0) return MDZ_OK;
    region = (region + 15) & ~(size_t)15;
    (void)NT;
    switch (eb) {
    case 1: return launch_shape<1>(quarters, r, b, bt->items, n_items, bt->blocks, bt->cursor, region, s);
    case 2: return launch_shape<2>(quarters, r, b, bt->items, n_items, bt->blocks, bt->cursor, region, s);
    case 4: return launch_shape<4>(quarters, r, b, bt->items, n_items, bt->blocks, bt->cursor, region, s);
    default: return launch_shape<8>(quarters, r, b, bt->items, n_items, bt->blocks, bt->cursor, region, s);
    }
}

}  // namespace mdz
