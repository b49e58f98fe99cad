// RLE v1 decoder for gfx950.
//
// Replaces the reference's reader-warp / decoder-warp inflate kernel
// (include/rle_v1/decoder.h:426-470, decoder_warp :275-423).  One workgroup
// of one wave64 decodes one chunk:
//   * all 64 lanes stage the chunk's interleaved payload into an LDS ring with
//     coalesced 16-byte loads (each payload byte is read from HBM once);
//   * lanes 0..31 each decode one lane stream straight out of the ring, using
//     the lane-address table to find their 4-byte words, and write their
//     lines with 16-byte stores;
//   * when every still-running lane needs a word beyond the staged window the
//     wave slides the window to the lowest word still needed and decodes on.
// Token grammar (include/rle_v1/decoder.h:296-411): int8 h < 0 -> -h LEB128
// literals; h >= 0 -> int8 delta, LEB128 base, h+3 values base + i*delta.
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

enum : uint32_t { kHdr = 0, kDelta = 1, kBase = 2, kRun = 3, kLit = 4, kDone = 5 };

template <int EB> struct VarintSpec;
template <> struct VarintSpec<1> { static constexpr uint32_t kMax = 2; };
template <> struct VarintSpec<2> { static constexpr uint32_t kMax = 3; };
template <> struct VarintSpec<4> { static constexpr uint32_t kMax = 5; };
template <> struct VarintSpec<8> { static constexpr uint32_t kMax = 10; };

// Gather the 7-bit groups of the first `len` bytes of x (len <= 8).
MDZ_DEV uint64_t gather7x8(uint64_t x) {
    return (x & 0x7full) | ((x >> 1) & (0x7full << 7)) | ((x >> 2) & (0x7full << 14)) |
           ((x >> 3) & (0x7full << 21)) | ((x >> 4) & (0x7full << 28)) |
           ((x >> 5) & (0x7full << 35)) | ((x >> 6) & (0x7full << 42)) |
           ((x >> 7) & (0x7full << 49));
}

// Decode one LEB128 value from the front of the buffer.  Returns the byte
// length (0 = not terminated within kMax bytes or within the bytes left).
template <int EB, int NB>
MDZ_DEV uint32_t varint_peek(const ByteBuf<NB>& b, uint64_t& val) {
    constexpr uint32_t kMax = VarintSpec<EB>::kMax;
    const uint64_t lo = b.lo();
    if constexpr (kMax <= 8) {
        constexpr uint64_t tmask = kMax == 8 ? 0x8080808080808080ull
                                            : ((1ull << (8 * kMax)) - 1) & 0x8080808080808080ull;
        uint64_t term = ~lo & tmask;
        if (term == 0) return 0;
        uint32_t len = (__builtin_ctzll(term) >> 3) + 1;
        if (len > b.nb) return 0;
        uint64_t v = gather7x8(lo);
        val = len >= 8 ? v : (v & ((1ull << (7 * len)) - 1));
        return len;
    } else {
        const uint64_t hi = b.hi();
        uint64_t tl = ~lo & 0x8080808080808080ull;
        uint32_t len;
        if (tl) {
            len = (__builtin_ctzll(tl) >> 3) + 1;
        } else {
            uint64_t th = ~hi & 0x8080ull;
            if (th == 0) return 0;
            len = 8 + (__builtin_ctzll(th) >> 3) + 1;
        }
        if (len > b.nb) return 0;
        uint64_t v = gather7x8(lo);
        if (len < 8) {
            v &= (1ull << (7 * len)) - 1;
        } else if (len > 8) {
            uint64_t h = (hi & 0x7f) | ((hi >> 1) & (0x7full << 7));
            h &= (1ull << (7 * (len - 8))) - 1;
            v |= h << 56;
        }
        val = v;
        return len;
    }
}

template <int EB, int SW>
__global__ __launch_bounds__(64) void rlev1_decode_kernel(
    const uint8_t* __restrict__ payload, const uint64_t* __restrict__ blk_off,
    const uint64_t* __restrict__ col_len, uint64_t n_chunks, uint64_t chunk_bytes,
    uint32_t line, uint8_t* __restrict__ dst, uint64_t dst_bytes, int32_t* __restrict__ status,
    uint32_t ring_words) {
    constexpr int NB = EB == 8 ? 16 : 8;
    constexpr uint32_t kMaxV = VarintSpec<EB>::kMax;
    constexpr uint32_t kRefillAt = NB - 4;

    extern __shared__ __align__(16) uint32_t smem[];
    __shared__ LaneTable T;
    uint32_t* ring = smem;

    const uint64_t c = blockIdx.x;
    if (c >= n_chunks) return;
    const uint32_t l = lane_id(), t = l & 31;
    const bool decoder = l < 32;

    const uint64_t c_off = blk_off[c];
    const uint64_t c_len = blk_off[c + 1] - c_off;
    const uint64_t cl = col_len[c * 32 + t];
    const uint32_t W = (uint32_t)((cl + 3) / 4);
    lane_table_build(T, W);
    const uint32_t total = T.total;
    const uint32_t chunk_words = (uint32_t)(c_len / 4);
    if ((uint64_t)total * 4 > c_len || (c_off & 15) || (c_len & 15)) {
        if (l == 0 && status) status[c] = MDZ_ST_INPUT_OVERRUN;
        return;
    }

    const u32x4* src = reinterpret_cast<const u32x4*>(payload + c_off);
    const uint64_t out_base = c * chunk_bytes;
    const uint64_t limit = dst_bytes > out_base ? min(dst_bytes - out_base, chunk_bytes) : 0;
    const uint32_t nvals = (uint32_t)(chunk_bytes / 32 / EB);

    LaneCursor cur;
    cur.init(T, t, decoder ? W : 0);
    ByteBuf<NB> bb;
    LineEmitter<EB, SW> em;
    em.init(dst + out_base, limit, t, line);
    uint32_t mode = decoder ? kHdr : kDone;
    uint32_t rem = 0, vo = 0;
    uint64_t rv = 0;
    int64_t rd = 0;
    int32_t err = MDZ_ST_OK;
    const uint32_t last_bytes = (uint32_t)(cl - 4ull * (W ? W - 1 : 0));

    uint32_t hi = min(chunk_words, ring_words);
    ring_stage(ring, ring_words, src, 0, hi);
    __syncthreads();

    const uint32_t rmask = ring_words - 1;
    for (;;) {
        // ---- decode as far as the staged window allows
        while (mode != kDone) {
            while (bb.nb <= kRefillAt && cur.r < cur.W) {
                uint32_t a = cur.addr();
                if (a >= hi) break;
                uint32_t w = ring[a & rmask];
                bb.append(w, cur.r + 1 == cur.W ? last_bytes : 4);
                cur.advance(T, t);
            }
            const uint32_t need = mode == kRun ? 0 : (mode == kLit || mode == kBase) ? kMaxV : 1;
            if (bb.nb < need && cur.r < cur.W) break;   // blocked on the window

            if (mode == kRun) {
                em.push(rv);
                rv += (uint64_t)rd;
                ++vo;
                if (--rem == 0) mode = kHdr;
            } else if (mode == kLit) {
                uint64_t v;
                uint32_t n = varint_peek<EB>(bb, v);
                if (n == 0) { err = bb.nb < kMaxV ? MDZ_ST_INPUT_OVERRUN : MDZ_ST_BAD_SYMBOL; mode = kDone; break; }
                bb.consume(n);
                em.push(v);
                ++vo;
                if (--rem == 0) mode = kHdr;
            } else if (mode == kHdr) {
                if (bb.nb == 0) { err = MDZ_ST_INPUT_OVERRUN; mode = kDone; break; }
                uint32_t h = (uint32_t)(bb.lo() & 0xff);
                bb.consume(1);
                if (h & 0x80) { rem = 256 - h; mode = kLit; }
                else { rem = h + 3; mode = kDelta; }
                if (rem > nvals - vo) { err = MDZ_ST_OUTPUT_OVERRUN; mode = kDone; break; }
            } else if (mode == kDelta) {
                if (bb.nb == 0) { err = MDZ_ST_INPUT_OVERRUN; mode = kDone; break; }
                rd = (int64_t)(int8_t)(bb.lo() & 0xff);
                bb.consume(1);
                mode = kBase;
            } else {   // kBase
                uint64_t v;
                uint32_t n = varint_peek<EB>(bb, v);
                if (n == 0) { err = bb.nb < kMaxV ? MDZ_ST_INPUT_OVERRUN : MDZ_ST_BAD_SYMBOL; mode = kDone; break; }
                bb.consume(n);
                rv = v;
                mode = kRun;
            }
            if (vo == nvals) mode = kDone;
        }

        // ---- slide the window to the lowest word any running lane needs
        uint32_t needw = (mode != kDone && cur.r < cur.W) ? cur.addr() : 0xffffffffu;
        uint32_t lo = wave_min_u32(needw);
        if (lo == 0xffffffffu) break;
        lo &= ~3u;
        uint32_t nhi = min(chunk_words, lo + ring_words);
        __syncthreads();
        if (nhi > hi) ring_stage(ring, ring_words, src, max(hi, lo), nhi);
        hi = max(hi, nhi);
        __syncthreads();
    }

    // ---- per-chunk status: the lowest lane's error wins
    uint64_t bad = __ballot(err != MDZ_ST_OK);
    if (status) {
        if (bad) {
            uint32_t first = (uint32_t)__builtin_ctzll(bad);
            int32_t e = __shfl(err, first, kWave);
            if (l == 0) status[c] = e;
        } else if (l == 0) {
            status[c] = MDZ_ST_OK;
        }
    }
}

template <int EB, int SW>
hipError_t launch(const mdz_lanes_t& in, void* dst, uint64_t dst_bytes, int32_t* status,
                  uint32_t ring_words, hipStream_t s) {
    size_t smem = (size_t)ring_words * 4;
    auto k = rlev1_decode_kernel<EB, SW>;
    hipError_t e = set_dynamic_lds((const void*)k, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((uint32_t)in.n_chunks), dim3(64), smem, s,
                       (const uint8_t*)in.payload, in.blk_off, in.col_len, in.n_chunks,
                       in.chunk_bytes, in.lane_line, (uint8_t*)dst, dst_bytes, status, ring_words);
    return hipGetLastError();
}

template <int EB>
hipError_t launch_eb(const mdz_lanes_t& in, void* dst, uint64_t dst_bytes, int32_t* status,
                     uint32_t ring_words, uint32_t sw, hipStream_t s) {
    if (sw == 16) return launch<EB, 16>(in, dst, dst_bytes, status, ring_words, s);
    if (sw == 8) return launch<EB, 8>(in, dst, dst_bytes, status, ring_words, s);
    if constexpr (EB <= 4)
        if (sw == 4) return launch<EB, 4>(in, dst, dst_bytes, status, ring_words, s);
    if constexpr (EB <= 2)
        if (sw == 2) return launch<EB, 2>(in, dst, dst_bytes, status, ring_words, s);
    if constexpr (EB == 1)
        if (sw == 1) return launch<EB, 1>(in, dst, dst_bytes, status, ring_words, s);
    return hipErrorInvalidValue;
}

}  // namespace

int rlev1_decompress(const mdz_lanes_t* in, void* dst, uint64_t dst_bytes, int32_t* status,
                     const mdz_lanes_tune_t* tune, hipStream_t s) {
    if (!in || !in->payload || !in->blk_off || !in->col_len || (!dst && dst_bytes))
        return MDZ_ERR_INVALID;
    const uint32_t eb = in->elem_bytes, line = in->lane_line;
    const uint64_t chunk = in->chunk_bytes;
    if (!(eb == 1 || eb == 2 || eb == 4 || eb == 8) || line == 0 || line % eb ||
        chunk % 32 || (chunk / 32) % line || (chunk / 32) / eb < 1 || chunk / 32 / eb > 0x7fffffffu)
        return MDZ_ERR_INVALID;
    if (in->n_chunks == 0) return MDZ_OK;
    if (in->n_chunks > 0x7fffffffull) return MDZ_ERR_UNSUPPORTED;
    if (((uintptr_t)in->payload & 15) != 0) return MDZ_ERR_INVALID;
    uint32_t ring = tune && tune->ring_words ? tune->ring_words : 16384;
    if (ring < 256 || (ring & (ring - 1)) || ring > 32768) return MDZ_ERR_INVALID;
    // widest store that divides the line and matches the output alignment
    uint32_t sw = 16;
    while (sw > eb && ((line % sw) || ((uintptr_t)dst % sw) || (chunk % sw))) sw >>= 1;
    hipError_t e;
    switch (eb) {
    case 1: e = launch_eb<1>(*in, dst, dst_bytes, status, ring, sw, s); break;
    case 2: e = launch_eb<2>(*in, dst, dst_bytes, status, ring, sw, s); break;
    case 4: e = launch_eb<4>(*in, dst, dst_bytes, status, ring, sw, s); break;
    default: e = launch_eb<8>(*in, dst, dst_bytes, status, ring, sw, s); break;
    }
    return record_hip(e, "rlev1_decode_kernel launch");
}

}  // namespace mdz
