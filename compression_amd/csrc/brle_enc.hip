// Byte RLE encode for gfx950: one wavefront per chunk.
//
// Replaces brle::compress_gpu and its kernels (reference include/brle/
// brle.h:232-303: one thread runs the byte-at-a-time encoder of :129-230 per
// 256 KiB chunk, then a scan and a shift kernel pack the chunks).  The output
// is byte-identical to that encoder's.
//
// The encoder is a greedy state machine over the chunk's bytes: a pending
// group is empty, a literal group (up to 128 bytes; a third equal byte in a
// row ends it and starts a run) or a run (up to 130 equal bytes).  Its
// decisions change only at the first three bytes of an equal-byte segment,
// at 128-literal and 130-run limits and at segment ends, so here:
//   1. per 4 KiB window, every lane compares 64 bytes with their
//      predecessors and publishes a 64-bit mask of segment starts (LDS);
//   2. lane 0 walks the window segment by segment (ctz over the masks),
//      stepping the state machine a constant number of times per segment
//      and by whole runs otherwise, and appends finished groups
//      {start, length, run?} to an LDS list;
//   3. the wave writes the list: a prefix sum over the group sizes places
//      each group, and each lane writes its group's header and bytes.
// Groups go to a per-chunk slot of the work buffer; a scan kernel turns the
// chunk sizes into ptr[] and a shift kernel packs the slots into `data`.
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

constexpr uint32_t kMinRun = 3, kMaxRun = 130, kMaxLit = 128;
constexpr uint32_t kWin = 4096;                 // bytes per window: 64 lanes x 64
constexpr uint32_t kMaxGroups = kWin / 2 + 4;   // a window ends at most this many groups

struct EncLds {
    uint64_t starts[kWave];                     // bit j of word l: byte 64 l + j starts a segment
    uint32_t g_start[kMaxGroups];               // chunk offset of the group's first byte
    uint32_t g_len[kMaxGroups];                 // bytes << 1 | run
};

enum : uint32_t { kEmpty = 0, kLit = 1, kRun = 2 };

struct Machine {
    uint32_t mode = kEmpty, n = 0, tail = 0, start = 0, groups = 0;

    MDZ_DEV void emit(EncLds& S, uint32_t at, uint32_t len, uint32_t run) {
        S.g_start[groups] = at;
        S.g_len[groups] = len << 1 | run;
        ++groups;
    }
    // Bytes [p, q) are equal; `fresh` != 0: byte p differs from byte p - 1.
    MDZ_DEV void feed(EncLds& S, uint32_t p, uint32_t q, uint32_t fresh) {
        if (fresh && mode == kRun) {            // a run ends where the byte changes
            emit(S, start, n, 1);
            mode = kEmpty;
        }
        uint32_t same = fresh ? 0u : 1u;
        while (p < q) {
            if (mode == kEmpty) {
                mode = kLit;
                start = p;
                n = 1;
                tail = 1;
                ++p;
                same = 1;
                continue;
            }
            if (mode == kRun) {
                const uint32_t take = min(q - p, kMaxRun - n);
                n += take;
                p += take;
                if (n == kMaxRun) {
                    emit(S, start, n, 1);
                    mode = kEmpty;
                }
                continue;
            }
            tail = same ? tail + 1 : 1;         // literal group: one byte
            if (tail == kMinRun) {
                if (n + 1 != kMinRun) {          // the group's last two bytes join the run
                    emit(S, start, n - (kMinRun - 1), 0);
                    start = p - (kMinRun - 1);
                }
                mode = kRun;
                n = kMinRun;
            } else if (++n == kMaxLit) {
                emit(S, start, n, 0);
                mode = kEmpty;
            }
            ++p;
            same = 1;
        }
    }
    MDZ_DEV void finish(EncLds& S) {
        if (mode != kEmpty) emit(S, start, n, mode == kRun ? 1u : 0u);
        mode = kEmpty;
    }
};

__global__ __launch_bounds__(kWave) void brle_encode_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes,
                                                            uint64_t chunk, uint64_t slot,
                                                            uint8_t* __restrict__ work,
                                                            uint64_t* __restrict__ sizes) {
    __shared__ EncLds S;
    const uint32_t l = lane_id();
    const uint64_t c = blockIdx.x;
    const uint64_t base = c * chunk;
    const uint32_t len = (uint32_t)min(chunk, in_bytes - base);
    const uint8_t* src = in + base;
    uint8_t* out = work + c * slot;
    Machine m;                                  // lane 0's
    uint64_t written = 0;
    for (uint32_t w0 = 0; w0 < len; w0 += kWin) {
        const uint32_t wl = min(kWin, len - w0);
        // 1. segment starts of this window: the lane's 64 bytes, loaded
        //    without branches (positions past the chunk read its last byte
        //    and are masked off)
        const uint32_t end = w0 + wl;
        const uint32_t lo = w0 + 64 * l;
        uint32_t prev = lo == 0 ? 0x100u : src[min(lo, len) - 1];
        uint64_t mask = 0;
#pragma unroll
        for (uint32_t j = 0; j < 64; ++j) {
            const uint32_t v = src[min(lo + j, len - 1)];
            if (v != prev) mask |= 1ull << j;
            prev = v;
        }
        S.starts[l] = lo >= end ? 0 : end - lo >= 64 ? mask : mask & ((1ull << (end - lo)) - 1);
        wave_sync();
        // 2. the state machine, one equal-byte piece at a time: a piece ends
        //    at the next set bit of the masks, or at the window's end
        if (l == 0) {
            m.groups = 0;
            uint32_t p = w0;
            const uint32_t nw = (wl + 63) / 64;
            uint32_t fresh = (uint32_t)(S.starts[0] & 1);
            for (uint32_t k = 0; k < nw; ++k) {
                uint64_t bits = S.starts[k];
                if (k == 0) bits &= ~1ull;               // byte w0 opens the first piece
                while (bits) {
                    const uint32_t q = w0 + 64 * k + (uint32_t)__builtin_ctzll(bits);
                    m.feed(S, p, q, fresh);
                    p = q;
                    fresh = 1;
                    bits &= bits - 1;
                }
            }
            m.feed(S, p, end, fresh);
            if (end == len) m.finish(S);
        }
        wave_sync();
        // 3. write the finished groups in order
        const uint32_t ng = __builtin_amdgcn_readfirstlane(m.groups);
        for (uint32_t g0 = 0; g0 < ng; g0 += kWave) {
            const uint32_t g = g0 + l;
            uint32_t at = 0, gl = 0, sz = 0;
            if (g < ng) {
                at = S.g_start[g];
                gl = S.g_len[g];
                sz = (gl & 1) ? 2 : 1 + (gl >> 1);
            }
            const uint32_t incl = wave_incl_scan_u32(sz);
            uint8_t* o = out + written + (incl - sz);
            if (g < ng) {
                const uint32_t n = gl >> 1;
                if (gl & 1) {
                    o[0] = (uint8_t)(n - kMinRun);
                    o[1] = src[at];
                } else {
                    o[0] = (uint8_t)(n + 127);
                    for (uint32_t i = 0; i < n; ++i) o[1 + i] = src[at + i];
                }
            }
            written += (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
        }
        wave_sync();
    }
    if (l == 0) sizes[c] = written;
}

// ptr[0] = 0, ptr[c + 1] = ptr[c] + sizes[c]: one workgroup, each thread
// summing a contiguous run of chunks, then a scan of the thread sums.
constexpr uint32_t kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void brle_scan_kernel(const uint64_t* __restrict__ sizes, uint64_t n,
                                                                 uint64_t* __restrict__ ptr) {
    __shared__ uint64_t part[kScanThreads / kWave];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (n + kScanThreads - 1) / kScanThreads;
    const uint64_t b = min(n, per * t), e = min(n, b + per);
    uint64_t s = 0;
    for (uint64_t i = b; i < e; ++i) s += sizes[i];
    const uint64_t incl = wave_incl_scan_u64(s);
    if (lane_id() == kWave - 1) part[t / kWave] = incl;
    __syncthreads();
    uint64_t before = 0;
    for (uint32_t w = 0; w < t / kWave; ++w) before += part[w];
    uint64_t run = before + incl - s;
    if (t == 0) ptr[0] = 0;
    for (uint64_t i = b; i < e; ++i) {
        run += sizes[i];
        ptr[i + 1] = run;
    }
}

// data[ptr[c], ptr[c + 1]) = work[c * slot, ...): dword stores at aligned
// destination addresses, each assembled from the two source dwords it spans.
constexpr uint32_t kShiftThreads = 256;

__global__ __launch_bounds__(kShiftThreads) void brle_shift_kernel(const uint8_t* __restrict__ work, uint64_t slot,
                                                                   const uint64_t* __restrict__ ptr,
                                                                   uint8_t* __restrict__ data) {
    const uint64_t c = blockIdx.x;
    const uint64_t d0 = ptr[c], n = ptr[c + 1] - d0;
    const uint8_t* s = work + c * slot;        // slot starts are 16-byte aligned
    uint8_t* d = data + d0;
    const uint32_t t = threadIdx.x;
    const uint64_t head = min<uint64_t>((4 - ((uintptr_t)d & 3)) & 3, n);
    for (uint64_t i = t; i < head; i += kShiftThreads) d[i] = s[i];
    const uint64_t nw = (n - head) / 4;
    const uint32_t sk = (uint32_t)(head & 3);          // source skew of each destination dword
    const uint32_t* s4 = reinterpret_cast<const uint32_t*>(s + (head & ~3ull));
    uint32_t* d4 = reinterpret_cast<uint32_t*>(d + head);
    for (uint64_t i = t; i < nw; i += kShiftThreads) {
        // source bytes head + 4i .. + 3 lie in dwords i and i + 1 of s4 (the
        // slot is padded, so the second read stays inside it)
        d4[i] = sk ? alignbyte(s4[i + 1], s4[i], sk) : s4[i];
    }
    for (uint64_t i = head + 4 * nw + t; i < n; i += kShiftThreads) d[i] = s[i];
}

}  // namespace

int brle_compress_bound(uint64_t in_bytes, uint64_t chunk, uint64_t* n_chunks, uint64_t* work_bytes,
                        uint64_t* data_cap) {
    if (in_bytes == 0 || chunk == 0 || chunk > 0xFFFFFFF0ull) return MDZ_ERR_INVALID;
    const uint64_t n = (in_bytes - 1) / chunk + 1;
    const uint64_t worst = chunk + (chunk + kMaxLit - 1) / kMaxLit + 1;
    const uint64_t slot = (worst + 16 + 15) & ~15ull;
    if (n_chunks) *n_chunks = n;
    if (work_bytes) *work_bytes = n * slot + 8 * n;
    if (data_cap) *data_cap = in_bytes + n * ((chunk + kMaxLit - 1) / kMaxLit + 1);
    return MDZ_OK;
}

int brle_compress(const void* in, uint64_t in_bytes, uint64_t chunk, void* work, void* data, uint64_t* ptr,
                  hipStream_t s) {
    uint64_t n = 0, wb = 0;
    int rc = brle_compress_bound(in_bytes, chunk, &n, &wb, nullptr);
    if (rc) return rc;
    if (!in || !work || !data || !ptr) return MDZ_ERR_INVALID;
    if (n > (1ull << 24)) return MDZ_ERR_INVALID;     // grid sizes stay below 2^32 threads
    const uint64_t slot = (wb - 8 * n) / n;
    uint64_t* sizes = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(work) + n * slot);
    hipLaunchKernelGGL(brle_encode_kernel, dim3((uint32_t)n), dim3(kWave), 0, s,
                       static_cast<const uint8_t*>(in), in_bytes, chunk, slot, static_cast<uint8_t*>(work), sizes);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return record_hip(e, "brle_encode_kernel launch");
    hipLaunchKernelGGL(brle_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, sizes, n, ptr);
    if ((e = hipGetLastError()) != hipSuccess) return record_hip(e, "brle_scan_kernel launch");
    hipLaunchKernelGGL(brle_shift_kernel, dim3((uint32_t)n), dim3(kShiftThreads), 0, s,
                       static_cast<const uint8_t*>(work), slot, ptr, static_cast<uint8_t*>(data));
    return record_hip(hipGetLastError(), "brle_shift_kernel launch");
}

}  // namespace mdz
