// Self-describing container file (MDZF) and its staged reader.
//
// One file carries everything a decode needs, in place of the reference's
// side files in the working directory (rle_v1 col_len / blk_offset under
// ./input_data, reference include/rle_v1/rlev1.h:81-87,110-173;
// include/brle/brle_trans.h:646-657) and the LZSS .meta blob
// (include/utils.h:46-118):
//
//   [0, 256)      header: magic "MDZF", version, codec, parameters, counts,
//                 section offsets; CRC-32 of bytes [0, 252) at 252
//   blk_off       u64[n + 1]: payload offset of chunk (or stream) i; the last
//                 entry as the codec defines it (RLE v2: uncompressed size)
//   col_len       u64[32 n]  (RLE v1, RLE v2, LZSS)
//   col_map       u8[32 n]   (LZSS)
//   raw_len       u64[n]     (Deflate: output bytes of each stream)
//   payload       the codec's chunk bytes
// Every section starts on a 4 KiB boundary, so the payload can be mapped
// and staged page by page.  All integers are little-endian.
//
// mdz_file_open maps the file and checks it; the tables are read from the
// mapping.  mdz_file_decode streams the payload to HBM through two pinned
// staging buffers: the host reads piece k+1 from the file (pread, a few
// threads) while piece k moves over PCIe on a copy stream and earlier pieces
// decode on the caller's stream; each piece's chunks are decoded as soon as
// its bytes have landed (an event orders the two streams).
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <new>
#include <thread>

#include "deflate_hdr_dev.h"
#include "mdz_internal.h"

namespace {

constexpr uint32_t kHdrBytes = 256, kCrcAt = 252, kVersion = 1;
constexpr uint64_t kAlign = 4096;

struct Header {                    // on-disk layout, 256 bytes
    char magic[4];
    uint32_t version, header_bytes, codec;
    uint32_t elem_bytes, lane_line, framing, reserved0;
    uint64_t n_chunks, chunk_bytes, raw_bytes, payload_bytes;
    uint64_t blk_off_at, col_len_at, col_map_at, raw_len_at, payload_at;
    uint8_t pad[kCrcAt - 32 - 8 * 9];
    uint32_t crc;
};
static_assert(sizeof(Header) == kHdrBytes, "header layout");

uint32_t crc32_of(const uint8_t* p, uint64_t n) {
    uint32_t r = 0xFFFFFFFFu;
    for (uint64_t i = 0; i < n; ++i) {
        r ^= p[i];
        for (int k = 0; k < 8; ++k) r = mdz::crc_times_x(r);
    }
    return ~r;
}

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

bool has_lanes(uint32_t c) { return c == MDZ_CODEC_RLEV1 || c == MDZ_CODEC_RLEV2 || c == MDZ_CODEC_LZSS; }

// a + b without wrap-around
bool add_ok(uint64_t a, uint64_t b, uint64_t* r) { return !__builtin_add_overflow(a, b, r); }
bool mul_ok(uint64_t a, uint64_t b, uint64_t* r) { return !__builtin_mul_overflow(a, b, r); }

// Parameter checks shared by the writer and the reader.
int check_info(const mdz_file_info_t& i) {
    if (i.n_chunks == 0 || i.n_chunks >= (1ull << 32)) return MDZ_ERR_INVALID;
    switch (i.codec) {
    case MDZ_CODEC_RLEV1:
        if (i.elem_bytes != 1 && i.elem_bytes != 2 && i.elem_bytes != 4 && i.elem_bytes != 8)
            return MDZ_ERR_INVALID;
        if (i.lane_line == 0 || i.lane_line % i.elem_bytes) return MDZ_ERR_INVALID;
        break;
    case MDZ_CODEC_RLEV2:
        if (i.elem_bytes != 1 && i.elem_bytes != 2 && i.elem_bytes != 4 && i.elem_bytes != 8)
            return MDZ_ERR_INVALID;
        if (i.lane_line == 0) return MDZ_ERR_INVALID;
        break;
    case MDZ_CODEC_BRLE:
    case MDZ_CODEC_LZSS:
        break;
    case MDZ_CODEC_DEFLATE:
        if (!mdz::inflate_hdr_ok((int)i.framing)) return MDZ_ERR_INVALID;
        return MDZ_OK;                      // chunk_bytes unused; raw_len carries sizes
    default:
        return MDZ_ERR_INVALID;
    }
    uint64_t cap;
    if (i.chunk_bytes == 0 || !mul_ok(i.n_chunks, i.chunk_bytes, &cap) || i.raw_bytes > cap)
        return MDZ_ERR_INVALID;
    return MDZ_OK;
}

// Payload extent of chunk c: [blk_off[c], end) with end = the next chunk's
// start, or the payload's end for the last chunk.
uint64_t chunk_end(const uint64_t* blk_off, uint64_t n, uint64_t payload_bytes, uint64_t c) {
    return c + 1 < n ? blk_off[c + 1] : payload_bytes;
}

int check_tables(const mdz_file_info_t& i, const uint64_t* blk_off, const uint64_t* raw_len) {
    for (uint64_t c = 0; c < i.n_chunks; ++c) {
        const uint64_t b = blk_off[c], e = chunk_end(blk_off, i.n_chunks, i.payload_bytes, c);
        if (b > e || e > i.payload_bytes) return MDZ_ERR_INVALID;
    }
    if (i.codec == MDZ_CODEC_RLEV2) {
        if (blk_off[i.n_chunks] != i.raw_bytes) return MDZ_ERR_INVALID;
    } else if (blk_off[i.n_chunks] > i.payload_bytes) {
        return MDZ_ERR_INVALID;
    }
    if (i.codec == MDZ_CODEC_DEFLATE) {
        uint64_t total = 0;
        for (uint64_t c = 0; c < i.n_chunks; ++c)
            if (!add_ok(total, raw_len[c], &total)) return MDZ_ERR_INVALID;
        if (total != i.raw_bytes) return MDZ_ERR_INVALID;
    }
    return MDZ_OK;
}

}  // namespace

struct mdz_file_s {
    int fd = -1;
    uint8_t* map = nullptr;
    uint64_t size = 0;
    mdz_file_info_t info{};
    const uint64_t *blk_off = nullptr, *col_len = nullptr, *raw_len = nullptr;
    const uint8_t* col_map = nullptr;
    uint64_t payload_at = 0;               // file offset of the payload
    uint64_t max_chunk = 0;                // largest chunk's payload bytes
    // device side, made on the first decode and kept for the next ones
    int device = -1;
    void *d_pay = nullptr, *d_blk = nullptr, *d_col = nullptr, *d_map = nullptr, *d_desc = nullptr;
    uint8_t* pinned[2] = {nullptr, nullptr};
    uint64_t pinned_bytes = 0;
    hipStream_t copy = nullptr;
    hipEvent_t landed[2] = {nullptr, nullptr};
    hipEvent_t start = nullptr;            // the caller's stream at a decode's start

    void release_device() {
        // decodes on the caller's streams may still read the device copies
        if (d_pay) (void)hipDeviceSynchronize();
        if (copy) (void)hipStreamSynchronize(copy);
        if (start) (void)hipEventDestroy(start);
        start = nullptr;
        for (int k = 0; k < 2; ++k) {
            if (landed[k]) (void)hipEventDestroy(landed[k]);
            if (pinned[k]) (void)hipHostFree(pinned[k]);
            landed[k] = nullptr;
            pinned[k] = nullptr;
        }
        if (copy) (void)hipStreamDestroy(copy);
        for (void** p : {&d_pay, &d_blk, &d_col, &d_map, &d_desc}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        copy = nullptr;
        pinned_bytes = 0;
        device = -1;
    }
    ~mdz_file_s() {
        release_device();
        if (map) munmap(map, size);
        if (fd >= 0) close(fd);
    }
};

namespace {

int device_setup(mdz_file_s* f, uint64_t stage) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return mdz::record_hip(e, "file: device");
    if (f->device != dev) f->release_device();
    const mdz_file_info_t& i = f->info;
    const uint64_t n = i.n_chunks;
    if (!f->d_pay) {
        // metadata tables are uploaded once; a failure leaves nothing behind
        auto undo = [&](int rc) {
            f->release_device();
            return rc;
        };
        f->device = dev;
        if (hipMalloc(&f->d_pay, i.payload_bytes + 16) != hipSuccess ||
            hipMalloc(&f->d_blk, 8 * (n + 1)) != hipSuccess)
            return undo(MDZ_ERR_NOMEM);
        if ((e = hipMemcpy(f->d_blk, f->blk_off, 8 * (n + 1), hipMemcpyHostToDevice)) != hipSuccess)
            return undo(mdz::record_hip(e, "file: blk_off upload"));
        if (f->col_len) {
            if (hipMalloc(&f->d_col, 256 * n) != hipSuccess) return undo(MDZ_ERR_NOMEM);
            if ((e = hipMemcpy(f->d_col, f->col_len, 256 * n, hipMemcpyHostToDevice)) != hipSuccess)
                return undo(mdz::record_hip(e, "file: col_len upload"));
        }
        if (f->col_map) {
            if (hipMalloc(&f->d_map, 32 * n) != hipSuccess) return undo(MDZ_ERR_NOMEM);
            if ((e = hipMemcpy(f->d_map, f->col_map, 32 * n, hipMemcpyHostToDevice)) != hipSuccess)
                return undo(mdz::record_hip(e, "file: col_map upload"));
        }
        if (i.codec == MDZ_CODEC_DEFLATE && hipMalloc(&f->d_desc, sizeof(mdz_inflate_in_t) * n) != hipSuccess)
            return undo(MDZ_ERR_NOMEM);
        if ((e = hipStreamCreateWithFlags(&f->copy, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&f->landed[0], hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&f->landed[1], hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&f->start, hipEventDisableTiming)) != hipSuccess)
            return undo(mdz::record_hip(e, "file: stream/events"));
    }
    const uint64_t want = std::max(stage, f->max_chunk);
    if (f->pinned_bytes < want) {
        for (int k = 0; k < 2; ++k) {
            if (f->pinned[k]) {
                (void)hipEventSynchronize(f->landed[k]);
                (void)hipHostFree(f->pinned[k]);
                f->pinned[k] = nullptr;
            }
            if (hipHostMalloc((void**)&f->pinned[k], want, hipHostMallocDefault) != hipSuccess)
                return MDZ_ERR_NOMEM;
        }
        f->pinned_bytes = want;
    }
    return MDZ_OK;
}

// Decodes chunks (or streams) [c0, c1), whose payload bytes are on the device.
int decode_range(mdz_file_s* f, uint64_t c0, uint64_t c1, uint8_t* dst, uint64_t dst_bytes,
                 void* status, hipStream_t s) {
    const mdz_file_info_t& i = f->info;
    const uint64_t n = c1 - c0;
    if (i.codec == MDZ_CODEC_DEFLATE)
        return mdz::inflate_batch(static_cast<const mdz_inflate_in_t*>(f->d_desc) + c0,
                                  static_cast<mdz_inflate_status_t*>(status) + c0, (uint32_t)n,
                                  (int)i.framing, nullptr, 0, nullptr, nullptr, s);
    const uint64_t o0 = c0 * i.chunk_bytes;
    const uint64_t nb = std::min(dst_bytes > o0 ? dst_bytes - o0 : 0, n * i.chunk_bytes);
    int32_t* st = status ? static_cast<int32_t*>(status) + c0 : nullptr;
    const uint64_t* blk = static_cast<const uint64_t*>(f->d_blk) + c0;
    if (i.codec == MDZ_CODEC_BRLE) {
        uint64_t mx = 0;
        for (uint64_t c = c0; c < c1; ++c)
            mx = std::max(mx, chunk_end(f->blk_off, i.n_chunks, i.payload_bytes, c) - f->blk_off[c]);
        return mdz::brle_decompress(blk, n, i.chunk_bytes, f->d_pay, dst + o0, nb, st,
                                    (uint32_t)std::min<uint64_t>(mx, 0xFFFFFFFFu), s);
    }
    mdz_lanes_t L = {f->d_pay, blk, static_cast<const uint64_t*>(f->d_col) + 32 * c0,
                     f->d_map ? static_cast<const uint8_t*>(f->d_map) + 32 * c0 : nullptr,
                     n, i.chunk_bytes, i.elem_bytes, i.lane_line, i.payload_bytes};
    switch (i.codec) {
    case MDZ_CODEC_RLEV1: return mdz::rlev1_decompress(&L, dst + o0, nb, st, nullptr, s);
    case MDZ_CODEC_RLEV2: return mdz::rlev2_decompress(&L, dst + o0, nb, st, nullptr, s);
    default: return mdz::lzss_decompress(&L, dst + o0, nb, st, s);
    }
}

// Reads file bytes [at, at + len) into a staging buffer: pread (the page
// cache copies straight into pinned memory, no page faults on the mapping),
// split over a few threads for large pieces.
int stage_read(int fd, uint8_t* dst, uint64_t at, uint64_t len) {
    constexpr int kMaxThreads = 4;
    const int nt = len >= (16ull << 20) ? kMaxThreads : 1;
    std::atomic<int> bad{0};
    auto part = [&](int t) {
        uint64_t b = len * t / nt;
        const uint64_t e = len * (t + 1) / nt;
        while (b < e) {
            const ssize_t r = pread(fd, dst + b, e - b, (off_t)(at + b));
            if (r <= 0) {
                bad = 1;
                return;
            }
            b += (uint64_t)r;
        }
    };
    std::thread th[kMaxThreads];
    for (int t = 1; t < nt; ++t) th[t] = std::thread(part, t);
    part(0);
    for (int t = 1; t < nt; ++t) th[t].join();
    return bad ? MDZ_ERR_INVALID : MDZ_OK;
}

}  // namespace

extern "C" {

int mdz_file_write(const char* path, const mdz_file_info_t* info, const void* payload,
                   const uint64_t* blk_off, const uint64_t* col_len, const uint8_t* col_map,
                   const uint64_t* raw_len) {
    if (!path || !info || !blk_off || (!payload && info->payload_bytes)) return MDZ_ERR_INVALID;
    const mdz_file_info_t& i = *info;
    int rc = check_info(i);
    if (rc) return rc;
    if (has_lanes(i.codec) && !col_len) return MDZ_ERR_INVALID;
    if (i.codec == MDZ_CODEC_LZSS && !col_map) return MDZ_ERR_INVALID;
    if (i.codec == MDZ_CODEC_DEFLATE && !raw_len) return MDZ_ERR_INVALID;
    if ((rc = check_tables(i, blk_off, raw_len))) return rc;
    const uint64_t n = i.n_chunks;
    Header h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, "MDZF", 4);
    h.version = kVersion;
    h.header_bytes = kHdrBytes;
    h.codec = i.codec;
    h.elem_bytes = i.elem_bytes;
    h.lane_line = i.lane_line;
    h.framing = i.framing;
    h.n_chunks = n;
    h.chunk_bytes = i.chunk_bytes;
    h.raw_bytes = i.raw_bytes;
    h.payload_bytes = i.payload_bytes;
    uint64_t at = align_up(kHdrBytes);
    h.blk_off_at = at;
    at = align_up(at + 8 * (n + 1));
    if (has_lanes(i.codec)) {
        h.col_len_at = at;
        at = align_up(at + 256 * n);
    }
    if (i.codec == MDZ_CODEC_LZSS) {
        h.col_map_at = at;
        at = align_up(at + 32 * n);
    }
    if (i.codec == MDZ_CODEC_DEFLATE) {
        h.raw_len_at = at;
        at = align_up(at + 8 * n);
    }
    h.payload_at = at;
    h.crc = crc32_of(reinterpret_cast<const uint8_t*>(&h), kCrcAt);

    FILE* fp = fopen(path, "wb");
    if (!fp) return MDZ_ERR_INVALID;
    bool ok = true;
    auto put = [&](uint64_t where, const void* p, uint64_t bytes) {
        if (!ok || !bytes) return;
        ok = fseeko(fp, (off_t)where, SEEK_SET) == 0 && fwrite(p, 1, bytes, fp) == bytes;
    };
    put(0, &h, sizeof h);
    put(h.blk_off_at, blk_off, 8 * (n + 1));
    if (h.col_len_at) put(h.col_len_at, col_len, 256 * n);
    if (h.col_map_at) put(h.col_map_at, col_map, 32 * n);
    if (h.raw_len_at) put(h.raw_len_at, raw_len, 8 * n);
    put(h.payload_at, payload, i.payload_bytes);
    if (ok && i.payload_bytes == 0) {           // the file still spans the payload offset
        ok = fseeko(fp, 0, SEEK_END) == 0 && ftruncate(fileno(fp), (off_t)h.payload_at) == 0;
    }
    ok = (fclose(fp) == 0) && ok;
    return ok ? MDZ_OK : MDZ_ERR_INVALID;
}

int mdz_file_open(const char* path, mdz_file_t* out) {
    if (!path || !out) return MDZ_ERR_INVALID;
    *out = nullptr;
    mdz_file_s* f = new (std::nothrow) mdz_file_s;
    if (!f) return MDZ_ERR_NOMEM;
    struct stat sb;
    f->fd = open(path, O_RDONLY);
    if (f->fd < 0 || fstat(f->fd, &sb) != 0 || (uint64_t)sb.st_size < kHdrBytes) {
        delete f;
        return MDZ_ERR_INVALID;
    }
    f->size = (uint64_t)sb.st_size;
    void* m = mmap(nullptr, f->size, PROT_READ, MAP_PRIVATE, f->fd, 0);
    if (m == MAP_FAILED) {
        delete f;
        return MDZ_ERR_INVALID;
    }
    f->map = static_cast<uint8_t*>(m);
    Header h;
    memcpy(&h, f->map, sizeof h);
    auto fail = [&](int rc) {
        delete f;
        return rc;
    };
    if (memcmp(h.magic, "MDZF", 4) || h.version != kVersion || h.header_bytes != kHdrBytes)
        return fail(MDZ_ERR_INVALID);
    if (crc32_of(f->map, kCrcAt) != h.crc) return fail(MDZ_ERR_CORRUPT);
    mdz_file_info_t& i = f->info;
    i.codec = h.codec;
    i.elem_bytes = h.elem_bytes;
    i.lane_line = h.lane_line;
    i.framing = h.framing;
    i.n_chunks = h.n_chunks;
    i.chunk_bytes = h.chunk_bytes;
    i.raw_bytes = h.raw_bytes;
    i.payload_bytes = h.payload_bytes;
    if (check_info(i)) return fail(MDZ_ERR_INVALID);
    const uint64_t n = i.n_chunks;
    // every section lies inside the file
    auto within = [&](uint64_t at, uint64_t bytes) {
        uint64_t e;
        return at >= kHdrBytes && add_ok(at, bytes, &e) && e <= f->size;
    };
    if (!within(h.blk_off_at, 8 * (n + 1)) || !within(h.payload_at, i.payload_bytes))
        return fail(MDZ_ERR_INVALID);
    if (h.blk_off_at % 8 || h.col_len_at % 8 || h.raw_len_at % 8) return fail(MDZ_ERR_INVALID);
    if (has_lanes(i.codec) && !within(h.col_len_at, 256 * n)) return fail(MDZ_ERR_INVALID);
    if (i.codec == MDZ_CODEC_LZSS && !within(h.col_map_at, 32 * n)) return fail(MDZ_ERR_INVALID);
    if (i.codec == MDZ_CODEC_DEFLATE && !within(h.raw_len_at, 8 * n)) return fail(MDZ_ERR_INVALID);
    f->blk_off = reinterpret_cast<const uint64_t*>(f->map + h.blk_off_at);
    f->col_len = has_lanes(i.codec) ? reinterpret_cast<const uint64_t*>(f->map + h.col_len_at) : nullptr;
    f->col_map = i.codec == MDZ_CODEC_LZSS ? f->map + h.col_map_at : nullptr;
    f->raw_len = i.codec == MDZ_CODEC_DEFLATE ? reinterpret_cast<const uint64_t*>(f->map + h.raw_len_at) : nullptr;
    f->payload_at = h.payload_at;
    if (check_tables(i, f->blk_off, f->raw_len)) return fail(MDZ_ERR_INVALID);
    for (uint64_t c = 0; c < n; ++c)
        f->max_chunk = std::max(f->max_chunk, chunk_end(f->blk_off, n, i.payload_bytes, c) - f->blk_off[c]);
    *out = f;
    return MDZ_OK;
}

int mdz_file_info(mdz_file_t f, mdz_file_info_t* info) {
    if (!f || !info) return MDZ_ERR_INVALID;
    *info = f->info;
    return MDZ_OK;
}

int mdz_file_raw_len(mdz_file_t f, uint64_t* raw_len) {
    if (!f || !raw_len) return MDZ_ERR_INVALID;
    if (f->info.codec != MDZ_CODEC_DEFLATE) return MDZ_ERR_UNSUPPORTED;
    memcpy(raw_len, f->raw_len, 8 * f->info.n_chunks);
    return MDZ_OK;
}

int mdz_file_decode(mdz_file_t f, void* dst, uint64_t dst_bytes, void* status, uint64_t stage_bytes,
                    void* stream) {
    if (!f || (!dst && f->info.raw_bytes)) return MDZ_ERR_INVALID;
    const mdz_file_info_t& i = f->info;
    if (i.codec == MDZ_CODEC_DEFLATE && (!status || dst_bytes < i.raw_bytes)) return MDZ_ERR_INVALID;
    const uint64_t stage = stage_bytes ? stage_bytes : (64ull << 20);
    int rc = device_setup(f, stage);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint8_t* out = static_cast<uint8_t*>(dst);
    const uint64_t n = i.n_chunks;
    hipError_t e;
    if (i.codec == MDZ_CODEC_DEFLATE) {            // stream descriptors for this dst
        mdz_inflate_in_t* d = static_cast<mdz_inflate_in_t*>(malloc(sizeof(mdz_inflate_in_t) * n));
        if (!d) return MDZ_ERR_NOMEM;
        uint64_t o = 0;
        for (uint64_t c = 0; c < n; ++c) {
            d[c].src = static_cast<uint8_t*>(f->d_pay) + f->blk_off[c];
            d[c].src_bytes = chunk_end(f->blk_off, n, i.payload_bytes, c) - f->blk_off[c];
            d[c].dst = out + o;
            d[c].dst_bytes = f->raw_len[c];
            o += f->raw_len[c];
        }
        // the previous decode's launches may still read the old descriptors
        e = hipStreamSynchronize(s);
        if (e == hipSuccess) e = hipMemcpy(f->d_desc, d, sizeof(mdz_inflate_in_t) * n, hipMemcpyHostToDevice);
        free(d);
        if (e != hipSuccess) return mdz::record_hip(e, "file: descriptors");
    }
    // a previous decode on `stream` may still read the device payload that
    // this one's copies overwrite
    if ((e = hipEventRecord(f->start, s)) != hipSuccess || (e = hipStreamWaitEvent(f->copy, f->start, 0)) != hipSuccess)
        return mdz::record_hip(e, "file: order");
    uint64_t c0 = 0, piece = 0;
    while (c0 < n) {
        // the piece: chunks [c0, c1) whose payload fits one staging buffer
        const uint64_t beg = f->blk_off[c0];
        uint64_t c1 = c0 + 1;
        while (c1 < n && chunk_end(f->blk_off, n, i.payload_bytes, c1) - beg <= f->pinned_bytes) ++c1;
        const uint64_t end = chunk_end(f->blk_off, n, i.payload_bytes, c1 - 1);
        const int k = (int)(piece & 1);
        // the staging buffer's previous copy must have left it
        if ((e = hipEventSynchronize(f->landed[k])) != hipSuccess) return mdz::record_hip(e, "file: stage wait");
        if ((rc = stage_read(f->fd, f->pinned[k], f->payload_at + beg, end - beg))) return rc;
        if ((e = hipMemcpyAsync(static_cast<uint8_t*>(f->d_pay) + beg, f->pinned[k], end - beg,
                                hipMemcpyHostToDevice, f->copy)) != hipSuccess ||
            (e = hipEventRecord(f->landed[k], f->copy)) != hipSuccess ||
            (e = hipStreamWaitEvent(s, f->landed[k], 0)) != hipSuccess)
            return mdz::record_hip(e, "file: stage copy");
        if ((rc = decode_range(f, c0, c1, out, dst_bytes, status, s))) return rc;
        c0 = c1;
        ++piece;
    }
    return MDZ_OK;
}

int mdz_file_close(mdz_file_t f) {
    delete f;
    return MDZ_OK;
}

}  // extern "C"
