// Stream headers and trailers around Deflate data: gzip header length,
// Adler-32 and CRC-32 arithmetic for wave-parallel trailer checks.
#pragma once

#include "mdz_device.h"

namespace mdz {

// gzip member header length (RFC 1952), or -1 when the header is rejected.
// Rules follow cuDF's parse_gzip_header (reference include/cudf_deflate.cuh:
// 1179-1217): at least 18 bytes, ID1 ID2 CM = 1f 8b 08, FEXTRA / FNAME /
// FCOMMENT / FHCRC skipped in that order with the same bounds checks, and at
// least one byte of Deflate data left before the 8-byte footer.  One
// difference: FEXTRA skips its 2-byte XLEN field as well as XLEN bytes
// (RFC 1952 2.3.1.1); cuDF skips XLEN bytes only, which starts the Deflate
// data two bytes early.  The reserved FLG bits are not checked (as cuDF).
__host__ __device__ inline int64_t gzip_header_len(const uint8_t* s, uint64_t n) {
    if (n < 18) return -1;
    if (s[0] != 0x1f || s[1] != 0x8b || s[2] != 0x08) return -1;
    const uint32_t flg = s[3];
    uint64_t h = 10;
    if (flg & 0x04) {                                    // FEXTRA
        h += 2 + ((uint32_t)s[10] | ((uint32_t)s[11] << 8));
        if (h >= n) return -1;
    }
    for (uint32_t bit = 0x08; bit <= 0x10; bit <<= 1) {  // FNAME, then FCOMMENT
        if (!(flg & bit)) continue;
        for (;;) {
            if (h >= n) return -1;
            if (s[h++] == 0) break;
        }
    }
    if (flg & 0x02) h += 2;                              // FHCRC (not verified)
    if (h + 8 >= n) return -1;
    return (int64_t)h;
}

// CRC-32 (ISO-HDLC: reflected, polynomial 0x04C11DB7) arithmetic in GF(2)[x]
// modulo P.  A residue is held reflected: bit 31 is the x^0 coefficient and
// bit 0 the x^31 one, so multiplying by x is a right shift that folds the
// x^32 term back in with 0xEDB88320.
constexpr uint32_t kCrcPoly = 0xEDB88320u;

__host__ __device__ constexpr uint32_t crc_times_x(uint32_t a) { return (a >> 1) ^ ((a & 1u) ? kCrcPoly : 0u); }

// a * b mod P: b * x^i is accumulated for every x^i term of a.
__host__ __device__ constexpr uint32_t crc_mulmod(uint32_t a, uint32_t b) {
    uint32_t p = 0;
    for (int i = 0; i < 32; ++i) {
        if (a & (0x80000000u >> i)) p ^= b;
        b = crc_times_x(b);
    }
    return p;
}

__host__ __device__ constexpr uint32_t crc_xpow(uint64_t e) {        // x^e mod P, square-and-multiply
    uint32_t r = 0x80000000u, sq = crc_times_x(0x80000000u);
    while (e) {
        if (e & 1) r = crc_mulmod(r, sq);
        sq = crc_mulmod(sq, sq);
        e >>= 1;
    }
    return r;
}

// Register update for one dword of message, four bytes low byte first:
// (r + w) * x^32 mod P (the shift-register form of the bytewise CRC).
__host__ __device__ inline uint32_t crc_dword(uint32_t r, uint32_t w) {
    r ^= w;
#pragma unroll
    for (int i = 0; i < 32; ++i) r = crc_times_x(r);
    return r;
}

constexpr uint32_t kAdlerMod = 65521u;

}  // namespace mdz
