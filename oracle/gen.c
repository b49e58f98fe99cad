/* mdz oracle: deterministic synthetic workloads.  TEST INFRASTRUCTURE ONLY.
 *
 * All sampling uses mdz_rng (splitmix64-seeded xoshiro256**).  Integer
 * ranges use mdz_rng_below; geometric lengths count Bernoulli trials, so no
 * floating point enters the integer workloads. */
#include "mdz_oracle.h"

/* Geometric on {1,2,...} with success probability 1/q. */
static uint64_t geom(mdz_rng* r, uint64_t q) {
    uint64_t k = 1;
    while (mdz_rng_below(r, q) != 0) ++k;
    return k;
}

/* RLE v1 column of u32: i.i.d. segments.  With p = 0.7 a run: length
 * clamp(Geom(1/16), 3, 130), base ~ U[0, 2^20), delta = 0 with p = 0.7 else
 * U{-128..127}, value_i = (u32)(base + i*delta).  Otherwise a literal segment:
 * length U{1..128}, values U[0, 2^20).  Truncated at n. */
void mdz_gen_rlev1_u32(uint32_t* out, uint64_t n, uint64_t seed) {
    mdz_rng r;
    mdz_rng_seed(&r, seed);
    uint64_t i = 0;
    while (i < n) {
        if (mdz_rng_below(&r, 10) < 7) {
            uint64_t len = geom(&r, 16);
            if (len < 3) len = 3;
            if (len > 130) len = 130;
            int64_t base = (int64_t)mdz_rng_below(&r, 1u << 20);
            int64_t d = mdz_rng_below(&r, 10) < 7 ? 0 : (int64_t)mdz_rng_below(&r, 256) - 128;
            for (uint64_t k = 0; k < len && i < n; ++k, ++i)
                out[i] = (uint32_t)(base + (int64_t)k * d);
        } else {
            uint64_t len = 1 + mdz_rng_below(&r, 128);
            for (uint64_t k = 0; k < len && i < n; ++k, ++i)
                out[i] = (uint32_t)mdz_rng_below(&r, 1u << 20);
        }
    }
}

/* Byte runs: each run is a fresh byte different from the previous run's,
 * length Geom(1/12).  Truncated at n. */
void mdz_gen_brle_u8(uint8_t* out, uint64_t n, uint64_t seed) {
    mdz_rng r;
    mdz_rng_seed(&r, seed);
    uint64_t i = 0;
    int prev = -1;
    while (i < n) {
        int b = (int)mdz_rng_below(&r, prev < 0 ? 256 : 255);
        if (prev >= 0 && b >= prev) ++b;
        prev = b;
        uint64_t len = geom(&r, 12);
        for (uint64_t k = 0; k < len && i < n; ++k, ++i) out[i] = (uint8_t)b;
    }
}

/* RLE v2 int64 column: i.i.d. segments weighted 15/35/35/15.
 *  SHORT_REPEAT: length U{3..10}, value U[0, 2^32).
 *  DELTA: length U{10..512}, start U[2^20, 2^40), stride U{-100..100}.
 *  DIRECT: length U{1..512}, width w in {1,2,4,8,12,16,24,32,64}, values U[0, 2^w).
 *  PATCHED_BASE: length U{16..128}, base width b U{8..16}, values U[0, 2^b),
 *    each replaced with p = 0.05 by an outlier of width o U{40..60},
 *    value U[2^(o-1), 2^o).
 * Truncated at n. */
void mdz_gen_rlev2_i64(uint64_t* out, uint64_t n, uint64_t seed) {
    static const uint32_t widths[9] = {1, 2, 4, 8, 12, 16, 24, 32, 64};
    mdz_rng r;
    mdz_rng_seed(&r, seed);
    uint64_t i = 0;
    while (i < n) {
        const uint64_t kind = mdz_rng_below(&r, 100);
        if (kind < 15) {
            const uint64_t len = 3 + mdz_rng_below(&r, 8);
            const uint64_t v = mdz_rng_below(&r, 1ull << 32);
            for (uint64_t k = 0; k < len && i < n; ++k, ++i) out[i] = v;
        } else if (kind < 50) {
            const uint64_t len = 10 + mdz_rng_below(&r, 503);
            const uint64_t start = (1ull << 20) + mdz_rng_below(&r, (1ull << 40) - (1ull << 20));
            const int64_t stride = (int64_t)mdz_rng_below(&r, 201) - 100;
            for (uint64_t k = 0; k < len && i < n; ++k, ++i)
                out[i] = (uint64_t)((int64_t)start + (int64_t)k * stride);
        } else if (kind < 85) {
            const uint64_t len = 1 + mdz_rng_below(&r, 512);
            const uint32_t w = widths[mdz_rng_below(&r, 9)];
            for (uint64_t k = 0; k < len && i < n; ++k, ++i)
                out[i] = w == 64 ? mdz_rng_next(&r) : mdz_rng_below(&r, 1ull << w);
        } else {
            const uint64_t len = 16 + mdz_rng_below(&r, 113);
            const uint32_t b = 8 + (uint32_t)mdz_rng_below(&r, 9);
            for (uint64_t k = 0; k < len && i < n; ++k, ++i) {
                if (mdz_rng_below(&r, 100) < 5) {
                    const uint32_t o = 40 + (uint32_t)mdz_rng_below(&r, 21);
                    out[i] = (1ull << (o - 1)) + mdz_rng_below(&r, 1ull << (o - 1));
                } else {
                    out[i] = mdz_rng_below(&r, 1ull << b);
                }
            }
        }
    }
}
