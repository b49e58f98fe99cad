/* Deterministic synthetic workloads (configs 1-4 of the benchmark).
 *
 * All sampling uses mdz_rng (splitmix64-seeded xoshiro256**).  Integer
 * ranges use mdz_rng_below; geometric lengths count Bernoulli trials, so no
 * floating point enters the integer workloads. */
#include "mdz_host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

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


static double std_normal(mdz_rng* r) {
    double u1 = mdz_rng_unit(r), u2 = mdz_rng_unit(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Gamma(a) for a < 1: Marsaglia-Tsang for Gamma(a+1), times U^(1/a). */
static double gamma_small(mdz_rng* r, double a) {
    const double d = a + 1.0 - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    double g;
    for (;;) {
        const double z = std_normal(r);
        double v = 1.0 + c * z;
        if (v <= 0) continue;
        v = v * v * v;
        const double u = mdz_rng_unit(r);
        if (u < 1.0 - 0.0331 * z * z * z * z || log(u) < 0.5 * z * z + d * (1.0 - v + log(v))) {
            g = d * v;
            break;
        }
    }
    double u = mdz_rng_unit(r);
    if (u < 1e-300) u = 1e-300;
    return g * pow(u, 1.0 / a);
}

/* 64-symbol order-2 Markov text with rows ~ Dirichlet(0.3), mixed with
 * copies: at each step, with probability copy_p (once there is history) copy
 * L ~ U{4..256} bytes from distance D ~ U{1..min(4096, pos)}, byte by byte
 * (overlap allowed); otherwise emit one Markov symbol. */
enum { SYM = 64, ROWS = SYM * SYM };

double* mdz_markov_table(uint64_t seed) {
    mdz_rng r;
    mdz_rng_seed(&r, seed);
    double* cdf = (double*)malloc(sizeof(double) * ROWS * SYM);
    for (int row = 0; row < ROWS; ++row) {
        double* c = cdf + row * SYM, tot = 0;
        for (int k = 0; k < SYM; ++k) { c[k] = gamma_small(&r, 0.3); tot += c[k]; }
        double acc = 0;
        for (int k = 0; k < SYM; ++k) { acc += c[k] / tot; c[k] = acc; }
        c[SYM - 1] = 2.0;                            /* the last symbol catches rounding */
    }
    return cdf;
}

void mdz_markov_gen(const double* cdf, uint8_t* out, uint64_t n, uint64_t seed, double copy_p) {
    mdz_rng r;
    mdz_rng_seed(&r, seed);
    uint64_t pos = 0;
    while (pos < n) {
        if (pos > 0 && mdz_rng_unit(&r) < copy_p) {
            uint64_t len = 4 + mdz_rng_below(&r, 253);
            const uint64_t dmax = pos < 4096 ? pos : 4096;
            const uint64_t d = 1 + mdz_rng_below(&r, dmax);
            if (len > n - pos) len = n - pos;
            for (uint64_t i = 0; i < len; ++i, ++pos) out[pos] = out[pos - d];
            continue;
        }
        const uint32_t a = pos >= 2 ? out[pos - 2] : 0, b = pos >= 1 ? out[pos - 1] : 0;
        const double* c = cdf + (a * SYM + b) * SYM;
        const double u = mdz_rng_unit(&r);
        int lo = 0, hi = SYM - 1;                    /* first k with c[k] > u */
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (c[mid] > u) hi = mid; else lo = mid + 1;
        }
        out[pos++] = (uint8_t)lo;
    }
}

/* Config 3: the Markov text with copy probability 0.4; the table and the
 * sequence use seeds derived from `seed`. */
void mdz_gen_lzss_markov(uint8_t* out, uint64_t n, uint64_t seed) {
    double* cdf = mdz_markov_table(seed);
    mdz_markov_gen(cdf, out, n, seed ^ 0x9e3779b97f4a7c15ull, 0.4);
    free(cdf);
}

/* --------------------------------------------------------- config 4 ------ */

typedef struct {
    const double* cdf; uint64_t seed;
    const uint64_t* raw_off; const uint64_t* raw_len; const double* copy_p; const uint8_t* mode;
    uint8_t* raw; uint8_t** comp; uint64_t* comp_len;
} gen_ctx;

static void gen_task(void* v, uint64_t i) {
    gen_ctx* x = (gen_ctx*)v;
    uint8_t* r = x->raw + x->raw_off[i];
    const uint64_t n = x->raw_len[i];
    mdz_markov_gen(x->cdf, r, n, x->seed + 0x100000001b3ull * (i + 1), x->copy_p[i]);
    z_stream z;
    memset(&z, 0, sizeof z);
    const int level = x->mode[i];
    deflateInit2(&z, level, Z_DEFLATED, 15, 8, level == 1 ? Z_FIXED : Z_DEFAULT_STRATEGY);
    const uLong bound = deflateBound(&z, (uLong)n);
    uint8_t* out = (uint8_t*)malloc(bound + 16);
    z.next_in = r;
    z.avail_in = (uInt)n;
    z.next_out = out;
    z.avail_out = (uInt)bound;
    deflate(&z, Z_FINISH);
    x->comp_len[i] = z.total_out;
    deflateEnd(&z);
    x->comp[i] = out;
}

/* Config 4: `count` zlib streams.  Sizes floor(exp(U[ln min, ln max])) in
 * stream order; copy rate U[0.1, 0.7] per stream; modes by a seeded shuffle:
 * 80% level 6 (dynamic Huffman), 10% level 1 with Z_FIXED (fixed Huffman),
 * 10% level 0 (stored).  Streams are concatenated with each start aligned to
 * 16 bytes.  Outputs are allocated here (free with mdz_free). */
int mdz_gen_deflate_batch(uint64_t count, uint64_t seed, uint64_t min_size, uint64_t max_size,
                          uint8_t** src, uint64_t** src_off, uint64_t** src_len, uint8_t** raw,
                          uint64_t** raw_off, uint64_t** raw_len, uint8_t** mode, int threads) {
    if (!count || !min_size || max_size < min_size) return -1;
    mdz_rng r;
    mdz_rng_seed(&r, seed);
    uint64_t* rl = (uint64_t*)malloc(count * 8);
    uint64_t* ro = (uint64_t*)malloc((count + 1) * 8);
    double* cp = (double*)malloc(count * sizeof(double));
    uint8_t* md = (uint8_t*)malloc(count);
    const double lo = log((double)min_size), hi = log((double)max_size);
    ro[0] = 0;
    for (uint64_t i = 0; i < count; ++i) {
        uint64_t s = (uint64_t)floor(exp(lo + (hi - lo) * mdz_rng_unit(&r)));
        if (s < min_size) s = min_size;
        if (s > max_size) s = max_size;
        rl[i] = s;
        ro[i + 1] = ro[i] + s;
        cp[i] = 0.1 + 0.6 * mdz_rng_unit(&r);
    }
    const uint64_t n_fixed = (count + 5) / 10, n_stored = count / 10;
    for (uint64_t i = 0; i < count; ++i) md[i] = i < n_fixed ? 1 : i < n_fixed + n_stored ? 0 : 6;
    for (uint64_t i = count - 1; i > 0; --i) {      /* Fisher-Yates */
        const uint64_t j = mdz_rng_below(&r, i + 1);
        const uint8_t t = md[i]; md[i] = md[j]; md[j] = t;
    }
    double* cdf = mdz_markov_table(seed);
    uint8_t* rw = (uint8_t*)malloc(ro[count] + 1);
    uint8_t** comp = (uint8_t**)calloc(count, sizeof(uint8_t*));
    uint64_t* cl = (uint64_t*)malloc(count * 8);
    gen_ctx x = {cdf, seed, ro, rl, cp, md, rw, comp, cl};
    mdz_parallel_for(count, threads, gen_task, &x);
    uint64_t* so = (uint64_t*)malloc((count + 1) * 8);
    so[0] = 0;
    for (uint64_t i = 0; i < count; ++i) so[i + 1] = (so[i] + cl[i] + 15) & ~15ull;
    uint8_t* s = (uint8_t*)calloc(so[count] + 16, 1);
    for (uint64_t i = 0; i < count; ++i) {
        memcpy(s + so[i], comp[i], cl[i]);
        free(comp[i]);
    }
    free(comp);
    free(cdf);
    free(cp);
    *src = s; *src_off = so; *src_len = cl;
    *raw = rw; *raw_off = ro; *raw_len = rl; *mode = md;
    return 0;
}

