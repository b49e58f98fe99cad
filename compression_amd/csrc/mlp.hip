// Decode-cost model for the batched scheduler: a small MLP (F -> H -> H -> 1,
// ReLU) with hand-written forward, backward and Adam kernels, fp32.
//
// Features come from what a scheduler knows before decoding a Deflate stream
// (its descriptor and first block header, including what its code lengths
// imply); the target is log2 of the cycles the decoder logged for it
// (mdz_inflate_cost_t).  The model only orders
// work: decoded bytes never pass through it.
//
// Parameter layout (floats): W1[F][H] b1[H] W2[H][H] b2[H] w3[H] b3.
// Training is full-batch: backward stages each sample's activations and
// deltas in LDS, sums every parameter's outer-product terms over the
// workgroup's samples (one thread per parameter) and adds them with one
// atomic per parameter and block; Adam then updates in place.
#include "deflate_hdr_dev.h"
#include "mdz_device.h"
#include "mdz_internal.h"

namespace mdz {
namespace {

__constant__ uint8_t kOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

constexpr int F = MDZ_MLP_FEATURES, H = MDZ_MLP_HIDDEN;
constexpr int P_W1 = 0, P_B1 = F * H, P_W2 = P_B1 + H, P_B2 = P_W2 + H * H, P_W3 = P_B2 + H,
              P_B3 = P_W3 + H, P_N = P_B3 + 1;
static_assert(P_N == MDZ_MLP_PARAMS, "parameter count");
constexpr int kBlock = 256;

// A scheduler's view of a Deflate stream before decoding it: its sizes and
// its first block header.  A dynamic header's code lengths imply a symbol
// distribution (a code of length L stands for probability 2^-L), which gives
// the share of match symbols and the expected bits per symbol; with the
// compressed size these estimate how many symbols, matches and blocks the
// decoder will meet (zlib closes a block every 16,384 symbols at its default
// memory level).  Fixed blocks use the fixed code's lengths; stored streams
// have no symbols and a block per 65,535 bytes.
struct HeaderBits {                     // LSB-first reader over src[0, n)
    const uint8_t* p;
    uint64_t n, pos;
    bool bad;
    MDZ_DEV uint32_t get(uint32_t k) {
        uint32_t v = 0;
        for (uint32_t i = 0; i < k; ++i, ++pos) {
            const uint64_t b = pos >> 3;
            if (b >= n) { bad = true; return 0; }
            v |= (uint32_t)((p[b] >> (pos & 7)) & 1) << i;
        }
        return v;
    }
};

MDZ_DEV uint32_t len_extra_bits(uint32_t s) { return s < 265 || s == 285 ? 0 : (s - 261) / 4; }
MDZ_DEV uint32_t dist_extra_bits(uint32_t d) { return d < 4 ? 0 : (d - 2) / 2; }

// Expected bits per literal/length symbol and match share implied by code
// lengths lit[0..nlen) and dist[0..ndist); false if no code is present.
MDZ_DEV bool implied_stats(const uint8_t* lit, uint32_t nlen, const uint8_t* dist, uint32_t ndist,
                           float* bps, float* pmatch) {
    float w = 0.f, wm = 0.f, eb = 0.f, ex = 0.f, wd = 0.f, ed = 0.f;
    for (uint32_t s = 0; s < nlen; ++s) {
        if (!lit[s]) continue;
        const float q = __builtin_ldexpf(1.0f, -(int)lit[s]);
        w += q;
        eb += q * lit[s];
        if (s >= 257) {
            wm += q;
            ex += q * len_extra_bits(s);
        }
    }
    for (uint32_t d = 0; d < ndist; ++d) {
        if (!dist[d]) continue;
        const float q = __builtin_ldexpf(1.0f, -(int)dist[d]);
        wd += q;
        ed += q * (dist[d] + dist_extra_bits(d));
    }
    if (w <= 0.f) return false;
    *pmatch = wm / w;
    const float per_match = (wm > 0.f ? ex / wm : 0.f) + (wd > 0.f ? ed / wd : 0.f);
    *bps = eb / w + *pmatch * per_match;
    return true;
}

// Features of stream i (MDZ_MLP_FEATURES = 11): log2(1 + src bytes),
// log2(1 + dst bytes), their difference, the first block's type one-hot
// (stored, fixed, dynamic; all zero if unreadable), log2(1 + estimated
// symbols), the implied match share, log2(1 + estimated matches),
// log2(1 + estimated blocks), and the implied bits per symbol / 16.
__global__ void features_kernel(const mdz_inflate_in_t* __restrict__ in, uint32_t n, int parse_hdr,
                                float* __restrict__ x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const mdz_inflate_in_t d = in[i];
    const float ls = log2f(1.0f + (float)d.src_bytes), ld = log2f(1.0f + (float)d.dst_bytes);
    float* xi = x + (size_t)i * F;
    for (int k = 0; k < F; ++k) xi[k] = 0.f;
    xi[0] = ls;
    xi[1] = ld;
    xi[2] = ld - ls;
    const uint8_t* src = reinterpret_cast<const uint8_t*>(d.src);
    const int hdr = parse_hdr & 3;
    const int64_t at = hdr == MDZ_HDR_ZLIB ? 2 : hdr == MDZ_HDR_GZIP ? gzip_header_len(src, d.src_bytes) : 0;
    if (at < 0 || d.src_bytes <= (uint64_t)at) return;
    HeaderBits r{src + at, d.src_bytes - (uint64_t)at, 0, false};
    r.get(1);
    const uint32_t type = r.get(2);
    if (r.bad || type == 3) return;
    xi[3] = type == 0;
    xi[4] = type == 1;
    xi[5] = type == 2;
    const float bytes = (float)(d.src_bytes - (uint64_t)at);
    if (type == 0) {
        xi[9] = log2f(1.0f + bytes / 65535.0f);
        return;
    }
    uint8_t lens[320];
    uint32_t nlen = 288, ndist = 30;
    if (type == 1) {
        for (uint32_t s = 0; s < 288; ++s) lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
        for (uint32_t s = 0; s < 30; ++s) lens[288 + s] = 5;
    } else {
        nlen = r.get(5) + 257;
        ndist = r.get(5) + 1;
        const uint32_t ncode = r.get(4) + 4;
        uint8_t cl[19] = {};
        for (uint32_t k = 0; k < ncode; ++k) cl[kOrder[k]] = (uint8_t)r.get(3);
        if (r.bad || nlen > 286 || ndist > 30) return;
        uint32_t cnt[8] = {}, sym[19], offs[8] = {};
        for (uint32_t s = 0; s < 19; ++s) ++cnt[cl[s]];
        cnt[0] = 0;
        for (uint32_t k = 1; k < 7; ++k) offs[k + 1] = offs[k] + cnt[k];
        for (uint32_t s = 0; s < 19; ++s)
            if (cl[s]) sym[offs[cl[s]]++] = s;
        uint32_t idx = 0, prev = 0;
        while (idx < nlen + ndist) {
            int code = 0, first = 0, index = 0, s = -1;      // canonical decode
            for (uint32_t len = 1; len < 8; ++len) {
                code |= (int)r.get(1);
                const int c = (int)cnt[len];
                if (code - c < first) { s = (int)sym[index + (code - first)]; break; }
                index += c;
                first = (first + c) << 1;
                code <<= 1;
            }
            if (s < 0 || r.bad) return;
            uint32_t rep = 1, val = (uint32_t)s;
            if (s == 16) { if (idx == 0) return; val = prev; rep = 3 + r.get(2); }
            else if (s == 17) { val = 0; rep = 3 + r.get(3); }
            else if (s == 18) { val = 0; rep = 11 + r.get(7); }
            if (r.bad || idx + rep > nlen + ndist) return;
            for (uint32_t k = 0; k < rep; ++k, ++idx) lens[idx < nlen ? idx : 288 + (idx - nlen)] = (uint8_t)val;
            prev = val;
        }
    }
    float bps = 0.f, pm = 0.f;
    if (!implied_stats(lens, nlen, lens + 288, ndist, &bps, &pm) || bps <= 0.f) return;
    const float syms = 8.0f * bytes / bps;
    xi[6] = log2f(1.0f + syms);
    xi[7] = pm;
    xi[8] = log2f(1.0f + syms * pm);
    xi[9] = log2f(1.0f + syms / 16384.0f);
    xi[10] = bps / 16.0f;
}

MDZ_DEV void forward_one(const float* __restrict__ p, const float* xi, float* h1, float* h2, float& y) {
#pragma unroll 4
    for (int j = 0; j < H; ++j) {
        float a = p[P_B1 + j];
#pragma unroll
        for (int k = 0; k < F; ++k) a += xi[k] * p[P_W1 + k * H + j];
        h1[j] = a > 0.f ? a : 0.f;
    }
    float out = p[P_B3];
#pragma unroll 4
    for (int j = 0; j < H; ++j) {
        float a = p[P_B2 + j];
        for (int k = 0; k < H; ++k) a += h1[k] * p[P_W2 + k * H + j];
        h2[j] = a > 0.f ? a : 0.f;
        out += h2[j] * p[P_W3 + j];
    }
    y = out;
}

__global__ __launch_bounds__(kBlock) void forward_kernel(const float* __restrict__ p,
                                                         const float* __restrict__ x, uint32_t n,
                                                         float* __restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float xi[F], h1[H], h2[H], out;
#pragma unroll
    for (int k = 0; k < F; ++k) xi[k] = x[(size_t)i * F + k];
    forward_one(p, xi, h1, h2, out);
    y[i] = out;
}

// grad += d/dparams of (1/n) * sum_i (y_i - t_i)^2; loss_sum += sum (y - t)^2.
// Phase 1: one thread per sample runs the forward pass and the backward
// deltas and leaves x, h1, h2, d1 (dL/d pre-activation 1), d2 and dy in LDS.
// Phase 2: each thread owns parameters and sums their outer-product terms
// over the block's samples (W1 = x^T d1, W2 = h1^T d2, w3 = h2^T dy, biases
// the column sums), then adds them to grad with one atomic per parameter.
constexpr int kBwd = 96;                        // samples (and threads) per block (52 KB LDS)
__global__ __launch_bounds__(kBwd) void backward_kernel(const float* __restrict__ p,
                                                        const float* __restrict__ x,
                                                        const float* __restrict__ t, uint32_t n,
                                                        float* __restrict__ grad,
                                                        float* __restrict__ loss_sum) {
    __shared__ float xs[kBwd][F], h1s[kBwd][H], h2s[kBwd][H], d1s[kBwd][H], d2s[kBwd][H], dys[kBwd];
    __shared__ float ls;
    const uint32_t tid = threadIdx.x, i = blockIdx.x * kBwd + tid;
    if (tid == 0) ls = 0.f;
    __syncthreads();
    {
        float xi[F], h1[H], h2[H], y = 0.f, dy = 0.f;
#pragma unroll
        for (int k = 0; k < F; ++k) xi[k] = i < n ? x[(size_t)i * F + k] : 0.f;
        forward_one(p, xi, h1, h2, y);
        if (i < n) {
            const float r = y - t[i];
            dy = 2.f * r / (float)n;
            atomicAdd(&ls, r * r);
        }
        float d2[H];
#pragma unroll
        for (int j = 0; j < H; ++j) {
            d2[j] = h2[j] > 0.f ? dy * p[P_W3 + j] : 0.f;
            d2s[tid][j] = d2[j];
            h2s[tid][j] = i < n ? h2[j] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < H; ++k) {
            float d1 = 0.f;
#pragma unroll
            for (int j = 0; j < H; ++j) d1 += p[P_W2 + k * H + j] * d2[j];
            d1s[tid][k] = h1[k] > 0.f ? d1 : 0.f;
            h1s[tid][k] = h1[k];
        }
#pragma unroll
        for (int k = 0; k < F; ++k) xs[tid][k] = xi[k];
        dys[tid] = dy;
    }
    __syncthreads();
    const uint32_t ns = min((uint32_t)kBwd, n - blockIdx.x * kBwd);
    for (int k = tid; k < P_N; k += kBwd) {
        float g = 0.f;
        if (k < P_B1) {                                   // W1[m][j]
            const int m = k / H, j = k % H;
            for (uint32_t s = 0; s < ns; ++s) g += xs[s][m] * d1s[s][j];
        } else if (k < P_W2) {                            // b1[j]
            const int j = k - P_B1;
            for (uint32_t s = 0; s < ns; ++s) g += d1s[s][j];
        } else if (k < P_B2) {                            // W2[a][b]
            const int a = (k - P_W2) / H, b = (k - P_W2) % H;
            for (uint32_t s = 0; s < ns; ++s) g += h1s[s][a] * d2s[s][b];
        } else if (k < P_W3) {                            // b2[b]
            const int b = k - P_B2;
            for (uint32_t s = 0; s < ns; ++s) g += d2s[s][b];
        } else if (k < P_B3) {                            // w3[j]
            const int j = k - P_W3;
            for (uint32_t s = 0; s < ns; ++s) g += dys[s] * h2s[s][j];
        } else {                                          // b3
            for (uint32_t s = 0; s < ns; ++s) g += dys[s];
        }
        if (g != 0.f) atomicAdd(&grad[k], g);
    }
    if (tid == 0 && loss_sum) atomicAdd(loss_sum, ls);
}

__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ grad, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1t, float b2t) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P_N) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gk = grad[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk, vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / (1.f - b1t)) / (sqrtf(vk / (1.f - b2t)) + eps);
    grad[k] = 0.f;
}

__global__ void zero_kernel(float* a, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = 0.f;
}

}  // namespace

int mlp_features(const mdz_inflate_in_t* in, uint32_t n, int parse_hdr, float* x, hipStream_t s) {
    if (!inflate_hdr_ok(parse_hdr)) return MDZ_ERR_INVALID;
    if (!n) return MDZ_OK;
    if (!in || !x) return MDZ_ERR_INVALID;
    hipLaunchKernelGGL(features_kernel, dim3((n + 255) / 256), dim3(256), 0, s, in, n, parse_hdr, x);
    return record_hip(hipGetLastError(), "mlp features");
}

int mlp_forward(const float* params, const float* x, uint32_t n, float* y, hipStream_t s) {
    if (!n) return MDZ_OK;
    if (!params || !x || !y) return MDZ_ERR_INVALID;
    hipLaunchKernelGGL(forward_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, params, x,
                       n, y);
    return record_hip(hipGetLastError(), "mlp forward");
}

int mlp_grad(const float* params, const float* x, const float* t, uint32_t n, float* grad,
             float* loss_sum, hipStream_t s) {
    if (!params || !x || !t || !grad) return MDZ_ERR_INVALID;
    hipLaunchKernelGGL(zero_kernel, dim3((P_N + 255) / 256), dim3(256), 0, s, grad, P_N);
    if (loss_sum) hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(1), 0, s, loss_sum, 1);
    if (n)
        hipLaunchKernelGGL(backward_kernel, dim3((n + kBwd - 1) / kBwd), dim3(kBwd), 0, s, params,
                           x, t, n, grad, loss_sum);
    return record_hip(hipGetLastError(), "mlp backward");
}

int mlp_adam(float* params, float* grad, float* m, float* v, float lr, uint32_t step, hipStream_t s) {
    if (!params || !grad || !m || !v || step == 0) return MDZ_ERR_INVALID;
    const float b1t = powf(0.9f, (float)step), b2t = powf(0.999f, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((P_N + 255) / 256), dim3(256), 0, s, params, grad, m, v, lr,
                       b1t, b2t);
    return record_hip(hipGetLastError(), "mlp adam");
}

}  // namespace mdz
