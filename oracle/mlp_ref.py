"""CPU restatement of the decode-cost MLP (TEST INFRASTRUCTURE ONLY).

float64 numpy forward and backward for the parameter layout of
compression_amd/csrc/mlp.hip: W1[F][H] b1[H] W2[H][H] b2[H] w3[H] b3, ReLU
hidden layers, loss mean((y - t)^2).  The tests pin it with central finite
differences and check the HIP gradients against it.
"""
from __future__ import annotations

import numpy as np

F, H = 11, 32
P = F * H + H + H * H + H + H + 1


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    return h2 @ w3 + b3


def loss(p, x, t):
    r = forward(p, x) - np.asarray(t, np.float64)
    return float(np.mean(r * r))


def grad(p, x, t):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    a1 = x @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ w3 + b3
    dy = 2.0 * (y - t) / n
    g_b3 = dy.sum()
    g_w3 = h2.T @ dy
    d2 = np.outer(dy, w3) * (a2 > 0)
    g_b2 = d2.sum(0)
    g_W2 = h1.T @ d2
    d1 = (d2 @ W2.T) * (a1 > 0)
    g_b1 = d1.sum(0)
    g_W1 = x.T @ d1
    return np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])


def init(seed: int, target_mean: float = 0.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = np.zeros(P)
    o = 0
    p[o:o + F * H] = rng.normal(0, np.sqrt(2.0 / F), F * H); o += F * H
    p[o:o + H] = 0.1; o += H
    p[o:o + H * H] = rng.normal(0, np.sqrt(2.0 / H), H * H); o += H * H
    p[o:o + H] = 0.1; o += H
    p[o:o + H] = rng.normal(0, np.sqrt(1.0 / H), H); o += H
    p[o] = target_mean
    return p


# ---------------------------------------------------------------- features --
_ORDER = [16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15]


def _len_extra(s):
    return 0 if s < 265 or s == 285 else (s - 261) // 4


def _dist_extra(d):
    return 0 if d < 4 else (d - 2) // 2


def features(stream: bytes, dst_bytes: int, zlib_hdr: bool = True) -> np.ndarray:
    """The 11 scheduler features of one Deflate stream, restated from the
    format (RFC 1951 3.2.3-3.2.7) for checking the GPU feature kernel: sizes,
    first block type, and what a first dynamic (or fixed) block's code
    lengths imply -- match share p = sum over length codes of 2^-L divided by
    the sum over all codes, bits per symbol = E[L] + p * (E[length extra] +
    E[distance code + extra]), symbols = 8 * bytes / bits per symbol, matches
    = symbols * p, blocks = symbols / 16384 (stored: bytes / 65535)."""
    x = np.zeros(F)
    n = len(stream)
    x[0], x[1] = np.log2(1 + n), np.log2(1 + dst_bytes)
    x[2] = x[1] - x[0]
    at = 2 if zlib_hdr else 0
    if n <= at:
        return x
    data = stream[at:]
    pos = [0]

    def get(k):
        v = 0
        for i in range(k):
            b = pos[0] >> 3
            if b >= len(data):
                raise EOFError
            v |= ((data[b] >> (pos[0] & 7)) & 1) << i
            pos[0] += 1
        return v
    try:
        get(1)
        btype = get(2)
    except EOFError:
        return x
    if btype == 3:
        return x
    x[3 + btype] = 1
    nbytes = float(len(data))
    if btype == 0:
        x[9] = np.log2(1 + nbytes / 65535)
        return x
    if btype == 1:
        lit = [8] * 144 + [9] * 112 + [7] * 24 + [8] * 8
        dist = [5] * 30
    else:
        try:
            nlen, ndist, ncode = get(5) + 257, get(5) + 1, get(4) + 4
            cl = [0] * 19
            for k in range(ncode):
                cl[_ORDER[k]] = get(3)
            if nlen > 286 or ndist > 30:
                return x
            codes, code, nxt = {}, 0, {}
            cnt = [0] * 8
            for v in cl:
                if v:
                    cnt[v] += 1
            for L in range(1, 8):
                code = (code + cnt[L - 1]) << 1 if L > 1 else 0
                nxt[L] = code
            for s, L in enumerate(cl):
                if L:
                    codes[(L, nxt[L])] = s
                    nxt[L] += 1
            lens, prev = [], 0
            while len(lens) < nlen + ndist:
                c, s = 0, None
                for L in range(1, 8):
                    c = (c << 1) | get(1)
                    if (L, c) in codes:
                        s = codes[(L, c)]
                        break
                if s is None:
                    return x
                if s < 16:
                    lens.append(s)
                    prev = s
                elif s == 16:
                    if not lens:
                        return x
                    lens += [prev] * (3 + get(2))
                elif s == 17:
                    lens += [0] * (3 + get(3))
                    prev = 0
                else:
                    lens += [0] * (11 + get(7))
                    prev = 0
            if len(lens) > nlen + ndist:
                return x
        except EOFError:
            return x
        lit, dist = lens[:nlen], lens[nlen:]
    w = sum(2.0 ** -L for L in lit if L)
    if w <= 0:
        return x
    wm = sum(2.0 ** -L for s, L in enumerate(lit) if L and s >= 257)
    eb = sum(2.0 ** -L * L for L in lit if L)
    ex = sum(2.0 ** -L * _len_extra(s) for s, L in enumerate(lit) if L and s >= 257)
    wd = sum(2.0 ** -L for L in dist if L)
    ed = sum(2.0 ** -L * (L + _dist_extra(d)) for d, L in enumerate(dist) if L)
    p = wm / w
    bps = eb / w + p * ((ex / wm if wm > 0 else 0) + (ed / wd if wd > 0 else 0))
    if bps <= 0:
        return x
    syms = 8 * nbytes / bps
    x[6], x[7], x[8] = np.log2(1 + syms), p, np.log2(1 + syms * p)
    x[9], x[10] = np.log2(1 + syms / 16384), bps / 16
    return x
