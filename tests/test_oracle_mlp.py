"""CPU reference of the decode-cost MLP: gradients pinned by central finite
differences (float64)."""
import numpy as np

from oracle import mlp_ref as M


def test_layout():
    assert M.P == 11 * 32 + 32 + 32 * 32 + 32 + 32 + 1 == 1473


def test_grad_vs_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init(1, 20.0)
    x = np.column_stack([rng.uniform(12, 22, 64), rng.uniform(12, 22, 64), rng.uniform(-1, 4, 64),
                         np.eye(3)[rng.integers(0, 3, 64)], rng.uniform(0, 16, (64, M.F - 6))])
    t = rng.uniform(15, 28, 64)
    g = M.grad(p, x, t)
    eps = 1e-6
    idx = rng.choice(M.P, 120, replace=False)
    for k in idx:
        e = np.zeros(M.P)
        e[k] = eps
        fd = (M.loss(p + e, x, t) - M.loss(p - e, x, t)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])


def test_loss_zero_at_exact_fit():
    p = M.init(2)
    x = np.random.default_rng(3).normal(size=(10, M.F))
    t = M.forward(p, x)
    assert M.loss(p, x, t) == 0.0 and np.abs(M.grad(p, x, t)).max() == 0.0


def test_features_closed_forms():
    """The scheduler features on streams whose first header is known: a
    fixed block implies a match share of (23 * 2^-7 + 8 * 2^-8) over a
    complete code; a stored stream has no symbols and a block per 65,535
    bytes; a dynamic header written with chosen code lengths implies the
    share and bits per symbol computed from those lengths by hand."""
    import zlib
    from tests import deflate_util as DU
    data = bytes(range(256)) * 40
    c = zlib.compressobj(1, zlib.DEFLATED, 15, 8, zlib.Z_FIXED)
    x = M.features(c.compress(data) + c.flush(), len(data))
    assert x[4] == 1 and x[7] == (23 * 2.0 ** -7 + 8 * 2.0 ** -8)
    s = zlib.compress(data, 0)
    x = M.features(s, len(data))
    assert x[3] == 1 and x[6] == 0 and np.isclose(x[9], np.log2(1 + (len(s) - 2) / 65535))
    # dynamic: literals 0..127 of length 8, 256 and 257..284 (29) and 285 ...
    lit = [0] * 286
    for s_ in range(128):
        lit[s_] = 8
    lit[256] = 8
    for s_ in range(257, 286):
        lit[s_] = 7 if s_ < 287 - 32 else 8
    # make it complete: 129 codes of 8 bits + 29 of 7 or 8 -- pad with literals 128.. at 8 bits
    k = 128
    while sum(2.0 ** -L for L in lit if L) < 1:
        lit[k] = 8
        k += 1
    dist = [5] * 30
    w = DU.BitWriter()
    DU.dynamic_header(w, lit, dist, final=True)
    DU.write_symbols(w, DU.canonical(lit), DU.canonical(dist), [("E",)])
    x = M.features(DU.zlib_wrap(w.tobytes()), 1000)
    W = sum(2.0 ** -L for L in lit if L)
    Wm = sum(2.0 ** -L for s_, L in enumerate(lit) if L and s_ >= 257)
    assert x[5] == 1 and np.isclose(x[7], Wm / W)
    Eb = sum(2.0 ** -L * L for L in lit if L) / W
    Ex = sum(2.0 ** -L * ((s_ - 261) // 4 if 265 <= s_ < 285 else 0) for s_, L in enumerate(lit) if L and s_ >= 257) / Wm
    Ed = sum(2.0 ** -5 * (5 + (d - 2) // 2 if d >= 4 else 5) for d in range(30)) / (30 * 2.0 ** -5)
    assert np.isclose(x[10] * 16, Eb + Wm / W * (Ex + Ed))
