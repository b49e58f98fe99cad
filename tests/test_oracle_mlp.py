"""CPU reference of the decode-cost MLP: gradients pinned by central finite
differences (float64)."""
import numpy as np

from oracle import mlp_ref as M


def test_layout():
    assert M.P == 6 * 32 + 32 + 32 * 32 + 32 + 32 + 1 == 1313


def test_grad_vs_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init(1, 20.0)
    x = np.column_stack([rng.uniform(12, 22, 64), rng.uniform(12, 22, 64), rng.uniform(-1, 4, 64),
                         np.eye(3)[rng.integers(0, 3, 64)]])
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
    x = np.random.default_rng(3).normal(size=(10, 6))
    t = M.forward(p, x)
    assert M.loss(p, x, t) == 0.0 and np.abs(M.grad(p, x, t)).max() == 0.0
