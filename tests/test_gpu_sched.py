"""GPU: decode-cost MLP kernels vs the CPU reference, and the learned
scheduler against the compressed-size heuristic on decoder-logged costs."""
import numpy as np
import pytest

from oracle import mlp_ref as M
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _xt(n, seed):
    rng = np.random.default_rng(seed)
    x = np.column_stack([rng.uniform(12, 22, n), rng.uniform(12, 22, n), rng.uniform(-1, 4, n),
                         np.eye(3)[rng.integers(0, 3, n)]]).astype(np.float32)
    t = rng.uniform(15, 28, n).astype(np.float32)
    return x, t


def test_forward_and_grad_match_cpu(gpu):
    from compression_amd import sched
    p = M.init(4, 20.0).astype(np.float32)
    x, t = _xt(1000, 5)
    model = sched.CostModel(p)
    xd, td = gpu.DeviceBuffer.from_array(x), gpu.DeviceBuffer.from_array(t)
    y = model.predict(xd, len(t))
    ref_y = M.forward(p.astype(np.float64), x)
    assert np.allclose(y, ref_y, rtol=1e-5, atol=1e-3)             # fp32 vs float64 forward
    g, loss = model.grad(xd, td, len(t))
    ref_g = M.grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    scale = np.abs(ref_g).max()
    assert np.abs(g - ref_g).max() <= 1e-4 * scale                  # fp32 sums, any order
    assert abs(loss - M.loss(p, x, t)) <= 1e-4 * M.loss(p, x, t)


def test_adam_step_matches_cpu(gpu):
    from compression_amd import sched
    p = M.init(6, 20.0).astype(np.float32)
    x, t = _xt(300, 7)
    model = sched.CostModel(p)
    xd, td = gpu.DeviceBuffer.from_array(x), gpu.DeviceBuffer.from_array(t)
    model.train(xd, td, len(t), steps=1, lr=1e-2)
    g = M.grad(p.astype(np.float64), x, t)
    m, v = 0.1 * g, 0.001 * g * g
    ref = p - 1e-2 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.allclose(model.params(), ref, rtol=1e-4, atol=1e-5)


def _logged_batch(gpu, count, seed):
    from compression_amd import deflate
    b = O.gen_deflate_batch(count, seed, 4096, 1 << 20, threads=16)
    db = deflate.DeviceBatch(b.src, b.src_off, b.src_len, b.raw_len)
    db.enable_cost_log()
    deflate.inflate_sched(db)
    st = db.statuses()
    assert (st["status"] == 0).all()
    return b, db, db.costs()["cycles"].astype(np.float64)


def test_learned_order_beats_compressed_size(gpu):
    from compression_amd import deflate, sched
    _, train_db, train_cyc = _logged_batch(gpu, 1024, 0x4D445A24)
    xd = sched.features(train_db)
    n = train_db.count
    td = gpu.DeviceBuffer.from_array(np.log2(train_cyc).astype(np.float32))
    model = sched.CostModel(M.init(7, float(np.log2(train_cyc).mean())))
    hist = model.train(xd, td, n, steps=2000, lr=3e-3)
    assert hist[-1] < 0.25 * hist[0]

    test_b, test_db, cyc = _logged_batch(gpu, 1024, 0x4D445A34)
    pred = model.predict(sched.features(test_db), test_db.count)
    heur = test_b.src_len.astype(np.float64)
    rank = lambda a: np.argsort(np.argsort(a))
    corr_model = np.corrcoef(rank(pred), rank(cyc))[0, 1]
    corr_heur = np.corrcoef(rank(heur), rank(cyc))[0, 1]
    assert corr_model > corr_heur
    slots = deflate.slots()
    ms_model = sched.makespan(sched.lpt_order(pred), cyc, slots)
    ms_heur = sched.makespan(sched.lpt_order(heur), cyc, slots)
    ms_plain = sched.makespan(np.arange(test_db.count), cyc, slots)
    assert ms_model <= ms_heur and ms_model <= ms_plain
    # decoded bytes do not depend on the order
    out_before = test_db.outputs()
    test_db.set_order(sched.lpt_order(pred))
    test_db.dst.fill(0)
    deflate.inflate_sched(test_db)
    assert np.array_equal(test_db.outputs(), out_before)
