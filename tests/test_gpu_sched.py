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
                         np.eye(3)[rng.integers(0, 3, n)], rng.uniform(5, 18, n),
                         rng.uniform(0, 0.6, n), rng.uniform(3, 16, n), rng.uniform(0, 5, n),
                         rng.uniform(0.3, 1.5, n)]).astype(np.float32)
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


def test_features_match_oracle(gpu):
    """The GPU header probe equals the oracle's restatement on every stream
    of an uneven batch (dynamic, fixed and stored), to float32 rounding."""
    from compression_amd import deflate, sched
    b = O.gen_deflate_batch(256, 0x4D445A25, 4096, 1 << 20, threads=16)
    db = deflate.DeviceBatch(b.src, b.src_off, b.src_len, b.raw_len)
    got = sched.features(db).download(b.count * M.F * 4, np.float32).reshape(b.count, M.F)
    ref = np.stack([M.features(b.stream(i), int(b.raw_len[i])) for i in range(b.count)])
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-4)
    # first block type by mode: stored, fixed (Z_FIXED), and dynamic or fixed (level 6)
    assert (got[b.mode == 0, 3] == 1).all() and (got[b.mode == 1, 4] == 1).all()
    assert (got[b.mode == 6, 4] + got[b.mode == 6, 5] == 1).all()
    assert (got[b.mode == 6, 7] > 0).all() and (got[b.mode == 0, 6] == 0).all()


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
    # the model ranks streams by their measured decode cycles far better than
    # the compressed size does (measured: about 0.99 against 0.7)
    assert corr_model > 0.95 and corr_model > corr_heur + 0.15
    # list scheduling of the logged cycles on a queue 16 streams deep per
    # worker (64 workers): the learned order is within 1% of the order by the
    # true cycles and of the size order (longest-first by compressed size is
    # itself within about 0.5% of the true order on these batches, so no order
    # can gain more than that), and well ahead of index order
    ms = {k: sched.makespan(o, cyc, 64) for k, o in
          (("model", sched.lpt_order(pred)), ("size", sched.lpt_order(heur)),
           ("true", sched.lpt_order(cyc)), ("index", np.arange(test_db.count)))}
    assert ms["model"] <= 1.01 * ms["true"] and ms["model"] <= 1.01 * ms["size"]
    # (the margin over index order follows the spread of the logged cycles:
    # 1.19 to 1.3 across kernel versions)
    assert ms["index"] >= 1.15 * ms["model"]
    # decoded bytes do not depend on the order
    out_before = test_db.outputs()
    test_db.set_order(sched.lpt_order(pred))
    test_db.dst.fill(0)
    deflate.inflate_sched(test_db)
    assert np.array_equal(test_db.outputs(), out_before)
