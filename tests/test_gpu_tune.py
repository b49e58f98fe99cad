"""GPU: the autotuner picks a launch shape whose output is bit-exact, stores
it, and reuses it from its database without timing."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _pair(gpu):
    from compression_amd import brle
    from compression_amd.lanes import DeviceLanes
    x = O.gen_rlev2_i64(1 << 20, 0x4D445A02)                # 8 MiB, 64 chunks
    c2 = O.rlev2_encode(x, 131072, 64, threads=8)
    y = O.gen_brle_u8(1 << 22, 0x4D445A03)                  # 4 MiB, 32 chunks
    cb = O.brle_encode(y, 131072, threads=8)
    lanes = DeviceLanes(c2.payload, c2.blk_off, c2.col_len, c2.n_chunks, 131072, 8, 64)
    db = brle.DeviceBrle(cb.data, cb.ptr, cb.n_chunks, 131072)
    d1, s1 = gpu.DeviceBuffer(c2.raw_bytes), gpu.DeviceBuffer(4 * c2.n_chunks)
    d2, s2 = gpu.DeviceBuffer(cb.raw_bytes), gpu.DeviceBuffer(4 * cb.n_chunks)
    args = dict(rlev2=lanes, rlev2_dst=d1, rlev2_dst_bytes=c2.raw_bytes, rlev2_status=s1,
                brle=db, brle_dst=d2, brle_dst_bytes=cb.raw_bytes, brle_status=s2)
    return x, y, args


def test_tune_batch_bit_exact_and_reused(gpu, tmp_path):
    from compression_amd import tune
    x, y, args = _pair(gpu)
    tdb = tune.TuneDB(str(tmp_path / "tune.json"))
    b, res = tune.tune_batch(**args, db=tdb, reps=2, warmup=1)
    assert not res.from_db and res.ms > 0
    assert len(res.candidates) == 36
    assert all(t is not None for _, t in res.candidates)    # every shape decodes the same bytes
    for buf in (args["rlev2_dst"], args["brle_dst"]):
        buf.fill(0xEE)
    b.submit()
    assert (args["rlev2_status"].download(dtype=np.int32) == 0).all()
    assert (args["brle_status"].download(dtype=np.int32) == 0).all()
    assert np.array_equal(args["rlev2_dst"].download(args["rlev2_dst_bytes"]), x.view(np.uint8))
    assert np.array_equal(args["brle_dst"].download(args["brle_dst_bytes"]), y)
    # the second call takes the stored choice, with no timing
    b2, res2 = tune.tune_batch(**args, db=tune.TuneDB(tdb.path),
                               time_fn=lambda c: pytest.fail("timed again"))
    assert res2.from_db and res2.choice == res.choice


def test_tune_inflate_bit_exact(gpu, tmp_path):
    from compression_amd import deflate, sched, tune
    b = O.gen_deflate_batch(96, 0x4D445A04, 4096, 1 << 20)
    dbt = deflate.DeviceBatch(b.src, b.src_off, b.src_len, b.raw_len)
    orders = {"index": None, "size": sched.lpt_order(b.src_len.astype(np.float64))}
    blocks, res = tune.tune_inflate(dbt, orders, db=tune.TuneDB(str(tmp_path / "t.json")), reps=1)
    assert blocks >= 1 and res.choice["order"] in orders
    deflate.inflate_sched(dbt, blocks=blocks)
    _, ref, _, _, st = O.inflate_batch(b)
    out = dbt.outputs()
    assert (dbt.statuses()["status"] == 0).all() and (st == 0).all()
    for i in range(b.count):
        o, r = int(dbt.dst_off[i]), int(b.raw_len[:i].sum())
        assert np.array_equal(out[o:o + int(b.raw_len[i])], ref[r:r + int(b.raw_len[i])]), i
