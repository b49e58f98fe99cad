"""GPU: captured pipelines (mdz_graph_*).  A graph holds only kernel nodes,
and every replay -- with eager work and synchronisation between replays --
gives the eager launch's bytes and statuses."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("quarters", [False, True])
def test_batch_graph_replays(gpu, quarters):
    from compression_amd import batch, brle, graph
    from compression_amd.lanes import DeviceLanes
    x = O.gen_rlev2_i64(1 << 20, 0x4D445A02)
    c2 = O.rlev2_encode(x, 131072, 64, threads=8)
    y = O.gen_brle_u8(1 << 22, 0x4D445A03)
    cb = O.brle_encode(y, 131072, threads=8)
    lanes = DeviceLanes(c2.payload, c2.blk_off, c2.col_len, c2.n_chunks, 131072, 8, 64)
    dbr = brle.DeviceBrle(cb.data, cb.ptr, cb.n_chunks, 131072)
    d1, s1 = gpu.DeviceBuffer(c2.raw_bytes), gpu.DeviceBuffer(4 * c2.n_chunks)
    d2, s2 = gpu.DeviceBuffer(cb.raw_bytes), gpu.DeviceBuffer(4 * cb.n_chunks)
    b = batch.Batch(lanes, d1, c2.raw_bytes, s1, 0, dbr, d2, cb.raw_bytes, s2, quarters=quarters)
    b.set_order(b.size_order())
    b.submit()                                              # eager reference
    eager = (d1.download(), s1.download(dtype=np.int32), d2.download(), s2.download(dtype=np.int32))
    assert np.array_equal(eager[0], x.view(np.uint8)) and np.array_equal(eager[2], y)
    s = gpu.Stream()
    with graph.capture(s) as g:
        b.submit(s.ptr)
    n, k = g.nodes()
    assert n == k and k >= 2                                # kernels only (no memset nodes)
    for rep in range(3):
        for buf in (d1, d2, s1, s2):                        # eager work between replays
            buf.fill(0xA0 + rep)
        gpu.lib().mdz_device_sync()
        g.launch(s)
        s.sync()
        got = (d1.download(), s1.download(dtype=np.int32), d2.download(), s2.download(dtype=np.int32))
        assert all(np.array_equal(a, e) for a, e in zip(got, eager)), rep
        b.submit()                                          # an eager launch in between, too
        gpu.lib().mdz_device_sync()
    g.destroy()


def test_inflate_graph_replays(gpu):
    from compression_amd import deflate, graph
    bt = O.gen_deflate_batch(64, 0x4D445A04, 4096, 1 << 19)
    dbt = deflate.DeviceBatch(bt.src, bt.src_off, bt.src_len, bt.raw_len)
    deflate.inflate_sched(dbt)
    eager_st = dbt.statuses()
    assert (eager_st["status"] == 0).all()

    def written():          # each stream's written bytes (slot padding is never written)
        out = dbt.outputs()
        return [out[int(o):int(o) + int(n)] for o, n in zip(dbt.dst_off, eager_st["bytes_written"])]

    eager_out = written()
    s = gpu.Stream()
    with graph.capture(s) as g:
        deflate.inflate_sched(dbt, stream=s.ptr)
    n, k = g.nodes()
    assert n == k and k == 2                                # cursor reset + inflate kernel
    for rep in range(3):
        dbt.dst.fill(rep)
        dbt.status.fill(0xFF)
        gpu.lib().mdz_device_sync()
        g.launch(s)
        s.sync()
        assert all(np.array_equal(a, e) for a, e in zip(written(), eager_out)), rep
        assert np.array_equal(dbt.statuses(), eager_st), rep


def test_capture_needs_a_created_stream(gpu):
    from compression_amd import graph
    with pytest.raises(gpu.MdzError):
        with graph.capture(None):
            pass
