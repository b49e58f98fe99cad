"""GPU: the batched persistent decode (RLE v2 + byte RLE in one launch)
gives the per-codec decoders' bytes and statuses for any work order."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _parts(gpu, n_vals, n_bytes, ru=64, chunk=131072):
    from compression_amd import brle
    from compression_amd.lanes import DeviceLanes
    x = O.gen_rlev2_i64(n_vals, 0x4D445A02)
    c2 = O.rlev2_encode(x, chunk, ru, threads=8)
    y = O.gen_brle_u8(n_bytes, 0x4D445A03)
    cb = O.brle_encode(y, chunk, threads=8)
    lanes = DeviceLanes(c2.payload, c2.blk_off, c2.col_len, c2.n_chunks, chunk, 8, ru)
    db = brle.DeviceBrle(cb.data, cb.ptr, cb.n_chunks, chunk)
    return x, c2, lanes, y, cb, db


def _batch(gpu, lanes, c2, db, cb, quarters=False):
    from compression_amd import batch
    d1, s1 = gpu.DeviceBuffer(c2.raw_bytes), gpu.DeviceBuffer(4 * c2.n_chunks)
    d2, s2 = gpu.DeviceBuffer(cb.raw_bytes), gpu.DeviceBuffer(4 * cb.n_chunks)
    d1.fill(0xCC)
    d2.fill(0xCC)
    s1.fill(0x7F)
    s2.fill(0x7F)
    b = batch.Batch(lanes, d1, c2.raw_bytes, s1, 0, db, d2, cb.raw_bytes, s2, quarters=quarters)
    return b, (d1, s1, d2, s2)


@pytest.mark.parametrize("quarters", [True, False])
@pytest.mark.parametrize("order", ["index", "size", "random"])
def test_batch_matches_data(gpu, order, quarters):
    from compression_amd import batch
    x, c2, lanes, y, cb, db = _parts(gpu, 1 << 21, 1 << 23)
    b, (d1, s1, d2, s2) = _batch(gpu, lanes, c2, db, cb, quarters)
    if order == "size":
        b.set_order(b.size_order())
    elif order == "random":
        b.set_order(np.random.default_rng(1).permutation(b.all_items()))
    b.submit()
    assert (s1.download(dtype=np.int32) == 0).all() and (s2.download(dtype=np.int32) == 0).all()
    assert np.array_equal(d1.download(c2.raw_bytes), x.view(np.uint8))
    assert np.array_equal(d2.download(cb.raw_bytes), y)


def test_batch_small_grid_and_partial_items(gpu):
    from compression_amd import batch
    x, c2, lanes, y, cb, db = _parts(gpu, 1 << 20, 1 << 22)
    b, (d1, s1, d2, s2) = _batch(gpu, lanes, c2, db, cb)
    b.submit(blocks=3)                                    # 3 workgroups drain every item
    assert np.array_equal(d1.download(c2.raw_bytes), x.view(np.uint8))
    assert np.array_equal(d2.download(cb.raw_bytes), y)
    d2.fill(0xCC)
    b.set_order(batch.item(1, np.arange(0, cb.n_chunks, 2)))   # only even byte-RLE chunks
    b.submit()
    got = d2.download(cb.raw_bytes)
    for c in range(cb.n_chunks):
        sl = slice(c * cb.chunk, min((c + 1) * cb.chunk, cb.raw_bytes))
        if c % 2 == 0:
            assert np.array_equal(got[sl], y[sl])
        else:
            assert (got[sl] == 0xCC).all()


def test_batch_statuses_match_plain(gpu):
    from compression_amd import brle, rlev2
    x, c2, lanes, y, cb, db = _parts(gpu, 1 << 20, 1 << 22)
    data = cb.data.copy()
    data[int(cb.ptr[1]) + 3] = 0x7F                     # corrupt byte-RLE chunk 1
    db_bad = brle.DeviceBrle(data, cb.ptr, cb.n_chunks, cb.chunk)
    b, (d1, s1, d2, s2) = _batch(gpu, lanes, c2, db_bad, cb)
    b.submit()
    ps = gpu.DeviceBuffer(4 * cb.n_chunks)
    pd = gpu.DeviceBuffer(cb.raw_bytes)
    brle.decompress(db_bad, pd, cb.raw_bytes, ps)
    assert np.array_equal(s2.download(dtype=np.int32), ps.download(dtype=np.int32))
    assert (s1.download(dtype=np.int32) == 0).all()


def test_full_config2_batched(gpu):
    """Config 2 at full size through the one persistent launch, size order:
    the 1 GiB int64 column (8,192 chunks, READ_UNIT 512) and the 256 MiB
    byte-RLE stream (2,048 chunks), both bit-exact with the data and with the
    oracle's decode, every status 0."""
    from compression_amd import batch, brle
    from compression_amd.lanes import DeviceLanes
    x = O.gen_rlev2_i64(1 << 27, 0x4D445A02)
    c2 = O.rlev2_encode(x, 131072, 512, threads=16)
    y = O.gen_brle_u8(1 << 28, 0x4D445A03)
    cb = O.brle_encode(y, 131072, threads=16)
    lanes = DeviceLanes(c2.payload, c2.blk_off, c2.col_len, c2.n_chunks, 131072, 8, 512)
    db = brle.DeviceBrle(cb.data, cb.ptr, cb.n_chunks, 131072)
    assert c2.n_chunks == 8192 and cb.n_chunks == 2048
    b, (d1, s1, d2, s2) = _batch(gpu, lanes, c2, db, cb)
    b.set_order(b.size_order())
    b.submit()
    assert (s1.download(dtype=np.int32) == 0).all() and (s2.download(dtype=np.int32) == 0).all()
    got1 = d1.download(c2.raw_bytes)
    assert np.array_equal(got1, x.view(np.uint8))
    got2 = d2.download(cb.raw_bytes)
    assert np.array_equal(got2, y)
    rc, ref2, _ = O.brle_decode(cb)
    assert rc == 0 and np.array_equal(got2, ref2)
