"""GPU parity: RLE v2 decode through the C ABI vs the CPU oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _run(mdz, c, dst_bytes=None, payload=None, col_len=None, max_chunk=0, quarters=False,
         window=0):
    from compression_amd import rlev2
    from compression_amd.lanes import DeviceLanes
    pay = c.payload if payload is None else payload
    col = c.col_len if col_len is None else col_len
    lanes = DeviceLanes(pay, c.blk_off, col, c.n_chunks, c.chunk, c.elem_bytes, c.read_unit)
    n = c.raw_bytes if dst_bytes is None else dst_bytes
    dst = mdz.DeviceBuffer(max(n, 1))
    dst.fill(0xCD)
    st = mdz.DeviceBuffer(4 * c.n_chunks)
    rlev2.decompress(lanes, dst, n, st, max_chunk_bytes=max_chunk, quarters=quarters,
                     window=window)
    return dst.download(n), st.download(4 * c.n_chunks, np.int32)


@pytest.mark.parametrize("quarters", [True, False])
@pytest.mark.parametrize("dtype,ru", [(np.uint64, 1), (np.uint64, 8), (np.uint64, 64),
                                      (np.uint64, 512), (np.uint32, 64), (np.uint16, 32),
                                      (np.uint8, 16)])
def test_config2_shapes(gpu, dtype, ru, quarters):
    x = O.gen_rlev2_i64(1 << 18, 0x4D445A02).astype(dtype)
    c = O.rlev2_encode(x, 131072, ru, threads=8)
    got, st = _run(gpu, c, quarters=quarters)
    assert (st == 0).all()
    assert np.array_equal(got, x.view(np.uint8))


def test_full_config2_column(gpu):
    x = O.gen_rlev2_i64(1 << 27, 0x4D445A02)          # 1 GiB int64, 8192 chunks
    c = O.rlev2_encode(x, 131072, 512, threads=16)
    mc = int(np.diff(c.blk_off[:-1].astype(np.int64)).max(initial=0))
    mc = max(mc, c.payload_bytes - int(c.blk_off[-2]))
    got, st = _run(gpu, c, max_chunk=mc)
    assert c.n_chunks == 8192 and (st == 0).all()
    assert np.array_equal(got, x.view(np.uint8))


def test_ragged_and_clipped(gpu):
    x = O.gen_rlev2_i64(100_003, 7)
    c = O.rlev2_encode(x, 65536, 64, threads=8)
    got, st = _run(gpu, c)
    assert (st == 0).all() and np.array_equal(got, x.view(np.uint8))
    got, st = _run(gpu, c, dst_bytes=77_777)
    assert np.array_equal(got, x.view(np.uint8)[:77_777])


def test_edge_streams(gpu):
    rng = np.random.default_rng(2)
    cases = [np.zeros(1 << 16, np.uint64),                                   # B19: runs of zeros
             np.full(1 << 16, (1 << 64) - 1, np.uint64),
             rng.integers(0, 1 << 63, 1 << 16, dtype=np.uint64) * np.uint64(2) + np.uint64(1),
             np.arange(1 << 16, dtype=np.uint64)[::-1].copy(),              # decreasing ramps
             (rng.integers(0, 1 << 12, 1 << 16).astype(np.uint64) |
              (rng.integers(0, 100, 1 << 16) == 0).astype(np.uint64) << np.uint64(55))]  # patches
    for x in cases:
        c = O.rlev2_encode(x, 32768, 128, threads=8)
        got, st = _run(gpu, c)
        assert (st == 0).all() and np.array_equal(got, x.view(np.uint8))


def test_corrupt_status_matches_oracle(gpu):
    x = O.gen_rlev2_i64(1 << 16, 13)
    c = O.rlev2_encode(x, 32768, 64, threads=8)
    col = c.col_len.copy()
    col[3] = max(1, col[3] // 3)                    # chunk 0: a lane stream cut short
    pay = c.payload.copy()
    lo = int(c.blk_off[1])
    pay[lo:lo + 8] = 0xFF                           # chunk 1: damaged run headers
    got, st = _run(gpu, c, payload=pay, col_len=col)
    ref = O.RleV2Container(pay, col, c.blk_off, c.n_chunks, c.chunk, c.elem_bytes, c.read_unit,
                           c.raw_bytes, c.payload_bytes)
    _, _, ref_st = O.rlev2_decode(ref, threads=4)
    assert st[0] != 0 and st[1] != 0 and (st[2:] == 0).all()
    assert np.array_equal(st, ref_st)


def test_host_wrapper(gpu):
    from compression_amd import rlev2
    x = O.gen_rlev2_i64(90_000, 21)
    c = O.rlev2_encode(x, 65536, 128, threads=8)
    out = rlev2.decompress_gpu(c.payload.tobytes(), c.n_chunks, c.blk_off, c.col_len, 65536, 128, 8)
    assert out == x.tobytes()


@pytest.mark.parametrize("window", [8, 16, 32, 64])
@pytest.mark.parametrize("quarters", [True, False])
def test_consumer_windows(gpu, window, quarters):
    # every consumer window gives the same bytes: enough runs per chunk for
    # several record epochs (short segments at a small READ_UNIT)
    x = O.gen_rlev2_i64(1 << 18, 0x4D445A02)
    c = O.rlev2_encode(x, 131072, 8, threads=8)
    got, st = _run(gpu, c, quarters=quarters, window=window)
    assert (st == 0).all() and np.array_equal(got, x.view(np.uint8))


def test_bad_window_is_rejected(gpu):
    x = O.gen_rlev2_i64(1 << 14, 3)
    c = O.rlev2_encode(x, 32768, 64, threads=2)
    with pytest.raises(gpu.MdzError):
        _run(gpu, c, window=24)


@pytest.mark.parametrize("window", [8, 32])
@pytest.mark.parametrize("quarters", [True, False])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_corruption_matches_oracle(gpu, seed, quarters, window):
    # one random byte overwritten in most chunks (header or data) and a few
    # lane lengths changed: every chunk's status matches the oracle, and
    # chunks the oracle decodes cleanly are bit-exact
    x = O.gen_rlev2_i64(1 << 18, 40 + seed)
    c = O.rlev2_encode(x, 32768, 64, threads=8)
    rng = np.random.default_rng(seed)
    pay = c.payload.copy()
    col = c.col_len.copy()
    for ch in range(c.n_chunks):
        lo = int(c.blk_off[ch])
        hi = int(c.blk_off[ch + 1]) if ch + 1 < c.n_chunks else c.payload_bytes
        if rng.random() < 0.8 and hi > lo:
            pay[lo + int(rng.integers(0, hi - lo))] = int(rng.integers(0, 256))
        if rng.random() < 0.2:
            t = ch * 32 + int(rng.integers(0, 32))
            col[t] = max(0, int(col[t]) - int(rng.integers(1, 40)))
    got, st = _run(gpu, c, payload=pay, col_len=col, quarters=quarters, window=window)
    ref = O.RleV2Container(pay, col, c.blk_off, c.n_chunks, c.chunk, c.elem_bytes, c.read_unit,
                           c.raw_bytes, c.payload_bytes)
    _, ref_out, ref_st = O.rlev2_decode(ref, threads=4)
    assert np.array_equal(st, ref_st)
    assert (st != 0).any()
    for ch in np.flatnonzero(ref_st == 0):
        lo, hi = ch * c.chunk, min((ch + 1) * c.chunk, c.raw_bytes)
        assert np.array_equal(got[lo:hi], ref_out[lo:hi]), ch
