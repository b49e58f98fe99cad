"""GPU parity: lane-transposed byte RLE (brle_trans) through the C ABI vs the
CPU oracle -- bytes on good chunks, statuses on every chunk."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _run(mdz, c, dst_bytes=None, payload=None, col_len=None, col_map=None, blk_off=None):
    from compression_amd import brle
    from compression_amd.lanes import DeviceLanes
    lanes = DeviceLanes(c.payload if payload is None else payload,
                        c.blk_off if blk_off is None else blk_off,
                        c.col_len if col_len is None else col_len, c.n_chunks, c.chunk, 1, 4,
                        col_map=c.col_map if col_map is None else col_map)
    n = c.raw_bytes if dst_bytes is None else dst_bytes
    dst = mdz.DeviceBuffer(max(n, 1))
    dst.fill(0xA5)
    st = mdz.DeviceBuffer(4 * max(c.n_chunks, 1))
    brle.decompress_trans(lanes, dst, n, st)
    return dst.download(n), st.download(4 * c.n_chunks, np.int32)


@pytest.mark.parametrize("chunk", [128, 4096, 65536, 131072])
def test_generator_chunks(gpu, chunk):
    x = O.gen_brle_u8(24 * chunk, 0x4D445A02)
    c = O.brle_trans_encode(x, chunk, threads=8)
    got, st = _run(gpu, c)
    assert (st == 0).all() and np.array_equal(got, x)


def test_full_256MiB(gpu):
    """The 256 MiB byte-RLE column of the RLE v2 + byte RLE workload, in the
    lane-transposed format at 128 KiB chunks; every byte compared."""
    x = O.gen_brle_u8(256 << 20, 0x4D445A02)
    c = O.brle_trans_encode(x, 131072, threads=16)
    rc, ref, st_ref = O.brle_trans_decode(c)
    got, st = _run(gpu, c)
    assert np.array_equal(st, st_ref)
    assert np.array_equal(got, ref)
    if rc == 0:                    # no run of 256+ bytes: the round trip holds
        assert np.array_equal(got, x)


def test_edge_patterns(gpu):
    rng = np.random.default_rng(9)
    n = 8 * 65536
    cases = [rng.integers(0, 256, n).astype(np.uint8),                    # literals only
             np.repeat(rng.integers(0, 256, n // 255 + 1).astype(np.uint8), 255)[:n],   # longest runs
             np.repeat(rng.integers(0, 256, n // 3 + 1).astype(np.uint8), 3)[:n],       # shortest runs
             np.repeat(rng.integers(0, 256, n // 2 + 1).astype(np.uint8), 2)[:n],       # pairs: literals
             np.zeros(n, np.uint8)]                                        # counter wraps (lossy)
    for x in cases:
        c = O.brle_trans_encode(x, 65536, threads=8)
        rc, ref, st_ref = O.brle_trans_decode(c)
        got, st = _run(gpu, c)
        assert np.array_equal(st, st_ref)
        good = np.repeat(st_ref == 0, 65536)
        assert np.array_equal(got[good], ref[good])


def test_damaged_containers_match_oracle_statuses(gpu):
    rng = np.random.default_rng(4)
    chunk = 4096
    x = O.gen_brle_u8(64 * chunk, 0x44)
    c = O.brle_trans_encode(x, chunk, threads=8)
    pay = c.payload.copy()
    for _ in range(80):
        pay[rng.integers(0, pay.size)] = rng.integers(0, 256)
    bad = O.LzssContainer(pay, c.col_len, c.blk_off, c.col_map, c.n_chunks, chunk)
    rc, ref, st_ref = O.brle_trans_decode(bad)
    got, st = _run(gpu, c, payload=pay)
    assert np.array_equal(st, st_ref)
    assert (st_ref != 0).any() and (st_ref == 0).any()
    for i in np.nonzero(st_ref == 0)[0]:
        assert np.array_equal(got[i * chunk:(i + 1) * chunk], ref[i * chunk:(i + 1) * chunk])
    # metadata: col_map not a permutation, col_len out of order -> BAD_HEADER
    cm = c.col_map.copy()
    cm[32 * 3 + 1] = cm[32 * 3]
    cl = c.col_len.copy()
    cl[32 * 5] = cl[32 * 5 + 1] - 1                 # sorted position 0 shorter than 1
    for kw in ({"col_map": cm}, {"col_len": cl}):
        b2 = O.LzssContainer(c.payload, kw.get("col_len", c.col_len), c.blk_off, kw.get("col_map", c.col_map),
                             c.n_chunks, chunk)
        _, _, st_ref = O.brle_trans_decode(b2)
        _, st = _run(gpu, c, **kw)
        assert np.array_equal(st, st_ref) and (st == -11).sum() == 1
