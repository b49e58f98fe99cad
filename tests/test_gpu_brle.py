"""GPU parity: byte RLE decode through the C ABI vs the CPU oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _run(mdz, c, dst_bytes=None, data=None, ptr=None):
    from compression_amd import brle
    b = brle.DeviceBrle(c.data if data is None else data, c.ptr if ptr is None else ptr,
                        c.n_chunks, c.chunk)
    n = c.raw_bytes if dst_bytes is None else dst_bytes
    dst = mdz.DeviceBuffer(max(n, 1))
    dst.fill(0xEE)
    st = mdz.DeviceBuffer(4 * c.n_chunks)
    brle.decompress(b, dst, n, st)
    return dst.download(n), st.download(4 * c.n_chunks, np.int32)


@pytest.mark.parametrize("chunk", [4096, 65536, 131072, 262144])
def test_runs_geom12(gpu, chunk):
    x = O.gen_brle_u8(3_000_017, 0x4D445A03)
    c = O.brle_encode(x, chunk, threads=8)
    got, st = _run(gpu, c)
    assert (st == 0).all() and np.array_equal(got, x)


def test_full_256MiB(gpu):
    x = O.gen_brle_u8(1 << 28, 0x4D445A03)
    c = O.brle_encode(x, 131072, threads=16)
    got, st = _run(gpu, c)
    assert c.n_chunks == 2048 and (st == 0).all() and np.array_equal(got, x)


def test_edge_patterns(gpu):
    rng = np.random.default_rng(4)
    cases = [np.full(300_000, 7, np.uint8),                                  # runs of 130
             (np.arange(300_000) * 7 + 3).astype(np.uint8),                   # literal groups of 128
             rng.integers(0, 256, 300_000).astype(np.uint8),                  # incompressible
             rng.integers(0, 2, 300_000).astype(np.uint8),                    # short runs mixed
             np.repeat(rng.integers(0, 256, 3000).astype(np.uint8), rng.integers(1, 200, 3000))]
    for x in cases:
        c = O.brle_encode(x, 65536, threads=8)
        got, st = _run(gpu, c)
        assert (st == 0).all() and np.array_equal(got, x)


def test_clipped_output(gpu):
    x = O.gen_brle_u8(500_000, 8)
    c = O.brle_encode(x, 65536, threads=8)
    got, st = _run(gpu, c, dst_bytes=123_457)
    assert np.array_equal(got, x[:123_457])
    _, _, ref_st = O.brle_decode(c, out_bytes=123_457, threads=4)
    assert np.array_equal(st, ref_st)


def test_corrupt_status_matches_oracle(gpu):
    x = O.gen_brle_u8(400_000, 9)
    c = O.brle_encode(x, 65536, threads=8)
    data = c.data.copy()
    data[int(c.ptr[1]) + 5] = 0x7F          # chunk 1: a run of 130 somewhere -> too much output
    ptr = c.ptr.copy()
    ptr[3] -= 1                              # chunk 2 loses its last byte, chunk 3 gains one
    got, st = _run(gpu, c, data=data, ptr=ptr)
    ref = O.BrleContainer(data, ptr, c.n_chunks, c.chunk, c.raw_bytes)
    _, _, ref_st = O.brle_decode(ref, threads=4)
    assert np.array_equal(st, ref_st)
    assert st[0] == 0 and (st[1:4] != 0).any()


def test_host_wrapper(gpu):
    from compression_amd import brle
    x = O.gen_brle_u8(700_001, 10)
    c = O.brle_encode(x, 262144, threads=8)
    assert brle.decompress_gpu(c.to_bytes()) == x.tobytes()


@pytest.mark.parametrize("chunk,shift", [(4099, 3), (65536, 7), (131072, 1), (1000, 15)])
def test_unaligned_destination(gpu, chunk, shift):
    # chunk outputs start off 16-byte boundaries: blocks are cut at the
    # destination's own alignment, with byte stores only at chunk ends
    from compression_amd import brle
    x = O.gen_brle_u8(1_000_003, 11 + shift)
    c = O.brle_encode(x, chunk, threads=8)
    b = brle.DeviceBrle(c.data, c.ptr, c.n_chunks, c.chunk)
    dst = gpu.DeviceBuffer(c.raw_bytes + 64)
    dst.fill(0xEE)
    st = gpu.DeviceBuffer(4 * c.n_chunks)
    brle.decompress(b, dst.ptr + shift, c.raw_bytes, st)
    got = dst.download(c.raw_bytes + 64)
    assert (st.download(4 * c.n_chunks, np.int32) == 0).all()
    assert np.array_equal(got[shift:shift + c.raw_bytes], x)
    assert (got[:shift] == 0xEE).all() and (got[shift + c.raw_bytes:] == 0xEE).all()


def test_truncated_last_token(gpu):
    # each chunk loses its last 1..3 bytes: the final token overruns the
    # input; status and every byte before the stop point match the oracle
    x = O.gen_brle_u8(600_000, 12)
    c = O.brle_encode(x, 65536, threads=8)
    keep, ptr, pos = [], [0], 0
    rng = np.random.default_rng(5)
    for i in range(c.n_chunks):
        a, e = int(c.ptr[i]), int(c.ptr[i + 1])
        e -= int(rng.integers(0, 4))
        keep.append(c.data[a:e])
        pos += e - a
        ptr.append(pos)
    data = np.concatenate(keep)
    ptr = np.asarray(ptr, np.uint64)
    ref = O.BrleContainer(data, ptr, c.n_chunks, c.chunk, c.raw_bytes)
    rc, ref_out, ref_st = O.brle_decode(ref, threads=4)
    got, st = _run(gpu, c, data=data, ptr=ptr)
    assert np.array_equal(st, ref_st) and (st != 0).any()
    for i in range(c.n_chunks):
        # the bytes the oracle writes before it stops, and nothing after them
        a, e = int(ptr[i]), int(ptr[i + 1])
        _, o1 = O.brle_chunk_decode(data[a:e].tobytes(), c.chunk)
        lo = i * c.chunk
        assert got[lo:lo + len(o1)].tobytes() == o1, i
        hi = min(lo + c.chunk, c.raw_bytes)
        assert (got[lo + len(o1):hi] == 0xEE).all(), i
