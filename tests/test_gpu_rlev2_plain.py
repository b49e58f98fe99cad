"""GPU parity: plain RLE v2 (one run stream per chunk of int64 values,
include/rlev2/decoder.h:190-262) through the C ABI vs the CPU oracle --
bytes on good chunks, statuses on every chunk."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests.test_oracle_rlev2_plain import ORC_EXAMPLES

pytestmark = pytest.mark.gpu


def _container(streams, chunk, raw_bytes=None):
    """Builds [u32 n+1][u64 ptr[n+1]][u64 raw][streams] from chunk streams."""
    ptr = np.zeros(len(streams) + 1, np.uint64)
    ptr[1:] = np.cumsum([len(s) for s in streams])
    raw = len(streams) * chunk if raw_bytes is None else raw_bytes
    head = np.array([len(streams) + 1], np.uint32).tobytes() + ptr.tobytes() + np.uint64(raw).tobytes()
    return np.frombuffer(head + b"".join(streams), np.uint8).copy()


def _run(mdz, blob, chunk, out_bytes=None, data=None):
    from compression_amd import rlev2
    from compression_amd.host import RleV2PlainContainer
    c = RleV2PlainContainer(blob, chunk)
    d = c.data if data is None else data
    n = c.raw_bytes if out_bytes is None else out_bytes
    dd = mdz.DeviceBuffer.from_array(d if d.size else np.zeros(1, np.uint8), pad=16)
    dp = mdz.DeviceBuffer.from_array(c.ptr)
    dst = mdz.DeviceBuffer(max(n, 8))
    dst.fill(0xA5)
    st = mdz.DeviceBuffer(4 * max(c.n_chunks, 1))
    rlev2.decompress_plain(dd, d.size, dp, c.n_chunks, chunk, dst, n, st)
    return dst.download(n), st.download(4 * c.n_chunks, np.int32)


def _values_written(blob, chunk, c):
    """Bytes of chunk c's output that its stream writes (the oracle's count)."""
    from compression_amd.host import RleV2PlainContainer
    k = RleV2PlainContainer(blob, chunk)
    s = k.data[int(k.ptr[c]):int(k.ptr[c + 1])].tobytes()
    return 8 * len(O.rlev2_plain_chunk_decode(s, chunk // 8)[1])


def test_orc_spec_examples(gpu):
    streams = [bytes.fromhex(h) for _, _, h in ORC_EXAMPLES]
    blob = _container(streams, 4096)
    got, st = _run(gpu, blob, 4096)
    assert (st == 0).all()
    for i, (_, vals, _) in enumerate(ORC_EXAMPLES):
        v = got[i * 4096:i * 4096 + 8 * len(vals)].view(np.int64)
        assert v.tolist() == vals


def _col(n, seed):
    return O.gen_rlev2_i64(n, seed).astype(np.int64) & np.int64(0x7FFFFFFFFFFFFFFF)


@pytest.mark.parametrize("chunk", [4096, 8192, 32768, 40960])
def test_generated_columns(gpu, chunk):
    x = _col(3_000_000, 0x4D445A02)
    x[1000:200000] = np.sort(x[1000:200000])[::-1]          # descending stretch
    x[300000:400000] = np.random.default_rng(1).integers(0, 130, 100000)
    c = O.rlev2_plain_encode(x, chunk, threads=8)
    rc, ref, st_ref = O.rlev2_plain_decode(c.blob, chunk)
    assert rc == 0
    got, st = _run(gpu, c.blob, chunk)
    assert (st == 0).all()
    assert np.array_equal(got, ref) and np.array_equal(got, x.view(np.uint8))


def test_full_256MiB(gpu):
    """A 256 MiB column of the RLE v2 workload's generator in the plain
    format at the reference's 32 KiB chunks; every byte compared."""
    x = _col(32 << 20, 0x4D445A02)
    c = O.rlev2_plain_encode(x, 32768, threads=16)
    got, st = _run(gpu, c.blob, 32768)
    assert (st == 0).all() and np.array_equal(got, x.view(np.uint8))


def test_host_form_and_clipped_output(gpu):
    from compression_amd import rlev2
    x = _col(100_000, 3)
    raw = np.concatenate([x.view(np.uint8), np.array([9, 9, 9], np.uint8)])   # 3 unencoded tail bytes
    c = O.rlev2_plain_encode(raw, 32768)
    out = rlev2.decompress_plain_gpu(c.blob.tobytes())
    assert len(out) == raw.size and out[:x.nbytes] == x.tobytes()
    for n in (x.nbytes - 5, 32768 * 2 + 3, 1):                   # outputs cut mid-value
        rc, ref, st_ref = O.rlev2_plain_decode(c.blob, 32768, out_bytes=n)
        got, st = _run(gpu, c.blob, 32768, out_bytes=n)
        assert np.array_equal(st, st_ref) and np.array_equal(got, ref[:n])


def test_edge_streams(gpu):
    ex = bytes.fromhex(ORC_EXAMPLES[2][2])
    streams = [b"", bytes.fromhex("0a2710"), ex * 409, ex * 410, bytes([0x40, 0x00, 0x07]),
               O.rlev2_plain_chunk_encode(np.full(4096, 7, np.int64)),
               O.rlev2_plain_chunk_encode(np.arange(4096, dtype=np.int64) * 3 + (1 << 62))]
    blob = _container(streams, 32768)
    rc, ref, st_ref = O.rlev2_plain_decode(blob, 32768)
    got, st = _run(gpu, blob, 32768)
    assert np.array_equal(st, st_ref)
    assert st_ref.tolist() == [0, 0, 0, 1, 0, 0, 0]     # 410 * 10 values overrun a 4096-value chunk
    for i in range(len(streams)):                        # bytes past a chunk's values are not written
        if st_ref[i] == 0:
            n = _values_written(blob, 32768, i)
            assert np.array_equal(got[i * 32768:i * 32768 + n], ref[i * 32768:i * 32768 + n])


def test_damaged_containers_match_oracle_statuses(gpu):
    rng = np.random.default_rng(4)
    chunk = 4096
    x = _col(64 * chunk // 8, 0x44)
    c = O.rlev2_plain_encode(x, chunk, threads=8)
    seen = set()
    for trial in range(4):
        blob = c.blob.copy()
        for _ in range(400):
            blob[c.data_off + rng.integers(0, c.data.size)] = rng.integers(0, 256)
        rc, ref, st_ref = O.rlev2_plain_decode(blob, chunk)
        got, st = _run(gpu, blob, chunk)
        assert np.array_equal(st, st_ref)
        seen |= set(st_ref.tolist())
        for i in np.nonzero(st_ref == 0)[0]:
            n = _values_written(blob, chunk, i)
            assert np.array_equal(got[i * chunk:i * chunk + n], ref[i * chunk:i * chunk + n])
    assert {-10, 1, 2} <= seen                  # malformed symbol, output and input overruns
    # offset table out of order / past the data: BAD_HEADER for that chunk only
    blob = c.blob.copy()
    ptr = blob[4:4 + 8 * (c.n_chunks + 1)].view(np.uint64)
    ptr[5] = ptr[7]
    rc, ref, st_ref = O.rlev2_plain_decode(blob, chunk)
    got, st = _run(gpu, blob, chunk)
    assert np.array_equal(st, st_ref) and st[5] == -11
