"""GPU parity: LZSS decode through the C ABI vs the CPU oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests.lzss_util import assemble, build_container

pytestmark = pytest.mark.gpu


def _run(mdz, c, dst_bytes=None, payload=None, col_len=None, col_map=None, blk_off=None):
    from compression_amd import lzss
    from compression_amd.lanes import DeviceLanes
    lanes = DeviceLanes(c.payload if payload is None else payload,
                        c.blk_off if blk_off is None else blk_off,
                        c.col_len if col_len is None else col_len, c.n_chunks, c.chunk, 1, 4,
                        col_map=c.col_map if col_map is None else col_map)
    n = c.raw_bytes if dst_bytes is None else dst_bytes
    dst = mdz.DeviceBuffer(max(n, 1))
    dst.fill(0xA5)
    st = mdz.DeviceBuffer(4 * max(c.n_chunks, 1))
    lzss.decompress(lanes, dst, n, st)
    return dst.download(n), st.download(4 * c.n_chunks, np.int32)


@pytest.mark.parametrize("chunk", [4096, 32768, 65536, 131072])
def test_markov_chunks(gpu, chunk):
    x = O.gen_lzss_markov(24 * chunk, 0x4D445A03)
    c = O.lzss_encode(x, chunk, threads=8)
    got, st = _run(gpu, c)
    assert (st == 0).all() and np.array_equal(got, x)


def test_full_config3(gpu):
    x = O.gen_lzss_markov(1 << 30, 0x4D445A03)           # 1 GiB, 16384 chunks of 64 KiB
    c = O.lzss_encode(x, 65536, threads=16)
    got, st = _run(gpu, c)
    assert c.n_chunks == 16384 and (st == 0).all()
    assert np.array_equal(got, x)


def test_edge_patterns(gpu):
    rng = np.random.default_rng(6)
    n = 16 * 65536
    cases = [np.zeros(n, np.uint8),                                        # long matches
             rng.integers(0, 256, n).astype(np.uint8),                     # literals only
             np.tile(np.arange(97, dtype=np.uint8), n // 97 + 1)[:n],      # periodic
             rng.integers(0, 3, n).astype(np.uint8),
             np.repeat(rng.integers(0, 256, n // 50 + 1).astype(np.uint8), 50)[:n]]
    for x in cases:
        for chunk in (65536, 131072):
            c = O.lzss_encode(x, chunk, threads=8)
            got, st = _run(gpu, c)
            assert (st == 0).all() and np.array_equal(got, x)


def test_overlapping_matches_from_crafted_streams(gpu):
    """Streams the reference encoder never emits (dist < len) decode byte-serially."""
    chunk = 4096
    lane_out = chunk // 32
    rng = np.random.default_rng(3)
    streams, lanes = [], []
    for t in range(32):
        out, s = bytearray(), bytearray()
        while len(out) < lane_out:
            f, body = 0, bytearray()
            for b in range(8):
                if len(out) >= lane_out:
                    break
                if len(out) == 0 or lane_out - len(out) < 3 or rng.random() < 0.4:
                    v = int(rng.integers(0, 256))
                    f |= 1 << b
                    body.append(v)
                    out.append(v)
                else:
                    d = int(rng.integers(1, min(len(out), 8) + 1))
                    ln = int(rng.integers(3, min(34, lane_out - len(out)) + 1))
                    v = (d << 5) | (ln - 3)
                    body += bytes([v & 0xff, v >> 8])
                    for _ in range(ln):
                        out.append(out[-d])
            s.append(f)
            s += body
        streams.append(bytes(s))
        lanes.append(bytes(out[:lane_out]))
    pay, cl, bo, cm = build_container([streams])
    c = O.LzssContainer(pay, cl, bo, cm, 1, chunk)
    rc, ref, ref_st = O.lzss_decode(c, threads=1)
    got, st = _run(gpu, c)
    assert rc == 0 and (st == 0).all()
    assert np.array_equal(ref, np.frombuffer(assemble([lanes], chunk), np.uint8))
    assert np.array_equal(got, ref)


def test_clipped_output(gpu):
    x = O.gen_lzss_markov(10 * 65536, 12)
    c = O.lzss_encode(x, 65536, threads=8)
    got, st = _run(gpu, c, dst_bytes=200_003)
    assert np.array_equal(got, x[:200_003]) and (st == 0).all()


def test_corrupt_status_matches_oracle(gpu):
    x = O.gen_lzss_markov(12 * 4096, 9)
    c = O.lzss_encode(x, 4096, threads=4)
    pay, cl, cm, bo = c.payload.copy(), c.col_len.copy(), c.col_map.copy(), c.blk_off.copy()
    cm[32 * 1 + 3] = cm[32 * 1 + 4]                        # chunk 1: col_map not a permutation
    cl[32 * 2 + 5], cl[32 * 2 + 6] = cl[32 * 2 + 6], cl[32 * 2 + 5] + 1   # chunk 2: unsorted
    cl[32 * 3 + 31] = 0                                    # chunk 3: an empty lane stream
    rng = np.random.default_rng(9)
    for ch in range(4, 12):                                # chunks 4..11: corrupt payload bytes
        a, b = int(bo[ch]), int(bo[ch + 1])
        for _ in range(3):
            pay[int(rng.integers(a, b))] ^= int(rng.integers(1, 256))
    bo[11] += 4                                            # chunk 10/11 extents shift
    bad = O.LzssContainer(pay, cl, bo, cm, c.n_chunks, c.chunk)
    _, _, ref_st = O.lzss_decode(bad, threads=4)
    _, st = _run(gpu, c, payload=pay, col_len=cl, col_map=cm, blk_off=bo)
    assert np.array_equal(st, ref_st)
    assert st[0] == 0 and st[1] == -11 and st[2] == -11 and (st[4:] != 0).any()


def test_host_wrapper_and_meta(gpu):
    from compression_amd import lzss
    x = O.gen_lzss_markov(5 * 131072, 10)
    c = O.lzss_encode(x, 131072, threads=8)
    blobs = lzss.split_meta(c.meta_bytes())
    assert len(blobs) == 4
    assert lzss.decompress_gpu(c.payload.tobytes(), blobs) == x.tobytes()
    with pytest.raises(Exception):
        lzss.decompress_gpu(c.payload.tobytes(), blobs[:3])
