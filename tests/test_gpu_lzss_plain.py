"""GPU parity: plain (per-chunk) LZSS decode through the C ABI vs the CPU
oracle (bit-exact bytes on good chunks, equal statuses on damaged ones)."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _run(mdz, data, lens, n_chunks, chunk, dst_bytes, data_off=0, dst_off=0, data_bytes=None):
    """Decode with the stream placed data_off bytes into its buffer and the
    output dst_off bytes into its buffer (exercises unaligned staging/stores)."""
    from compression_amd import lzss
    data = np.ascontiguousarray(data, np.uint8)
    d = mdz.DeviceBuffer(data_off + data.size + 1)
    if data.size:
        d.upload(np.concatenate([np.zeros(data_off, np.uint8), data]))
    ln = mdz.DeviceBuffer.from_array(np.ascontiguousarray(lens, np.uint64))
    out = mdz.DeviceBuffer(dst_off + max(dst_bytes, 1) + 16)
    out.fill(0xA5)
    st = mdz.DeviceBuffer(4 * max(n_chunks, 1))
    st.fill(0x7F)
    lzss.decompress_plain(d.ptr + data_off, data.size if data_bytes is None else data_bytes, ln, n_chunks,
                          chunk, out.ptr + dst_off, dst_bytes, st)
    full = out.download(dst_off + dst_bytes + 16)
    assert (full[:dst_off] == 0xA5).all() and (full[dst_off + dst_bytes:] == 0xA5).all(), "write outside dst"
    return full[dst_off:dst_off + dst_bytes], st.download(4 * n_chunks, np.int32)


def _oracle(c, out_bytes):
    rc, ref, st = O.lzss_plain_decode(c, out_bytes)
    return ref, st


@pytest.mark.parametrize("chunk", [4096, 65536, 131072])
def test_markov_chunks(gpu, chunk):
    x = O.gen_lzss_markov(24 * chunk + 12345, 0x4D445A03)
    c = O.lzss_plain_encode(x, chunk, threads=8)
    got, st = _run(gpu, c.data, c.lens, c.n_chunks, chunk, x.size)
    assert (st == 0).all() and np.array_equal(got, x)


def test_full_config3_plain(gpu):
    """1 GiB Markov+copy stream, 16384 chunks of 64 KiB, every byte compared."""
    x = O.gen_lzss_markov(1 << 30, 0x4D445A03)
    c = O.lzss_plain_encode(x, 65536, threads=16)
    got, st = _run(gpu, c.data, c.lens, c.n_chunks, 65536, x.size)
    assert c.n_chunks == 16384 and (st == 0).all()
    assert np.array_equal(got, x)


@pytest.mark.parametrize("chunk", [100, 1000, 65536, 131072, 300000])
def test_edge_patterns(gpu, chunk):
    """Literal-only streams (several staged pieces per chunk), long runs
    (self-overlapping matches), periodic data, odd chunk sizes."""
    rng = np.random.default_rng(chunk)
    n = 6 * chunk + 77
    cases = [np.zeros(n, np.uint8),
             rng.integers(0, 256, n).astype(np.uint8),
             np.tile(np.arange(97, dtype=np.uint8), n // 97 + 1)[:n],
             rng.integers(0, 3, n).astype(np.uint8),
             np.repeat(rng.integers(0, 256, n // 50 + 1).astype(np.uint8), 50)[:n]]
    for x in cases:
        c = O.lzss_plain_encode(x, chunk, threads=8)
        got, st = _run(gpu, c.data, c.lens, c.n_chunks, chunk, x.size)
        assert (st == 0).all() and np.array_equal(got, x)


@pytest.mark.parametrize("data_off,dst_off", [(1, 0), (3, 5), (2, 16), (0, 7)])
def test_unaligned_buffers(gpu, data_off, dst_off):
    x = O.gen_lzss_markov(9 * 65536 + 999, 0x77)
    c = O.lzss_plain_encode(x, 65536, threads=8)
    got, st = _run(gpu, c.data, c.lens, c.n_chunks, 65536, x.size, data_off, dst_off)
    assert (st == 0).all() and np.array_equal(got, x)


def test_crafted_overlaps_and_long_chains(gpu):
    """Streams the encoder never emits: distances shorter than the length
    (byte-serial overlap), distance-1 runs across windows, and copies of
    copies all the way back through 2047-byte distances."""
    rng = np.random.default_rng(11)
    chunk = 70000
    streams, outs = [], []
    for t in range(12):
        out, s = bytearray(), bytearray()
        while len(out) < chunk:
            f, body = 0, bytearray()
            for b in range(8):
                if len(out) >= chunk:
                    break
                if len(out) == 0 or chunk - len(out) < 3 or rng.random() < (0.02 if t % 3 else 0.4):
                    v = int(rng.integers(0, 256))
                    f |= 1 << b
                    body.append(v)
                    out.append(v)
                else:
                    hi = min(len(out), 2047)
                    d = int(rng.integers(1, min(hi, 8) + 1)) if t % 2 else int(rng.integers(max(1, hi - 40), hi + 1))
                    ln = int(rng.integers(3, min(34, chunk - len(out)) + 1))
                    v = (d << 5) | (ln - 3)
                    body += bytes([v & 0xff, v >> 8])
                    for _ in range(ln):
                        out.append(out[-d])
            s.append(f)
            s += body
        streams.append(bytes(s))
        outs.append(bytes(out))
    data = np.frombuffer(b"".join(streams), np.uint8)
    lens = np.cumsum([len(s) for s in streams]).astype(np.uint64)
    for s, o in zip(streams, outs):
        assert O.lzss_lane_decode(s, chunk) == (0, o)
    got, st = _run(gpu, data, lens, len(streams), chunk, chunk * len(streams))
    assert (st == 0).all() and got.tobytes() == b"".join(outs)


def test_damaged_streams_match_oracle_statuses(gpu):
    """Random byte flips, truncations and a stream that runs short: the GPU
    reports the oracle's status for every chunk and its bytes on good ones."""
    rng = np.random.default_rng(5)
    chunk = 32768
    x = O.gen_lzss_markov(64 * chunk, 0x99)
    c = O.lzss_plain_encode(x, chunk, threads=8)
    data = c.data.copy()
    lens = c.lens.copy()
    for _ in range(60):
        data[rng.integers(0, data.size)] = rng.integers(0, 256)
    bad = O.LzssPlainContainer(data, lens, c.n_chunks, chunk, c.raw_bytes)
    ref, st_ref = _oracle(bad, x.size)
    got, st = _run(gpu, data, lens, c.n_chunks, chunk, x.size)
    assert np.array_equal(st, st_ref), (st, st_ref)
    assert (st_ref != 0).any() and (st_ref == 0).any()
    for i in np.nonzero(st_ref == 0)[0]:
        assert np.array_equal(got[i * chunk:(i + 1) * chunk], ref[i * chunk:(i + 1) * chunk])
    # every chunk cut short by 1..40 bytes: lens keep the order, streams shrink
    L = np.diff(np.concatenate([[0], c.lens.astype(np.int64)]))
    cut = rng.integers(1, 41, c.n_chunks)
    parts = [c.data[int(c.lens[i] - L[i]):int(c.lens[i] - cut[i])] for i in range(c.n_chunks)]
    d2 = np.concatenate(parts)
    l2 = np.cumsum([p.size for p in parts]).astype(np.uint64)
    t = O.LzssPlainContainer(d2, l2, c.n_chunks, chunk, c.raw_bytes)
    _, st_ref = _oracle(t, x.size)
    _, st = _run(gpu, d2, l2, c.n_chunks, chunk, x.size)
    assert np.array_equal(st, st_ref) and (st_ref != 0).sum() > c.n_chunks // 2


def test_output_limits_and_headers(gpu):
    """dst_bytes shorter than the chunks (the last decoded chunk clipped, the
    ones past it empty and OK), longer than the streams hold (INPUT_OVERRUN),
    lens out of order or past data_bytes (BAD_HEADER)."""
    chunk = 4096
    x = O.gen_lzss_markov(10 * chunk - 500, 0x31)
    c = O.lzss_plain_encode(x, chunk, threads=8)
    for dst_bytes in (0, 1, 5 * chunk - 3, x.size):
        got, st = _run(gpu, c.data, c.lens, c.n_chunks, chunk, dst_bytes)
        _, st_ref = _oracle(c, dst_bytes)
        assert np.array_equal(st, st_ref)
        for i in range(c.n_chunks):                # a clip inside a match is OUTPUT_OVERRUN
            lo, hi = i * chunk, min((i + 1) * chunk, dst_bytes)
            if st[i] == 0:
                assert np.array_equal(got[lo:hi], x[lo:hi])
            else:
                assert st[i] == 1 and hi < (i + 1) * chunk
    got, st = _run(gpu, c.data, c.lens, c.n_chunks, chunk, 10 * chunk)
    _, st_ref = _oracle(c, 10 * chunk)
    assert np.array_equal(st, st_ref) and st[-1] == 2 and (st[:-1] == 0).all()
    lens = c.lens.copy()
    lens[3], lens[4] = lens[4], lens[3]
    _, st = _run(gpu, c.data, lens, c.n_chunks, chunk, x.size)
    assert st[4] == -11
    _, st = _run(gpu, c.data, c.lens, c.n_chunks, chunk, x.size, data_bytes=int(c.lens[-1]) - 1)
    assert st[-1] == -11 and (st[:-1] == 0).all()
    # an empty stream for a non-empty chunk
    _, st = _run(gpu, np.zeros(0, np.uint8), np.zeros(2, np.uint64), 2, chunk, 2 * chunk)
    assert (st == 2).all()
