"""The oracle's container decoders on damaged input: random payload bit
flips, and lane streams cut short or lengthened within their chunk's extent.
Each decode must return a status from the codes below for every chunk and
never read or write outside its buffers; oracle/asan.sh runs this file (with
the rest of the CPU suite) under AddressSanitizer and UBSan."""
import numpy as np
import pytest

from oracle import pyoracle as O

CODES = {0, 1, 2, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10, -11, -12}


def _flip(a, rng, k):
    a = a.copy()
    if a.size:
        for _ in range(k):
            a[int(rng.integers(0, a.size))] ^= np.uint8(1 << int(rng.integers(0, 8)))
    return a


def _lanes(c, rng):
    """col_len of one chunk changed by -64..+16 bytes, kept within the chunk's
    extent (blk_off[c+1] - blk_off[c] bytes in all)."""
    cl = c.col_len.copy()
    ch = int(rng.integers(0, max(c.n_chunks - 1, 1)))
    ext = int(c.blk_off[ch + 1] - c.blk_off[ch])
    t = int(rng.integers(0, 32))
    v = int(cl[32 * ch + t]) + int(rng.integers(-64, 17))
    v = max(0, v)
    rest = int(cl[32 * ch:32 * ch + 32].sum()) - int(cl[32 * ch + t])
    cl[32 * ch + t] = min(v, max(0, ext - rest))
    return cl


@pytest.mark.parametrize("seed", range(3))
def test_rlev1_damaged(seed):
    rng = np.random.default_rng(seed)
    c = O.rlev1_encode(O.gen_rlev1_u32(1 << 16, 40 + seed), 32768, 32)
    for _ in range(20):
        d = O.RleV1Container(_flip(c.payload, rng, 3), _lanes(c, rng), c.blk_off, c.n_chunks, c.chunk,
                             c.elem_bytes, c.line, c.raw_bytes)
        rc, _, st = O.rlev1_decode(d)
        assert set(int(s) for s in st) <= CODES


@pytest.mark.parametrize("seed", range(3))
def test_rlev2_damaged(seed):
    rng = np.random.default_rng(seed)
    c = O.rlev2_encode(O.gen_rlev2_i64(1 << 15, 50 + seed), 32768, 64)
    for _ in range(20):
        d = O.RleV2Container(_flip(c.payload, rng, 3), _lanes(c, rng), c.blk_off, c.n_chunks, c.chunk,
                             c.elem_bytes, c.read_unit, c.raw_bytes, c.payload_bytes)
        rc, _, st = O.rlev2_decode(d)
        assert set(int(s) for s in st) <= CODES


@pytest.mark.parametrize("seed", range(3))
def test_brle_damaged(seed):
    rng = np.random.default_rng(seed)
    c = O.brle_encode(O.gen_brle_u8(1 << 18, 60 + seed), 32768)
    for _ in range(20):
        d = O.BrleContainer(_flip(c.data, rng, 4), c.ptr, c.n_chunks, c.chunk, c.raw_bytes)
        rc, _, st = O.brle_decode(d)
        assert set(int(s) for s in st) <= CODES


@pytest.mark.parametrize("seed", range(3))
def test_lzss_damaged(seed):
    rng = np.random.default_rng(seed)
    c = O.lzss_encode(O.gen_lzss_markov(1 << 18, 70 + seed), 32768)
    for _ in range(20):
        d = O.LzssContainer(_flip(c.payload, rng, 4), c.col_len, c.blk_off, c.col_map, c.n_chunks, c.chunk)
        rc, _, st = O.lzss_decode(d)
        assert set(int(s) for s in st) <= CODES


def test_inflate_damaged_and_truncated():
    import zlib
    rng = np.random.default_rng(5)
    for i in range(40):
        raw = O.gen_lzss_markov(int(rng.integers(100, 30000)), 80 + i).tobytes()
        z = np.frombuffer(zlib.compress(raw, int(rng.choice([0, 1, 6, 9]))), np.uint8)
        bad = _flip(z, rng, 3)[: int(rng.integers(1, z.size + 1))].tobytes()
        st, _ = O.inflate(bad, len(raw) + int(rng.integers(-50, 50)) % (len(raw) + 1))
        assert st in CODES
