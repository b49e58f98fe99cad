"""CPU oracle: RLE v1 (lane-transposed) -- KATs, invariants, independent decoder."""
import numpy as np
import pytest

from oracle import pyoracle as O


# ---- known-answer vectors, hand-derived from include/rle_v1/encoder.h ----
@pytest.mark.parametrize("vals,dtype,expect", [
    ([5, 5, 5, 5], np.uint32, "010005"),            # run: h=1, d=0, base 5
    ([1, 1000, 7], np.uint32, "fd01e80707"),        # 3 literals, header back-patched
    ([10, 9, 8, 7], np.uint32, "fc0a090807"),       # u32: negative delta never starts a run
    ([10, 9, 8, 7], np.uint16, "01ff0a"),           # u16: promoted to int, run with d=-1
    ([10, 9, 8, 7], np.uint64, "01ff0a"),           # u64: int64 conversion wraps, run with d=-1
    ([7, 7], np.uint32, "fe0707"),                  # a lone pair ends as 2 literals
])
def test_lane_kat(vals, dtype, expect):
    assert O.rlev1_lane_encode(np.array(vals, dtype)).hex() == expect


def test_run_cap_129():
    # 200 equal values: a run of 129 (header 126 = 0x7e), then the rest
    s = O.rlev1_lane_encode(np.full(200, 3, np.uint32))
    assert s[:3] == bytes([0x7E, 0x00, 0x03])
    assert s[3:] == bytes([71 - 3, 0x00, 0x03])
    st, v = O.rlev1_lane_decode(s, 200, np.uint32)
    assert st == 0 and (v == 3).all()


def test_literal_groups_of_127():
    vals = (np.arange(300, dtype=np.uint32) * 1000) % 100003
    s = O.rlev1_lane_encode(vals)
    assert s[0] == 0x81   # -127
    st, v = O.rlev1_lane_decode(s, 300, np.uint32)
    assert st == 0 and np.array_equal(v, vals)


# ---- an independent decoder: serial de-interleave, then a bytewise parse ----
def py_deinterleave(chunk: bytes, col_len):
    words = [(c + 3) // 4 for c in col_len]
    lanes = [bytearray() for _ in range(32)]
    pos = 0
    for r in range(max(words)):
        for t in range(32):
            if r < words[t]:
                lanes[t] += chunk[pos:pos + 4]
                pos += 4
    return [bytes(lanes[t][:col_len[t]]) for t in range(32)]


def py_leb(s, p):
    v, sh = 0, 0
    while True:
        b = s[p]
        p += 1
        v |= (b & 0x7F) << sh
        sh += 7
        if b < 0x80:
            return v, p


def py_lane_decode(s: bytes, n: int, bits: int):
    out, p, m = [], 0, (1 << bits) - 1
    while len(out) < n:
        h = s[p]
        p += 1
        if h >= 0x80:
            for _ in range(256 - h):
                v, p = py_leb(s, p)
                out.append(v & m)
        else:
            d = s[p] - 256 if s[p] >= 0x80 else s[p]
            p += 1
            base, p = py_leb(s, p)
            out.extend(((base + i * d) & m) for i in range(h + 3))
    return out


@pytest.mark.parametrize("dtype,chunk,line", [(np.uint32, 4096, 32), (np.uint16, 2048, 16),
                                              (np.uint8, 1024, 8), (np.uint64, 8192, 64)])
def test_container_vs_independent_decoder(dtype, chunk, line):
    rng = np.random.default_rng(7)
    n = (3 * chunk + chunk // 3) // np.dtype(dtype).itemsize
    base = rng.integers(0, 200, n).astype(dtype)
    base[::7] = base[::7] + 3            # mix runs and literals
    base[100:300] = 42
    c = O.rlev1_encode(base, chunk, line, threads=2)
    raw = base.view(np.uint8)
    padded = np.zeros(c.n_chunks * chunk, np.uint8)
    padded[:raw.size] = raw
    eb = np.dtype(dtype).itemsize
    for ci in range(c.n_chunks):
        cl = [int(x) for x in c.col_len[32 * ci:32 * ci + 32]]
        lo, hi = int(c.blk_off[ci]), int(c.blk_off[ci + 1])
        # structural invariants (encoder.h:540-559)
        assert lo % 128 == 0 and hi - lo == (sum((x + 3) // 4 * 4 for x in cl) + 127) // 128 * 128
        lanes = py_deinterleave(c.payload[lo:hi].tobytes(), cl)
        for t in range(32):
            vals = py_lane_decode(lanes[t], chunk // 32 // eb, 8 * eb)
            got = np.array(vals, dtype).view(np.uint8)
            expect = np.concatenate([padded[ci * chunk + (ln * 32 + t) * line:
                                            ci * chunk + (ln * 32 + t + 1) * line]
                                     for ln in range(chunk // 32 // line)])
            assert np.array_equal(got, expect), (ci, t)
    rc, out, st = O.rlev1_decode(c, threads=2)
    assert rc == 0 and np.array_equal(out, raw)


def test_config1_generator_roundtrip_small():
    x = O.gen_rlev1_u32(1 << 18, 0x4D445A01)
    assert x.max() < (1 << 32)
    for line in (32, 128, 512):
        c = O.rlev1_encode(x, 131072, line, threads=4)
        assert c.n_chunks == 8
        rc, out, _ = O.rlev1_decode(c, threads=4)
        assert rc == 0 and np.array_equal(out, x.view(np.uint8))


def test_decode_reports_truncation():
    s = O.rlev1_lane_encode(np.arange(64, dtype=np.uint32) * 977)
    st, _ = O.rlev1_lane_decode(s[:-2], 64, np.uint32)
    assert st == 2
    st, _ = O.rlev1_lane_decode(s, 63, np.uint32)   # last token overruns the lane
    assert st in (0, 1)


def _splitmix(x):
    M = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & M
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return x, z ^ (z >> 31)


def test_rng_matches_published_generators():
    # splitmix64 from 0 starts e220a8397b1dcdaf (Vigna's reference output)
    _, z = _splitmix(0)
    assert z == 0xE220A8397B1DCDAF
    # xoshiro256** per its published definition, seeded by 4 splitmix64 draws
    M = (1 << 64) - 1
    x, s = 0x4D445A01, []
    for _ in range(4):
        x, z = _splitmix(x)
        s.append(z)
    rotl = lambda v, k: ((v << k) | (v >> (64 - k))) & M
    ref = []
    for _ in range(6):
        ref.append(rotl((s[1] * 5) & M, 7) * 9 & M)
        t = (s[1] << 17) & M
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t
        s[3] = rotl(s[3], 45)
    assert O.rng_stream(0x4D445A01, 6).tolist() == ref
