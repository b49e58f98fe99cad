"""CPU oracle: RLE v2 (ORC RLEv2, theirs side) -- KATs, independent decoder,
sub-mode coverage, PATCHED_BASE with wide patches."""
import numpy as np
import pytest

from oracle import pyoracle as O

DEC_W = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24,
         26, 28, 30, 32, 40, 48, 56, 64]


def closest(bits):
    for w in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24,
              26, 28, 30, 32, 40, 48, 56, 64):
        if bits <= w:
            return w


class Bits:
    """MSB-first bit reader, byte aligned at each run (ORC RLEv2)."""

    def __init__(self, s, p):
        self.s, self.p, self.acc, self.n = s, p, 0, 0

    def read(self, w):
        while self.n < w:
            self.acc = (self.acc << 8) | self.s[self.p]
            self.p += 1
            self.n += 8
        self.n -= w
        v = self.acc >> self.n
        self.acc &= (1 << self.n) - 1
        return v


def uvar(s, p):
    v, sh = 0, 0
    while True:
        b = s[p]
        p += 1
        v |= (b & 0x7F) << sh
        sh += 7
        if b < 0x80:
            return v, p


def py_decode(s: bytes, n: int, bits=64, modes=None):
    """Independent RLEv2 decoder (ORC spec shape + the reference's unsigned,
    raw-delta variant).  Patches use the full-width mask."""
    M = (1 << bits) - 1
    out, p = [], 0
    while len(out) < n:
        f = s[p]
        p += 1
        kind = f >> 6
        if modes is not None:
            modes[kind] = modes.get(kind, 0) + 1
        if kind == 0:                               # SHORT_REPEAT
            nb, cnt = ((f >> 3) & 7) + 1, (f & 7) + 3
            v = int.from_bytes(s[p:p + nb], "big")
            p += nb
            out += [v & M] * cnt
            continue
        ln = (((f & 1) << 8) | s[p]) + 1
        p += 1
        w = DEC_W[(f >> 1) & 0x1F]
        if kind == 1:                               # DIRECT
            br = Bits(s, p)
            out += [br.read(w) & M for _ in range(ln)]
            p = br.p
        elif kind == 3:                             # DELTA
            base, p = uvar(s, p)
            z, p = uvar(s, p)
            d = (z >> 1) ^ -(z & 1)
            cur = base & M
            vals = [cur, (cur + d) & M]
            cur = vals[-1]
            br = Bits(s, p)
            sign = 1 if (d & M) != 0 else -1
            for _ in range(ln - 2):
                cur = (cur + (sign * br.read(w) if (f >> 1) & 0x1F else d)) & M
                vals.append(cur)
            p = br.p if (f >> 1) & 0x1F else p
            out += vals
        else:                                       # PATCHED_BASE
            t3, t4 = s[p], s[p + 1]
            p += 2
            bw, pw = ((t3 >> 5) & 7) + 1, DEC_W[t3 & 0x1F]
            pgw, pll = ((t4 >> 5) & 7) + 1, t4 & 0x1F
            base = int.from_bytes(s[p:p + bw], "big")
            p += bw
            br = Bits(s, p)
            vals = [br.read(w) for _ in range(ln)]
            br2 = Bits(s, br.p)
            gap = 0
            for _ in range(pll):
                e = br2.read(closest(pw + pgw))
                gap += e >> pw
                vals[gap] |= (e & ((1 << pw) - 1)) << w
            p = br2.p
            out += [(v + base) & M for v in vals]
    return out, p


# ---- known-answer vectors (hand-derived from the RLE v2 header grammar) ----
@pytest.mark.parametrize("vals,expect", [
    ([7, 7, 7, 7, 7], "0207"),                                          # SHORT_REPEAT
    (list(range(0, 21, 2)), "4e0a00020406080a0c0e101214"),             # DIRECT (safe-subtract quirk)
    (list(range(1000, 1021, 2)), "c00ae80704"),                         # DELTA, fixed
])
def test_kat(vals, expect):
    s = O.rlev2_lane_encode(vals)
    assert s.hex() == expect
    assert py_decode(s, len(vals))[0] == vals


def test_short_repeat_value_bytes_and_long_repeat_as_delta():
    s = O.rlev2_lane_encode([0x123456] * 10)
    assert s == bytes([(2 << 3) | 7, 0x12, 0x34, 0x56])                # 3 value bytes, count 10
    s = O.rlev2_lane_encode([9] * 40)                                   # > 10 repeats: fixed DELTA
    assert s[0] >> 6 == 3 and (s[0] >> 1) & 0x1F == 0 and s[1] == 39


def test_patched_base_and_wide_patch():
    rng = np.random.default_rng(5)
    vals = [int(v) for v in rng.integers(0, 1 << 12, 100)]
    vals[17] = (1 << 50) + 12345                 # one 51-bit outlier
    vals[60] = (1 << 45) + 7
    s = O.rlev2_lane_encode(vals)
    assert s[0] >> 6 == 2                        # PATCHED_BASE
    pw = DEC_W[s[2] & 0x1F]
    assert pw >= 32                              # patch width the reference's 32-bit mask truncates
    got, used = py_decode(s, len(vals))
    assert got == vals and used == len(s)
    st, dec = O.rlev2_lane_decode(s, len(vals))
    assert st == 0 and dec.tolist() == vals
    # the reference decoder's 32-bit patch mask (decoder_new.h:533) would
    # keep only the low 32 bits of each patch: documented divergence
    trunc = list(vals)
    for i in (17, 60):
        red = vals[i] - min(vals)
        w = DEC_W[(s[0] >> 1) & 0x1F]
        lowbits = red & ((1 << w) - 1)
        patch = (red >> w) & 0xFFFFFFFF
        trunc[i] = ((lowbits | (patch << w)) + min(vals)) & ((1 << 64) - 1)
    assert trunc != vals


def test_random_lanes_vs_independent_decoder():
    rng = np.random.default_rng(11)
    modes = {}
    x = O.gen_rlev2_i64(20000, 77)
    for start in range(0, 20000, 2500):
        seg = [int(v) for v in x[start:start + 2500]]
        s = O.rlev2_lane_encode(seg)
        got, used = py_decode(s, len(seg), modes=modes)
        assert got == seg and used == len(s)
        st, dec = O.rlev2_lane_decode(s, len(seg))
        assert st == 0 and dec.tolist() == seg
    assert set(modes) == {0, 1, 2, 3}            # all four sub-modes occur
    for _ in range(30):                          # small-width randoms, repeats, ramps
        n = int(rng.integers(1, 600))
        kind = rng.integers(0, 3)
        if kind == 0:
            seg = rng.integers(0, 4, n)
        elif kind == 1:
            seg = np.repeat(rng.integers(0, 1 << 40, max(1, n // 7)), 7)[:n]
        else:
            seg = np.cumsum(rng.integers(-3, 4, n)) + (1 << 33)
        seg = [int(v) & ((1 << 64) - 1) for v in seg]
        s = O.rlev2_lane_encode(seg)
        assert py_decode(s, len(seg))[0] == seg


@pytest.mark.parametrize("dtype,ru", [(np.uint64, 1), (np.uint64, 64), (np.uint64, 128),
                                      (np.uint32, 8), (np.uint16, 16), (np.uint8, 4)])
def test_container_roundtrip(dtype, ru):
    x = O.gen_rlev2_i64(70_000, 3).astype(dtype)
    c = O.rlev2_encode(x, 32768, ru, threads=4)
    assert int(c.blk_off[-1]) == x.nbytes       # the reference keeps the raw size in blk_off[n]
    assert all(int(c.blk_off[i]) % 128 == 0 for i in range(c.n_chunks))
    rc, out, st = O.rlev2_decode(c, threads=4)
    assert rc == 0 and np.array_equal(out, x.view(np.uint8))


def test_decode_status_on_truncation():
    s = O.rlev2_lane_encode([int(v) for v in O.gen_rlev2_i64(3000, 9)])
    st, _ = O.rlev2_lane_decode(s[: len(s) // 2], 3000)
    assert st == 2
    st, _ = O.rlev2_lane_decode(s, 100)           # first run longer than the lane
    assert st in (0, 1)
