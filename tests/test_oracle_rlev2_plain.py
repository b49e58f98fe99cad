"""Plain RLE v2 (one run stream per chunk of int64 values,
include/rlev2/decoder.h:190-262, encoder.h:808-1058 HEAD side): the oracle
decoder and the host encoder against the ORC RLEv2 specification's worked
examples, an independent Python decoder, round trips and crafted faults."""
import numpy as np
import pytest

from oracle import pyoracle as O

# The four worked examples of the ORC file-format specification (integer
# run-length encoding, version 2): the encoded bytes and the values they
# stand for.  They pin the decoder's field layout and bit order, and, as the
# reference encoder follows ORC's writer, the encoder's run choices too.
ORC_EXAMPLES = [
    ("short_repeat", [10000] * 5, "0a2710"),
    ("direct", [23713, 43806, 57005, 48879], "5e035ca1ab1edeadbeef"),
    ("delta", [2, 3, 5, 7, 11, 13, 17, 19, 23, 29], "c6090202224242 46".replace(" ", "")),
    ("patched_base", [2030, 2000, 2020, 1000000, 2040, 2050, 2060, 2070, 2080, 2090, 2100, 2110, 2120,
                      2130, 2140, 2150, 2160, 2170, 2180, 2190],
     "8e132b2107d01e00147028323c46505a646e78828c96a0aab4befce8"),
]


@pytest.mark.parametrize("name,vals,hexs", ORC_EXAMPLES, ids=[e[0] for e in ORC_EXAMPLES])
def test_orc_spec_examples_decode(name, vals, hexs):
    st, got = O.rlev2_plain_chunk_decode(bytes.fromhex(hexs), 4096)
    assert st == 0 and got.tolist() == vals


@pytest.mark.parametrize("name,vals,hexs", ORC_EXAMPLES, ids=[e[0] for e in ORC_EXAMPLES])
def test_orc_spec_examples_encode(name, vals, hexs):
    assert O.rlev2_plain_chunk_encode(vals).hex() == hexs


# ---------------------------------------------------- independent decoder --
W = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 26, 28, 30,
     32, 40, 48, 56, 64]


def closest(b):
    return next(x for x in W if x >= max(b, 1))


def py_decode(s: bytes) -> list[int]:
    """ORC RLEv2 as the specification states it, on unsigned 64-bit values
    (valid streams only)."""
    out, p, M = [], 0, (1 << 64) - 1

    def bits(p, n, w):   # n MSB-first fields of w bits from byte p
        acc = int.from_bytes(s[p:p + (n * w + 7) // 8], "big")
        tot = ((n * w + 7) // 8) * 8
        return [(acc >> (tot - (i + 1) * w)) & ((1 << w) - 1) for i in range(n)], p + (n * w + 7) // 8

    def varint(p):
        v, sh = 0, 0
        while True:
            b = s[p]
            p += 1
            v |= (b & 0x7f) << sh
            sh += 7
            if not b & 0x80:
                return v & M, p

    while p < len(s):
        f = s[p]
        k = f >> 6
        if k == 0:
            nb, cnt = ((f >> 3) & 7) + 1, (f & 7) + 3
            out += [int.from_bytes(s[p + 1:p + 1 + nb], "big")] * cnt
            p += 1 + nb
            continue
        w, n = W[(f >> 1) & 31], (((f & 1) << 8) | s[p + 1]) + 1
        if k == 1:
            v, p = bits(p + 2, n, w)
            out += v
        elif k == 3:
            base, q = varint(p + 2)
            zz, q = varint(q)
            d = (zz >> 1) ^ -(zz & 1)
            vals = [base, (base + d) & M]
            if (f >> 1) & 31:
                ds, q = bits(q, n - 2, w)
                for x in ds:
                    vals.append((vals[-1] + x if d > 0 else vals[-1] - x) & M)
            else:
                for _ in range(n - 2):
                    vals.append((vals[-1] + d) & M)
            out += vals
            p = q
        else:
            t, fo = s[p + 2], s[p + 3]
            bw, pw, pgw, pll = ((t >> 5) & 7) + 1, W[t & 31], ((fo >> 5) & 7) + 1, fo & 31
            base = int.from_bytes(s[p + 4:p + 4 + bw], "big")
            v, q = bits(p + 4 + bw, n, w)
            pl, q = bits(q, pll, closest(pw + pgw))
            g = 0
            for e in pl:
                g += e >> pw
                v[g] |= (e & ((1 << pw) - 1)) << w
            out += [(x + base) & M for x in v]
            p = q
    return out


def _col(n, seed):
    return O.gen_rlev2_i64(n, seed).astype(np.int64) & np.int64(0x7FFFFFFFFFFFFFFF)


def test_random_chunks_three_ways():
    """Encoder output decoded by the C oracle and by the Python decoder: both
    give the input back."""
    rng = np.random.default_rng(5)
    for seed in range(12):
        x = _col(4096, 100 + seed)
        if seed % 3 == 1:                      # descending stretches: negative delta bases
            x = np.sort(x)[::-1].copy()
        if seed % 3 == 2:                      # small values with zeros (no INT8 safety test)
            x = rng.integers(0, 130, 4096).astype(np.int64)
        s = O.rlev2_plain_chunk_encode(x)
        st, got = O.rlev2_plain_chunk_decode(s, 4096)
        assert st == 0 and np.array_equal(got, x)
        assert py_decode(s) == x.astype(np.uint64).tolist()


def test_descending_runs_use_negative_delta_base():
    x = np.array([1000 - 3 * i - (i % 2) for i in range(40)], np.int64)
    s = O.rlev2_plain_chunk_encode(x)
    assert s[0] >> 6 == 3                      # one DELTA run
    zz = s[4]                                  # base 1000 is a 2-byte varint at [2:4]
    assert zz & 1                              # zigzag of a negative initial delta
    st, got = O.rlev2_plain_chunk_decode(s, 64)
    assert st == 0 and np.array_equal(got, x)


def test_container_layout_and_round_trip():
    x = _col(3 * 4096 + 1000, 7)
    raw = x.view(np.uint8)
    tail = np.concatenate([raw, np.array([1, 2, 3], np.uint8)])     # 3 bytes past the last int64
    c = O.rlev2_plain_encode(tail, 32768)
    assert c.n_chunks == 4 and c.ptr[0] == 0 and c.raw_bytes == tail.size
    assert np.all(np.diff(c.ptr.astype(np.int64)) > 0) and int(c.ptr[-1]) == c.data.size
    rc, out, st = O.rlev2_plain_decode(c.blob, 32768)
    assert rc == 0 and (st == 0).all()
    assert np.array_equal(out[:raw.size], raw)     # the last in_bytes % 8 bytes are not encoded


def test_negative_values_rejected():
    with pytest.raises(ValueError):
        O.rlev2_plain_encode(np.array([5, -1, 7], np.int64), 32768)


def test_faults():
    dec = O.rlev2_plain_chunk_decode
    ok = O.rlev2_plain_chunk_encode(np.arange(100, 140, dtype=np.int64))       # one DELTA run
    assert dec(ok, 64)[0] == 0
    assert dec(ok[:1], 64)[0] == 2                           # header cut: input overrun
    assert dec(ok, 39)[0] == 1                               # 40 values into 39: output overrun
    assert dec(bytes([0xC0, 0x00, 0x05, 0x02]), 64)[0] == -10        # DELTA of length 1
    assert dec(bytes([0xC0, 0x05]) + b"\xff" * 10 + b"\x00", 64)[0] == -10   # 11-byte varint
    d = O.rlev2_plain_chunk_encode([23713, 43806, 57005, 48879])
    assert dec(d[:-1], 64)[0] == 2                           # DIRECT data cut
    pb = bytearray.fromhex(ORC_EXAMPLES[3][2])
    assert dec(bytes(pb[:-1]), 64)[0] == 2                   # patch list cut
    pb = pb[:-2] + b"\xff" * 13                              # 7 entries, gaps of 3: 21 >= 20 values
    pb[3] = 0x27
    assert dec(bytes(pb), 64)[0] == -10
    assert dec(bytes.fromhex("0a27"), 64)[0] == 2            # SHORT_REPEAT value cut
    # values before the failing run are decoded
    st, got = dec(bytes.fromhex("0a2710") + bytes([0xC0, 0x05]), 64)
    assert st == 2 and got.tolist() == [10000] * 5
