"""CPU oracle: byte RLE (include/brle/brle.h) -- KATs, special cases, roundtrip."""
import numpy as np
import pytest

from oracle import pyoracle as O


def py_decode(s: bytes) -> bytes:
    out, p = bytearray(), 0
    while p < len(s):
        c = s[p]
        p += 1
        if c > 127:
            out += s[p:p + c - 127]
            p += c - 127
        else:
            out += bytes([s[p]]) * (c + 3)
            p += 1
    return bytes(out)


@pytest.mark.parametrize("data,expect", [
    (b"aaaab", "01618062"),           # run of 4 'a' (c=1), literal 'b' (c=128)
    (b"ab", "816162"),
    (b"aaa", "0061"),
    (b"abbb", "80610062"),            # literal then run of 3
    (b"", ""),
])
def test_kat(data, expect):
    assert O.brle_chunk_encode(data).hex() == expect


def test_all_equal_runs_of_130():
    s = O.brle_chunk_encode(b"\x07" * 300)
    # 130 + 130 + 40 -> headers 127, 127, 37
    assert s == bytes([127, 7, 127, 7, 37, 7])


def test_no_repeat_literal_groups_of_128():
    data = bytes((i * 7 + 3) % 256 for i in range(300))
    s = O.brle_chunk_encode(data)
    assert s[0] == 255 and s[129] == 255 and s[258] == 127 + 44
    assert py_decode(s) == data


def test_random_roundtrip_and_independent_decoder():
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = int(rng.integers(1, 3000))
        data = rng.integers(0, 3, n).astype(np.uint8).tobytes()
        s = O.brle_chunk_encode(data)
        assert py_decode(s) == data
        st, got = O.brle_chunk_decode(s, n)
        assert st == 0 and got == data


def test_container_and_status():
    x = O.gen_brle_u8(1_000_003, 0x4D445A02)
    c = O.brle_encode(x, 65536, threads=4)
    assert c.n_chunks == 16 and c.ptr[0] == 0
    rc, out, st = O.brle_decode(c, threads=4)
    assert rc == 0 and np.array_equal(out, x)
    # run lengths average about 12 -> about 2 bytes per 12
    assert 0.12 < c.data.size / x.size < 0.22
    st, _ = O.brle_chunk_decode(bytes([5]), 100)       # run header without its byte
    assert st == 2
    st, _ = O.brle_chunk_decode(bytes([5, 1]), 4)       # run of 8 into 4 bytes
    assert st == 1
