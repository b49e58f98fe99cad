"""CPU oracle: LZSS lane-transposed (include/lzss/lzss_gpu2.h) -- KATs, the
fast matcher against the literal find_match scan, an independent decoder and
container builder, statuses."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests.lzss_util import build_container, lane_view, py_lane_decode, read_container


# ---- known answers (C8, C9 hand-derived from find_match's rules; the third
# by hand: literals until a 3-byte match fits inside the history)
@pytest.mark.parametrize("lane,expect", [
    (b"abcabc", "3f616263616263"),                       # offset 0 (oldest byte) never tried
    (b"abcdefabcdef", "7f61626364656661c20000"),          # 7 literals, match len 5 dist 6, reserved byte
    (b"aaaaaaaa", "2f61616161600061"),                    # match needs len <= dist: (3,3) at byte 4
])
def test_kat(lane, expect):
    for naive in (False, True):
        assert O.lzss_lane_encode(lane, naive=naive).hex() == expect
    st, out = O.lzss_lane_decode(bytes.fromhex(expect), len(lane))
    assert st == 0 and out == lane
    assert py_lane_decode(bytes.fromhex(expect), len(lane)) == (0, lane)


def _tokens(s: bytes, out_len: int):
    toks, p, o = [], 0, 0
    while o < out_len:
        f = s[p]
        p += 1
        for b in range(8):
            if o >= out_len:
                break
            if f >> b & 1:
                toks.append(("L", o, s[p]))
                p += 1
                o += 1
            else:
                v = s[p] | s[p + 1] << 8
                p += 2
                toks.append(("M", o, (v & 31) + 3, v >> 5))
                o += (v & 31) + 3
    return toks


def test_tie_goes_to_oldest_and_cap_34():
    lane = bytes([200]) + b"abcd" + bytes(range(100, 140)) + b"abcd" + bytes(range(140, 180)) + b"abcdZ"
    toks = _tokens(O.lzss_lane_encode(lane), len(lane))
    m = [t for t in toks if t[0] == "M"]
    assert m[0][1:] == (45, 4, 44)                 # "abcd" again, from byte 1
    assert m[1][1:] == (89, 4, 88)                 # two 4-byte candidates: the older wins
    rep = bytes(range(1, 61))
    toks = _tokens(O.lzss_lane_encode(rep + bytes([0]) + rep), 121)
    assert [t for t in toks if t[0] == "M"][0][2] == 34    # a 60-byte repeat starts with a 34-byte match


@pytest.mark.parametrize("seed", range(6))
def test_fast_matcher_equals_literal_scan(seed):
    rng = np.random.default_rng(seed)
    for _ in range(20):
        n = int(rng.integers(1, 4097))
        k = int(rng.integers(1, 6))
        if rng.random() < 0.3:
            lane = np.repeat(rng.integers(0, k + 1, n // 5 + 1), 5)[:n].astype(np.uint8).tobytes()
        else:
            lane = rng.integers(0, k + 1, n).astype(np.uint8).tobytes()
        fast = O.lzss_lane_encode(lane)
        assert fast == O.lzss_lane_encode(lane, naive=True)
        assert py_lane_decode(fast, n) == (0, lane)


def test_markov_lanes_roundtrip_and_token_mix():
    x = O.gen_lzss_markov(1 << 18, 7)
    n_lit = n_mat = 0
    for c in range(4):
        ch = x[c * 65536:(c + 1) * 65536]
        for t in range(0, 32, 5):
            lane = lane_view(ch, t)
            s = O.lzss_lane_encode(lane)
            assert py_lane_decode(s, len(lane)) == (0, lane)
            toks = _tokens(s, len(lane))
            n_lit += sum(1 for q in toks if q[0] == "L")
            n_mat += sum(1 for q in toks if q[0] == "M")
    assert n_mat > 0.2 * (n_lit + n_mat)          # the Markov text compresses per lane


def test_generator_shape():
    x = O.gen_lzss_markov(1 << 20, 3)
    assert x.max() < 64
    assert np.array_equal(x, O.gen_lzss_markov(1 << 20, 3))
    assert not np.array_equal(x[:4096], O.gen_lzss_markov(4096, 4))


@pytest.mark.parametrize("chunk", [4096, 65536, 131072])
def test_container_matches_independent_builder(chunk):
    x = O.gen_lzss_markov(3 * chunk + 1000, 11)          # trailing partial chunk is dropped
    c = O.lzss_encode(x, chunk, threads=4)
    assert c.n_chunks == 3
    streams = [[O.lzss_lane_encode(lane_view(x[i * chunk:(i + 1) * chunk], t)) for t in range(32)]
               for i in range(3)]
    pay, cl, bo, cm = build_container(streams)
    assert np.array_equal(c.col_len, cl) and np.array_equal(c.col_map, cm)
    assert np.array_equal(c.blk_off, bo) and np.array_equal(c.payload, pay)
    for i in range(3):
        L = c.col_len[32 * i: 32 * i + 32]
        assert (np.diff(L.astype(np.int64)) <= 0).all()
        assert sorted(c.col_map[32 * i: 32 * i + 32].tolist()) == list(range(32))
    rc, out, st = O.lzss_decode(c, threads=4)
    assert rc == 0 and np.array_equal(out, x[:3 * chunk])
    assert read_container(c.payload, c.col_len, c.blk_off, c.col_map, 3) == streams


def test_sort_ties_to_higher_lane():
    chunk = np.zeros(4096, np.uint8)                      # every lane identical -> equal lengths
    c = O.lzss_encode(chunk, 4096, threads=1)
    assert c.col_map.tolist() == list(range(31, -1, -1))


def test_meta_layout():
    x = O.gen_lzss_markov(2 * 4096, 2)
    c = O.lzss_encode(x, 4096, threads=2)
    m = c.meta_bytes()
    k = int(np.frombuffer(m[:8], "<u8")[0])
    lens = np.frombuffer(m[8:8 + 8 * k], "<u8")
    assert k == 4 and lens.tolist() == [8, 8 * 3, 8 * 64, 64]
    assert len(m) == 8 + 8 * k + int(lens.sum())


def test_statuses():
    lane = lane_view(O.gen_lzss_markov(4096, 5), 3)
    s = O.lzss_lane_encode(lane)
    assert O.lzss_lane_decode(s[: len(s) // 2], len(lane))[0] == 2        # truncated
    assert O.lzss_lane_decode(b"\x00\x40\x00", 4)[0] == -9               # distance 2 before any output
    assert O.lzss_lane_decode(b"\x01a\x20\x00", 4)[0] == 0               # overlapping match: a x4
    assert O.lzss_lane_decode(b"\x01a\x20\x00", 4)[1] == b"aaaa"
    assert O.lzss_lane_decode(b"\x01a\x21\x00", 4)[0] == 1               # len 4 after 1 byte: overrun
    rng = np.random.default_rng(1)
    for _ in range(200):                                                 # random corruption
        b = bytearray(s)
        for _ in range(int(rng.integers(1, 4))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        b = bytes(b[: int(rng.integers(1, len(b) + 1))])
        st, out = O.lzss_lane_decode(b, len(lane))
        pst, pout = py_lane_decode(b, len(lane))
        assert st == pst
        if st == 0:
            assert out == pout


def test_container_statuses():
    x = O.gen_lzss_markov(4 * 4096, 8)
    c = O.lzss_encode(x, 4096, threads=2)
    cm = c.col_map.copy()
    cm[32 + 3] = cm[32 + 4]                                              # chunk 1: not a permutation
    cl = c.col_len.copy()
    cl[64 + 5], cl[64 + 6] = cl[64 + 6], cl[64 + 5] + 1                  # chunk 2: not sorted
    bad = O.LzssContainer(c.payload, cl, c.blk_off, cm, c.n_chunks, c.chunk)
    rc, out, st = O.lzss_decode(bad, threads=2)
    assert st[0] == 0 and st[1] == -11 and st[2] == -11 and st[3] == 0
    assert np.array_equal(out[:4096], x[:4096])


# ------------------------------------------------------------ plain format --
@pytest.mark.parametrize("size,chunk", [(0, 4096), (1, 4096), (4096, 4096), (10_000, 4096),
                                        (3 * 65536 + 12_345, 65536)])
def test_plain_container_layout_and_roundtrip(size, chunk):
    """Plain LZSS (compress_func + kernel_shift_data, lzss_gpu2.h:1450-1566,
    1926-1948): each chunk's stream is the chunk run through the lane grammar,
    streams are packed back to back, lens is the inclusive prefix of their
    lengths; the last chunk may be short.  Decoded by the independent Python
    decoder and by the oracle."""
    x = O.gen_lzss_markov(size, 0x51 + size)
    c = O.lzss_plain_encode(x, chunk)
    n = -(-size // chunk)
    assert c.n_chunks == n and c.lens.size == n
    pos = 0
    for i in range(n):
        part = x[i * chunk:(i + 1) * chunk].tobytes()
        s = c.stream(i)
        assert s == O.lzss_lane_encode(part)
        assert int(c.lens[i]) == pos + len(s)
        pos += len(s)
        st, got = py_lane_decode(s, len(part))
        assert st == 0 and got == part
    rc, got, status = O.lzss_plain_decode(c)
    assert rc == 0 and (status == 0).all() and got.tobytes() == x.tobytes()


def test_plain_stream_equals_literal_scan():
    """The plain chunk stream of the fast matcher equals the one from the
    literal find_match scan (lzss_gpu2.h:70-109) on whole chunks."""
    x = O.gen_lzss_markov(4096 * 3, 0x77)
    c = O.lzss_plain_encode(x, 4096)
    for i in range(3):
        part = x[i * 4096:(i + 1) * 4096].tobytes()
        assert c.stream(i) == O.lzss_lane_encode(part, naive=True)


def test_plain_statuses():
    """Damaged plain streams: cut short -> input overrun (2); a distance past
    the chunk start -> bad distance (-9); statuses per chunk."""
    x = O.gen_lzss_markov(3 * 4096, 0x78)
    c = O.lzss_plain_encode(x, 4096)
    d = c.data.copy()
    lens = c.lens.copy()
    lens[0] -= 5                                    # chunk 0 loses its tail; chunk 1 starts early
    bad = O.LzssPlainContainer(d, lens, c.n_chunks, c.chunk, c.raw_bytes)
    rc, _, st = O.lzss_plain_decode(bad)
    assert st[0] == 2 and st[2] == 0
    first = O.lzss_plain_encode(np.frombuffer(b"ab" * 100, np.uint8), 4096)
    s = bytearray(first.stream(0))
    # the first match token: distance field set past the output produced so far
    flags, q = s[0], 1
    for b in range(8):
        if flags >> b & 1:
            q += 1
            continue
        v = s[q] | s[q + 1] << 8
        v = (v & 31) | (2000 << 5)
        s[q], s[q + 1] = v & 255, v >> 8
        break
    c2 = O.LzssPlainContainer(np.frombuffer(bytes(s), np.uint8), np.array([len(s)], np.uint64), 1, 4096, 200)
    assert O.lzss_plain_decode(c2)[2][0] == -9
