"""Lane-transposed byte RLE (include/brle/brle_trans.h): known answers for the
host encoder, an independent Python restatement of encoder and decoder, and
the lane container shared with LZSS."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests.lzss_util import assemble, build_container, lane_view, read_container


# Known answers, derived by hand from compress_func (brle_trans.h:401-562):
# stream = flag byte, then tokens; bit b of a flag (LSB first) 1 = literal,
# 0 = run (count, value); a token completing a group of 8 reserves the next
# flag byte, the lane's final run does not; the counter is a uint8_t.
KATS = [
    (b"abcd", "0f61626364"),                       # four literals
    (b"aaaaab", "02056162"),                       # run (5,'a'), then literal 'b' (bit 1)
    (b"baaaa", "01620461"),                        # literal, then the final run (4,'a')
    (bytes(range(8)), "ff0001020304050607" + "00"),  # a full group reserves a flag byte
    (b"aab", "07616162"),                          # runs of two stay literals
    (b"a" * 300, "002c61"),                        # 256 lost to the counter wrap: (44,'a')
    (b"a" * 256 + b"b", "0162"),                   # wrap to 0, 'b' restarts the count
    (b"aaa" * 1 + b"b" * 3 + b"c" * 3, "00036103620363"),  # three runs, final one unflagged
]


@pytest.mark.parametrize("lane,expect", KATS)
def test_lane_encoder_known_answers(lane, expect):
    assert O.brle_trans_lane_encode(lane).hex() == expect


def py_encode(lane: bytes) -> bytes:
    """Independent restatement of compress_func for one lane."""
    out = bytearray([0])
    hpos, hdr, blocks = 0, 0, 0
    counter, prev = 0, None

    def token_done():
        nonlocal blocks, hdr, hpos
        blocks += 1
        if blocks == 8:
            out[hpos] = hdr
            hdr = 0
            hpos = len(out)
            out.append(0)
            blocks = 0
        if blocks:
            out[hpos] = hdr

    def literals(v, k):
        nonlocal hdr
        for _ in range(k):
            hdr |= 1 << blocks
            out.append(v)
            token_done()

    for v in lane:
        if counter == 0:
            prev = v
        if prev == v:
            counter = (counter + 1) & 0xff
            continue
        if counter >= 3:
            out += bytes([counter, prev])
            token_done()
        else:
            literals(prev, counter)
        counter, prev = 1, v
    if counter >= 3:
        out += bytes([counter, prev])
    else:
        literals(prev, counter)
    return bytes(out)


def py_decode(s: bytes, out_len: int):
    """Independent decoder with the oracle's statuses (0 ok, 1 output
    overrun, 2 input overrun)."""
    out, p = bytearray(), 0
    while len(out) < out_len:
        if p >= len(s):
            return 2, bytes(out)
        f = s[p]
        p += 1
        for b in range(8):
            if len(out) >= out_len:
                break
            if f >> b & 1:
                if p >= len(s):
                    return 2, bytes(out)
                out.append(s[p])
                p += 1
            else:
                if p + 2 > len(s):
                    return 2, bytes(out)
                cnt, v = s[p], s[p + 1]
                p += 2
                for _ in range(cnt):
                    if len(out) >= out_len:
                        return 1, bytes(out)
                    out.append(v)
    return 0, bytes(out)


def _lanes(rng, n, long_runs=False):
    out = bytearray()
    while len(out) < n:
        k = int(rng.integers(1, 600 if long_runs else 40))
        out += bytes([int(rng.integers(0, 4))]) * k
    return bytes(out[:n])


def test_encoder_matches_restatement():
    rng = np.random.default_rng(1)
    for i in range(300):
        lane = _lanes(rng, int(rng.integers(1, 3000)), long_runs=i % 3 == 0)
        assert O.brle_trans_lane_encode(lane) == py_encode(lane)


def test_decoders_agree_on_valid_and_damaged_streams():
    rng = np.random.default_rng(2)
    for i in range(400):
        lane = _lanes(rng, int(rng.integers(1, 2000)))
        s = bytearray(O.brle_trans_lane_encode(lane))
        assert py_decode(bytes(s), len(lane)) == (0, lane)
        assert O.brle_trans_lane_decode(bytes(s), len(lane)) == (0, lane)
        for _ in range(int(rng.integers(1, 4))):
            s[int(rng.integers(0, len(s)))] = int(rng.integers(0, 256))
        cut = bytes(s[: int(rng.integers(0, len(s) + 1))])
        for t in (bytes(s), cut):
            n = len(lane) + int(rng.integers(-5, 6))
            n = max(n, 0)
            st, o = O.brle_trans_lane_decode(t, n)
            pst, po = py_decode(t, n)
            assert st == pst and (st != 0 or o == po)


@pytest.mark.parametrize("chunk", [128, 4096, 65536])
def test_container_roundtrip(chunk):
    """Lanes are chunk bytes 128k + 4t + j; the container is the LZSS lane
    container (build_container restates its sort and interleave)."""
    rng = np.random.default_rng(chunk)
    x = np.frombuffer(_lanes(rng, 6 * chunk), np.uint8)
    c = O.brle_trans_encode(x, chunk, threads=4)
    assert c.n_chunks == 6
    streams = [[O.brle_trans_lane_encode(lane_view(x[k * chunk:(k + 1) * chunk], t)) for t in range(32)]
               for k in range(6)]
    pay, cl, bo, cm = build_container(streams)
    assert np.array_equal(c.col_len, cl) and np.array_equal(c.blk_off, bo) and np.array_equal(c.col_map, cm)
    assert np.array_equal(c.payload, pay)
    assert read_container(c.payload, c.col_len, c.blk_off, c.col_map, c.n_chunks) == streams
    dec = [[py_decode(s, chunk // 32)[1] for s in ch] for ch in streams]
    assert assemble(dec, chunk) == x.tobytes()
    rc, got, st = O.brle_trans_decode(c)
    assert rc == 0 and (st == 0).all() and got.tobytes() == x.tobytes()


def test_generator_data_roundtrip():
    """The byte-RLE column generator (runs of mean length 12) round-trips;
    a run of 256 or more would not (the encoder's uint8 counter)."""
    x = O.gen_brle_u8(1 << 22, 0x4D445A02)
    c = O.brle_trans_encode(x, 65536, threads=8)
    rc, got, st = O.brle_trans_decode(c)
    assert rc == 0 and np.array_equal(got, x)


def test_long_runs_lose_bytes_per_wrap():
    lane = b"q" * 700
    s = O.brle_trans_lane_encode(lane)
    assert py_decode(s, 700 - 512)[0] == 0 and O.brle_trans_lane_decode(s, 700)[0] == 2
