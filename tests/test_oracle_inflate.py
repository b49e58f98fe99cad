"""CPU oracle: zlib inflate wrapper (status mapping, header check), the
config-4 batch generator, committed fixtures."""
import json
import os
import zlib

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.deflate_util import crafted_faults

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "deflate_fixtures.json")))


def test_kats_c10_c12():
    # hand-checked against RFC 1950/1951; the compressor bytes
    # below are zlib 1.2.11's
    kat = {"C10_stored_abc": "78010103 00fcff616263024d0127",
           "C11_fixed_abcabcabc": "78014b4c4a4e042300113d0373",
           "C12_empty_level6": "789c030000000001"}
    for v in FIX["valid"]:
        if v["name"] in kat:
            assert v["stream"] == kat[v["name"]].replace(" ", "")
        st, out = O.inflate(bytes.fromhex(v["stream"]), len(v["raw"]) // 2)
        assert st == 0 and out.hex() == v["raw"]
    dyn = next(v for v in FIX["valid"] if v["name"] == "dynamic_range64x64")
    s = bytes.fromhex(dyn["stream"])
    assert len(s) == 104 and (s[2] >> 1) & 3 == 2          # BTYPE = 2


@pytest.mark.parametrize("fault", FIX["faults"], ids=lambda f: f["name"])
def test_fault_status(fault):
    st, _ = O.inflate(bytes.fromhex(fault["stream"]), fault["cap"])
    assert st == fault["status"]


def test_crafted_builders_match_fixtures():
    got = {n: (s.hex(), c, e) for n, s, c, e in crafted_faults()}
    for f in FIX["faults"]:
        assert got[f["name"]] == (f["stream"], f["cap"], f["status"])


def test_oracle_vs_python_zlib():
    rng = np.random.default_rng(0)
    for _ in range(40):
        n = int(rng.integers(0, 200_000))
        k = int(rng.integers(1, 256))
        data = rng.integers(0, k, n).astype(np.uint8).tobytes()
        if rng.random() < 0.5:
            data = data[: n // 3] * 3
        level = int(rng.choice([0, 1, 6, 9]))
        s = zlib.compress(data, level)
        st, out = O.inflate(s, len(data))
        assert st == 0 and out == data
        raw = zlib.compressobj(level, zlib.DEFLATED, -15)
        rs = raw.compress(data) + raw.flush()
        assert O.inflate(rs, len(data), zlib_hdr=False) == (0, data)
        if len(data) > 10:
            assert O.inflate(s, len(data) - 1)[0] == 1                # output full early
            assert O.inflate(s[: len(s) // 2], len(data))[0] == 2     # truncated


def test_generator_batch():
    b = O.gen_deflate_batch(200, 0x4D445A04, 4096, 1 << 20)
    assert b.count == 200
    assert (b.raw_len >= 4096).all() and (b.raw_len <= 1 << 20).all()
    counts = np.bincount(b.mode, minlength=7)
    assert counts[1] == 20 and counts[0] == 20 and counts[6] == 160
    for i in range(b.count):
        s = b.stream(i)
        btype = (s[2] >> 1) & 3
        assert btype == {0: 0, 1: 1, 6: 2}[int(b.mode[i])] or (b.mode[i] == 6 and btype == 1)
        assert int(b.src_off[i]) % 16 == 0
    for i in range(0, b.count, 17):
        assert zlib.decompress(b.stream(i)) == b.raw_stream(i)
    rc, out, dst_off, written, st = O.inflate_batch(b)
    assert rc == 0 and (st == 0).all() and (written == b.raw_len).all()
    assert np.array_equal(out[: b.raw.size], b.raw)
    b2 = O.gen_deflate_batch(200, 0x4D445A04, 4096, 1 << 20)
    assert np.array_equal(b.src, b2.src) and np.array_equal(b.raw_len, b2.raw_len)
