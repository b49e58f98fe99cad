"""CPU oracle: gzip framing by cuDF's parse_gzip_header rules
(include/cudf_deflate.cuh:1179-1217) and the trailer checks (Adler-32,
CRC-32, ISIZE).  Members are built field by field, so every header length is
known by construction, and zlib itself (wbits 31) accepts each one; trailer
values come from zlib's adler32 / crc32."""
import struct
import zlib

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.deflate_util import GZIP_FIELD_CASES, gzip_member

VERIFY = 0x100


@pytest.mark.parametrize("fields", GZIP_FIELD_CASES, ids=lambda f: "+".join(sorted(f)) or "plain")
def test_header_lengths_by_construction(fields):
    data = np.random.default_rng(len(str(fields))).integers(0, 9, 5000).astype(np.uint8).tobytes()
    s, hlen = gzip_member(data, **fields)
    assert zlib.decompress(s, 31) == data              # a valid member, header CRC included
    assert O.gzip_header_len(s) == hlen
    assert O.inflate(s, len(data), hdr=2) == (0, data)
    assert O.inflate(s, len(data), hdr=2 | VERIFY) == (0, data)


def test_rejected_headers():
    s, _ = gzip_member(b"")                             # 10 + 2 + 8 = 20 bytes
    assert len(s) == 20 and O.gzip_header_len(s) == 10
    assert O.gzip_header_len(s[:18]) == -1              # header + footer leave no Deflate byte
    assert O.gzip_header_len(s[:17]) == -1              # under 18 bytes
    for k, v in ((0, 0x1e), (1, 0x8c), (2, 0x07)):      # ID1, ID2, CM
        b = bytearray(s); b[k] = v
        assert O.gzip_header_len(bytes(b)) == -1
    n, _ = gzip_member(b"x" * 100, name=b"abc")
    cut = n[:10] + b"y" * 30                            # FNAME never terminated
    assert O.gzip_header_len(cut) == -1
    e, _ = gzip_member(b"x" * 100, extra=b"q" * 8)
    b = bytearray(e); b[10:12] = struct.pack("<H", len(e))   # XLEN past the end
    assert O.gzip_header_len(bytes(b)) == -1
    st, out = O.inflate(bytes(b), 100, hdr=2)
    assert st == -1 and out == b""                      # cuDF's status for a rejected header


def test_fextra_counts_xlen_field():
    # RFC 1952: XLEN (2 bytes) then XLEN bytes; cuDF skips only XLEN bytes
    s, hlen = gzip_member(b"payload" * 50, extra=b"\x01" * 6)
    assert hlen == 10 + 2 + 6 and O.gzip_header_len(s) == hlen


def test_footer_is_not_deflate_input():
    data = b"0123456789" * 300
    s, _ = gzip_member(data, level=0)                   # stored: input overrun is visible
    assert O.inflate(s, len(data), hdr=2) == (0, data)
    assert O.inflate(s[:-3] + s[-1:], len(data), hdr=2)[0] == 2    # footer bytes eaten by data


@pytest.mark.parametrize("level", [0, 1, 6, 9])
def test_trailer_checks(level):
    rng = np.random.default_rng(level)
    data = (rng.integers(0, 4, 70_000).astype(np.uint8).tobytes() + bytes(range(256)) * 30)
    z = zlib.compress(data, level)
    assert O.inflate(z, len(data), hdr=1 | VERIFY) == (0, data)
    for k in range(1, 5):                               # each Adler-32 byte
        b = bytearray(z); b[-k] ^= 0x20
        assert O.inflate(bytes(b), len(data), hdr=1 | VERIFY)[0] == -13
        assert O.inflate(bytes(b), len(data), hdr=1)[0] == 0      # unchecked framing
    assert O.inflate(z[:-1], len(data), hdr=1 | VERIFY)[0] == 2  # trailer cut off
    assert O.inflate(z + b"\0" * 5, len(data), hdr=1 | VERIFY)[0] == 0   # padding after it
    g, _ = gzip_member(data, level)
    for k in range(1, 9):                               # CRC-32 bytes, then ISIZE bytes
        b = bytearray(g); b[-k] ^= 0x01
        assert O.inflate(bytes(b), len(data), hdr=2 | VERIFY)[0] == -13
    if level == 0:
        b = bytearray(z); b[100] ^= 0xFF                # stored data byte: output differs
        assert O.inflate(bytes(b), len(data), hdr=1)[0] == 0
        assert O.inflate(bytes(b), len(data), hdr=1 | VERIFY)[0] == -13


def test_checksum_closed_forms():
    # Adler-32 of n zero bytes is (n mod 65521) << 16 | 1; CRC-32 check value
    for n in (0, 1, 1023, 1024, 65521, 70000):
        z = zlib.compress(bytes(n))
        assert struct.unpack(">I", z[-4:])[0] == ((n % 65521) << 16) | 1
        assert O.inflate(z, n, hdr=1 | VERIFY)[0] == 0
    assert zlib.crc32(b"123456789") == 0xCBF43926
    g, _ = gzip_member(b"123456789")
    assert struct.unpack("<I", g[-8:-4])[0] == 0xCBF43926
    assert O.inflate(g, 9, hdr=2 | VERIFY) == (0, b"123456789")
