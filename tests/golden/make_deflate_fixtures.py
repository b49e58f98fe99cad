"""Writes tests/golden/deflate_fixtures.json: small zlib streams with their
inflated bytes (as produced by this machine's zlib, version recorded) and the
crafted fault streams with zlib's status for each.  Run from the repo root:

    python tests/golden/make_deflate_fixtures.py
"""
import json
import os
import sys
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O                     # noqa: E402
from tests.deflate_util import crafted_faults        # noqa: E402


def fixed(data: bytes, level=1) -> bytes:
    c = zlib.compressobj(level, zlib.DEFLATED, 15, 8, zlib.Z_FIXED)
    return c.compress(data) + c.flush()


def main():
    valid = [
        ("C10_stored_abc", zlib.compress(b"abc", 0), b"abc"),
        ("C11_fixed_abcabcabc", fixed(b"abcabcabc"), b"abcabcabc"),
        ("C12_empty_level6", zlib.compress(b"", 6), b""),
        ("dynamic_range64x64", zlib.compress(bytes(range(64)) * 64, 6), bytes(range(64)) * 64),
        ("stored_block", zlib.compress(bytes(range(256)) * 32, 0), bytes(range(256)) * 32),
        ("overlap_dist1", zlib.compress(b"\x07" * 5000, 9), b"\x07" * 5000),
    ]
    out = {"zlib_version": zlib.ZLIB_VERSION, "valid": [], "faults": []}
    for name, s, raw in valid:
        st, dec = O.inflate(s, len(raw))
        assert st == 0 and dec == raw
        out["valid"].append({"name": name, "stream": s.hex(), "raw": raw.hex()})
    for name, s, cap, exp in crafted_faults():
        st, _ = O.inflate(s, cap)
        assert st == exp, name
        out["faults"].append({"name": name, "stream": s.hex(), "cap": cap, "status": st})
    with open(os.path.join(HERE, "deflate_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
