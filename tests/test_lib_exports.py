"""The C-ABI library loads and exports every symbol include/mdz.h declares.
No compute calls: this runs without a GPU."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mdz.h")
LIB = os.path.join(ROOT, "compression_amd", "libmdz.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdz_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "mdz_rlev1_decompress" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libmdz.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(mdz_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    import compression_amd._lib as L
    missing = [n for n in declared() if n not in L.SIGNATURES]
    assert not missing, missing
    lib = L.lib()       # dlopen only; no HIP call is made
    for n in declared():
        assert getattr(lib, n) is not None


def test_inflate_framing_checked_on_host():
    """Bad parse_hdr values are refused before any device call."""
    import compression_amd._lib as L
    lib = L.lib()
    for bad in (3, 0x100, 0x200, 0x103, -1):
        assert lib.mdz_inflate_batch(None, None, 0, bad, None) == -1
        assert lib.mdz_cost_features(None, 0, bad, None, None) == -1
    assert lib.mdz_copy_uncompressed(None, -1, None) == -1
    assert lib.mdz_copy_uncompressed(None, 0, None) == 0


def test_encoder_bounds_checked_on_host():
    """The encoders' size queries and argument checks make no device call."""
    import ctypes as C
    import compression_amd._lib as L
    lib = L.lib()
    n, w, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
    # RLE v1: 64 MiB of u32 in 128 KiB chunks, 128-byte lines -> 512 chunks
    assert lib.mdz_rlev1_compress_bound(64 << 20, 131072, 4, 128, C.byref(n), C.byref(w), C.byref(p)) == 0
    assert n.value == 512 and w.value > 0 and p.value >= 512 * 131072
    for bad in ((0, 131072, 4, 128), (1000, 131072, 3, 128), (1000, 131072, 4, 126), (1000, 100, 4, 4)):
        assert lib.mdz_rlev1_compress_bound(*bad, C.byref(n), C.byref(w), C.byref(p)) == -1
    assert lib.mdz_rlev2_compress_bound(1 << 30, 131072, 8, 512, C.byref(n), C.byref(w), C.byref(p)) == 0
    assert n.value == 8192
    assert lib.mdz_rlev2_compress_bound(1000, 131072, 8, 3, C.byref(n), C.byref(w), C.byref(p)) == -1
    assert lib.mdz_brle_compress_bound(300_000, 262144, C.byref(n), C.byref(w), C.byref(p)) == 0
    assert n.value == 2 and p.value >= 300_000
    assert lib.mdz_brle_compress_bound(0, 262144, C.byref(n), C.byref(w), C.byref(p)) == -1
    # LZSS: whole chunks only, a multiple of 128 up to 1 MiB
    assert lib.mdz_lzss_compress_bound(5 * 65536 + 7, 65536, C.byref(n), C.byref(w), C.byref(p)) == 0
    assert n.value == 5
    for chunk in (100, 2 << 20):
        assert lib.mdz_lzss_compress_bound(1 << 22, chunk, C.byref(n), C.byref(w), C.byref(p)) == -1
    assert lib.mdz_lzss_compress_bound(1000, 65536, C.byref(n), C.byref(w), C.byref(p)) == -1
