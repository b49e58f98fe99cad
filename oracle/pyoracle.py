"""ctypes binding of the CPU oracle (libmdz_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by bench.py's ``cpu_baseline`` leg.  The product
libraries never import this module.  The container encoders and workload
generators it re-exports come from the product's host library
(compression_amd.host); the decoders here are the checkers.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from compression_amd.host import (BrleContainer, DeflateBatch, LzssContainer, LzssPlainContainer,  # noqa: F401
                                  RleV1Container, brle_trans_encode, brle_trans_lane_encode, lzss_plain_encode,
                                  RleV2Container, _ptr, _take, brle_chunk_encode, brle_encode,
                                  default_threads, gen_brle_u8, gen_deflate_batch, gen_lzss_markov,
                                  gen_markov, gen_rlev1_u32, gen_rlev2_i64, host_cores, lzss_encode,
                                  rlev1_encode, rlev1_lane_encode, rlev2_encode, rlev2_lane_encode,
                                  rlev2_plain_chunk_encode, rlev2_plain_encode,
                                  rng_stream)
from compression_amd.host import lzss_lane_encode as _lzss_lane_encode

_HERE = os.path.dirname(os.path.abspath(__file__))
# MDZ_ORACLE_LIB names another build (the sanitizer build of oracle/asan.sh)
_LIB_PATH = os.environ.get("MDZ_ORACLE_LIB") or os.path.join(_HERE, "libmdz_oracle.so")
_lib = None

u8p = C.POINTER(C.c_uint8)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path.  The
    host library it links is built first if missing."""
    from compression_amd import host
    host.lib()
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.mdz_ref_rlev1_lane_decode.restype = C.c_int
        L.mdz_ref_rlev1_lane_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64]
        L.mdz_ref_rlev1_decode.restype = C.c_int
        L.mdz_ref_rlev1_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64,
                                           C.c_int, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]
        L.mdz_ref_brle_chunk_decode.restype = C.c_int
        L.mdz_ref_brle_chunk_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, u64p]
        L.mdz_ref_brle_decode.restype = C.c_int
        L.mdz_ref_brle_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                          C.c_uint64, C.c_void_p, C.c_int]
        L.mdz_ref_rlev2_lane_decode.restype = C.c_int
        L.mdz_ref_rlev2_lane_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64]
        L.mdz_ref_rlev2_plain_chunk_decode.restype = C.c_int
        L.mdz_ref_rlev2_plain_chunk_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, u64p]
        L.mdz_ref_rlev2_plain_decode.restype = C.c_int
        L.mdz_ref_rlev2_plain_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64,
                                                 C.c_void_p, C.c_uint64, C.c_int]
        L.mdz_ref_rlev2_decode.restype = C.c_int
        L.mdz_ref_rlev2_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64,
                                           C.c_int, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]
        L.mdz_ref_lzss_lane_encode_naive.restype = C.c_uint64
        L.mdz_ref_lzss_lane_encode_naive.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_ref_lzss_lane_decode.restype = C.c_int
        L.mdz_ref_lzss_lane_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_ref_lzss_plain_decode.restype = C.c_int
        L.mdz_ref_lzss_plain_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                                C.c_uint64, C.c_void_p, C.c_int]
        L.mdz_ref_brle_trans_lane_decode.restype = C.c_int
        L.mdz_ref_brle_trans_lane_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_ref_brle_trans_decode.restype = C.c_int
        L.mdz_ref_brle_trans_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                                C.c_int]
        L.mdz_ref_lzss_decode.restype = C.c_int
        L.mdz_ref_lzss_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                          C.c_int]
        L.mdz_ref_inflate.restype = C.c_int
        L.mdz_ref_inflate.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, u64p, C.c_int]
        L.mdz_ref_gzip_header_len.restype = C.c_int64
        L.mdz_ref_gzip_header_len.argtypes = [C.c_void_p, C.c_uint64]
        L.mdz_ref_inflate_batch.restype = C.c_int
        L.mdz_ref_inflate_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_int]
        L.mdz_zlib_version.restype = C.c_char_p
        _lib = L
    return _lib


def lzss_lane_encode(lane: bytes, naive: bool = False) -> bytes:
    """The host encoder, or (naive=True) the oracle's literal find_match scan."""
    if not naive:
        return _lzss_lane_encode(lane)
    a = np.frombuffer(bytes(lane), np.uint8) if len(lane) else np.zeros(1, np.uint8)
    cap = len(lane) + len(lane) // 8 + 16
    out = np.zeros(cap, np.uint8)
    n = lib().mdz_ref_lzss_lane_encode_naive(_ptr(a), len(lane), _ptr(out), cap)
    assert n <= cap
    return out[:n].tobytes()


def rlev1_lane_decode(stream: bytes, n_vals: int, dtype) -> tuple[int, np.ndarray]:
    src = np.frombuffer(stream, np.uint8).copy() if stream else np.zeros(1, np.uint8)
    out = np.zeros(n_vals, dtype)
    st = lib().mdz_ref_rlev1_lane_decode(_ptr(src), len(stream), out.dtype.itemsize, _ptr(out), n_vals)
    return st, out

def rlev1_decode(c: RleV1Container, out_bytes: int | None = None, threads: int | None = None):
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    pay = c.payload if c.payload.size else np.zeros(1, np.uint8)
    rc = lib().mdz_ref_rlev1_decode(_ptr(pay), _ptr(c.blk_off), _ptr(c.col_len), c.n_chunks, c.chunk,
                                    c.elem_bytes, c.line, _ptr(out), out_bytes, _ptr(status),
                                    threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]

def rlev2_lane_decode(stream: bytes, n_vals: int, elem_bytes: int = 8) -> tuple[int, np.ndarray]:
    src = np.frombuffer(stream, np.uint8).copy() if stream else np.zeros(1, np.uint8)
    out = np.zeros(max(n_vals, 1), np.uint64)
    st = lib().mdz_ref_rlev2_lane_decode(_ptr(src), len(stream), elem_bytes, _ptr(out), n_vals)
    return st, out[:n_vals]

def rlev2_decode(c: RleV2Container, out_bytes: int | None = None, threads: int | None = None):
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    pay = c.payload if c.payload.size else np.zeros(1, np.uint8)
    rc = lib().mdz_ref_rlev2_decode(_ptr(pay), _ptr(c.blk_off), _ptr(c.col_len), c.n_chunks, c.chunk,
                                    c.elem_bytes, c.read_unit, _ptr(out), out_bytes, _ptr(status),
                                    threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]

def rlev2_plain_chunk_decode(stream: bytes, cap: int) -> tuple[int, np.ndarray]:
    """One plain RLE v2 chunk stream -> (status, int64 values decoded)."""
    src = np.frombuffer(bytes(stream) + b"\0", np.uint8)
    out = np.zeros(max(cap, 1), np.uint64)
    got = C.c_uint64()
    st = lib().mdz_ref_rlev2_plain_chunk_decode(_ptr(src), len(stream), _ptr(out), cap, C.byref(got))
    return st, out[:got.value].view(np.int64)

def rlev2_plain_decode(blob: np.ndarray, chunk: int, out_bytes: int | None = None, threads: int | None = None):
    """A plain RLE v2 container -> (rc, output bytes, per-chunk statuses)."""
    blob = np.ascontiguousarray(blob, np.uint8)
    n = int(blob[:4].view(np.uint32)[0]) - 1 if blob.size >= 4 else 0
    if out_bytes is None:
        out_bytes = int(blob[4 + 8 * (n + 1):12 + 8 * (n + 1)].view(np.uint64)[0])
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(n, 1), np.int32)
    rc = lib().mdz_ref_rlev2_plain_decode(_ptr(blob), blob.size, chunk, _ptr(out), out_bytes, _ptr(status),
                                          max(n, 0), threads or default_threads())
    return rc, out[:out_bytes], status[:max(n, 0)]

def brle_chunk_decode(stream: bytes, cap: int) -> tuple[int, bytes]:
    src = np.frombuffer(stream, np.uint8).copy() if stream else np.zeros(1, np.uint8)
    out = np.zeros(max(cap, 1), np.uint8)
    got = C.c_uint64()
    st = lib().mdz_ref_brle_chunk_decode(_ptr(src), len(stream), _ptr(out), cap, C.byref(got))
    return st, out[:got.value].tobytes()

def brle_decode(c: BrleContainer, out_bytes: int | None = None, threads: int | None = None):
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    data = c.data if c.data.size else np.zeros(1, np.uint8)
    rc = lib().mdz_ref_brle_decode(_ptr(data), _ptr(c.ptr), c.n_chunks, c.chunk, _ptr(out), out_bytes,
                                   _ptr(status), threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]

def lzss_lane_decode(stream: bytes, out_len: int) -> tuple[int, bytes]:
    a = np.frombuffer(bytes(stream) + b"\0", np.uint8)
    out = np.zeros(max(out_len, 1), np.uint8)
    st = lib().mdz_ref_lzss_lane_decode(_ptr(a), len(stream), _ptr(out), out_len)
    return st, out[:out_len].tobytes()

def lzss_decode(c: LzssContainer, out_bytes: int | None = None, threads: int | None = None):
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    pay = c.payload if c.payload.size else np.zeros(1, np.uint8)
    cl = c.col_len if c.col_len.size else np.zeros(1, np.uint64)
    cm = c.col_map if c.col_map.size else np.zeros(1, np.uint8)
    rc = lib().mdz_ref_lzss_decode(_ptr(pay), c.payload.size, _ptr(c.blk_off), _ptr(cl), _ptr(cm),
                                   c.n_chunks, c.chunk, _ptr(out), out_bytes, _ptr(status),
                                   threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]

def brle_trans_lane_decode(stream: bytes, out_len: int) -> tuple[int, bytes]:
    a = np.frombuffer(bytes(stream) + b"\0", np.uint8)
    out = np.zeros(max(out_len, 1), np.uint8)
    st = lib().mdz_ref_brle_trans_lane_decode(_ptr(a), len(stream), _ptr(out), out_len)
    return st, out[:out_len].tobytes()

def brle_trans_decode(c: LzssContainer, out_bytes: int | None = None, threads: int | None = None):
    """Lane-transposed byte RLE container -> (rc, bytes, status[n])."""
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    pay = c.payload if c.payload.size else np.zeros(1, np.uint8)
    cl = c.col_len if c.col_len.size else np.zeros(1, np.uint64)
    cm = c.col_map if c.col_map.size else np.zeros(1, np.uint8)
    rc = lib().mdz_ref_brle_trans_decode(_ptr(pay), c.payload.size, _ptr(c.blk_off), _ptr(cl), _ptr(cm),
                                         c.n_chunks, c.chunk, _ptr(out), out_bytes, _ptr(status),
                                         threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]

def lzss_plain_decode(c, out_bytes: int | None = None, threads: int | None = None):
    """Plain LZSS container (host.LzssPlainContainer) -> (rc, bytes, status[n])."""
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    data = c.data if c.data.size else np.zeros(1, np.uint8)
    lens = c.lens if c.lens.size else np.zeros(1, np.uint64)
    rc = lib().mdz_ref_lzss_plain_decode(_ptr(data), _ptr(lens), c.n_chunks, c.chunk, _ptr(out), out_bytes,
                                         _ptr(status), threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]


def zlib_version() -> str:
    return lib().mdz_zlib_version().decode()

def inflate(stream: bytes, cap: int, zlib_hdr: bool = True, hdr: int | None = None) -> tuple[int, bytes]:
    """hdr: 0 raw, 1 zlib, 2 gzip, | 0x100 trailer check (default: zlib_hdr)."""
    a = np.frombuffer(bytes(stream) + b"\0", np.uint8)
    out = np.zeros(max(cap, 1), np.uint8)
    w = C.c_uint64()
    mode = int(zlib_hdr) if hdr is None else int(hdr)
    st = lib().mdz_ref_inflate(_ptr(a), len(stream), _ptr(out), cap, C.byref(w), mode)
    return st, out[:w.value].tobytes()

def gzip_header_len(stream: bytes) -> int:
    a = np.frombuffer(bytes(stream) + b"\0", np.uint8)
    return lib().mdz_ref_gzip_header_len(_ptr(a), len(stream))

def inflate_batch(b: DeflateBatch, caps=None, zlib_hdr: bool = True, threads: int | None = None):
    caps = b.raw_len.copy() if caps is None else np.ascontiguousarray(caps, np.uint64)
    dst_off = np.zeros(b.count, np.uint64)
    dst_off[1:] = np.cumsum(caps)[:-1]
    out = np.zeros(max(int(caps.sum()), 1), np.uint8)
    written = np.zeros(b.count, np.uint64)
    status = np.zeros(b.count, np.int32)
    rc = lib().mdz_ref_inflate_batch(_ptr(b.src), _ptr(b.src_off), _ptr(b.src_len), _ptr(out),
                                     _ptr(dst_off), _ptr(caps), b.count, int(zlib_hdr), _ptr(written),
                                     _ptr(status), threads or default_threads())
    return rc, out, dst_off, written, status
