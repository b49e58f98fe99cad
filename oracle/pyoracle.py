"""ctypes binding of the CPU oracle (libmdz_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by bench.py's ``cpu_baseline`` leg.  The product library
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmdz_oracle.so")
_lib = None

u8p = C.POINTER(C.c_uint8)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
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
        L.mdz_rng_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.mdz_rng_next.restype = C.c_uint64
        L.mdz_rng_next.argtypes = [C.c_void_p]
        L.mdz_host_cores.restype = C.c_int
        L.mdz_free.argtypes = [C.c_void_p]
        L.mdz_ref_rlev1_lane_encode.restype = C.c_uint64
        L.mdz_ref_rlev1_lane_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64]
        L.mdz_ref_rlev1_lane_decode.restype = C.c_int
        L.mdz_ref_rlev1_lane_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64]
        L.mdz_ref_rlev1_encode.restype = C.c_int
        L.mdz_ref_rlev1_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_uint32,
                                           C.POINTER(u8p), u64p, C.POINTER(u64p), C.POINTER(u64p),
                                           u64p, C.c_int]
        L.mdz_ref_rlev1_decode.restype = C.c_int
        L.mdz_ref_rlev1_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64,
                                           C.c_int, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]
        L.mdz_ref_brle_chunk_encode.restype = C.c_uint64
        L.mdz_ref_brle_chunk_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.mdz_ref_brle_chunk_decode.restype = C.c_int
        L.mdz_ref_brle_chunk_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, u64p]
        L.mdz_ref_brle_encode.restype = C.c_int
        L.mdz_ref_brle_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(u8p),
                                          C.POINTER(u64p), u64p, C.c_int]
        L.mdz_ref_brle_decode.restype = C.c_int
        L.mdz_ref_brle_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                          C.c_uint64, C.c_void_p, C.c_int]
        L.mdz_ref_rlev2_lane_encode.restype = C.c_uint64
        L.mdz_ref_rlev2_lane_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_ref_rlev2_lane_decode.restype = C.c_int
        L.mdz_ref_rlev2_lane_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64]
        L.mdz_ref_rlev2_encode.restype = C.c_int
        L.mdz_ref_rlev2_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_uint32,
                                           C.POINTER(u8p), u64p, C.POINTER(u64p), C.POINTER(u64p),
                                           u64p, C.c_int]
        L.mdz_ref_rlev2_decode.restype = C.c_int
        L.mdz_ref_rlev2_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64,
                                           C.c_int, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]
        L.mdz_gen_rlev2_i64.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.mdz_gen_rlev1_u32.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.mdz_gen_brle_u8.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        for f in (L.mdz_ref_lzss_lane_encode, L.mdz_ref_lzss_lane_encode_naive):
            f.restype = C.c_uint64
            f.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_ref_lzss_lane_decode.restype = C.c_int
        L.mdz_ref_lzss_lane_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_ref_lzss_encode.restype = C.c_int
        L.mdz_ref_lzss_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(u8p), u64p,
                                          C.POINTER(u64p), C.POINTER(u64p), C.POINTER(u8p), u64p,
                                          C.c_int]
        L.mdz_ref_lzss_decode.restype = C.c_int
        L.mdz_ref_lzss_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                          C.c_int]
        L.mdz_gen_lzss_markov.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.mdz_ref_inflate.restype = C.c_int
        L.mdz_ref_inflate.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, u64p, C.c_int]
        L.mdz_ref_inflate_batch.restype = C.c_int
        L.mdz_ref_inflate_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_int]
        L.mdz_gen_deflate_batch.restype = C.c_int
        L.mdz_gen_deflate_batch.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                            C.POINTER(u8p), C.POINTER(u64p), C.POINTER(u64p),
                                            C.POINTER(u8p), C.POINTER(u64p), C.POINTER(u64p),
                                            C.POINTER(u8p), C.c_int]
        L.mdz_zlib_version.restype = C.c_char_p
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def host_cores() -> int:
    return int(lib().mdz_host_cores())


def default_threads() -> int:
    return max(1, min(host_cores(), 16))


# ------------------------------------------------------------------- PRNG --
def rng_stream(seed: int, n: int) -> np.ndarray:
    state = (C.c_uint64 * 4)()
    lib().mdz_rng_seed(state, seed)
    return np.array([lib().mdz_rng_next(state) for _ in range(n)], dtype=np.uint64)


# ----------------------------------------------------------------- RLE v1 --
_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def rlev1_lane_encode(vals: np.ndarray) -> bytes:
    vals = np.ascontiguousarray(vals)
    eb = vals.dtype.itemsize
    cap = len(vals) * 11 + 16
    out = np.zeros(cap, np.uint8)
    n = lib().mdz_ref_rlev1_lane_encode(_ptr(vals), len(vals), eb, _ptr(out), cap)
    return out[:n].tobytes()


def rlev1_lane_decode(stream: bytes, n_vals: int, dtype) -> tuple[int, np.ndarray]:
    src = np.frombuffer(stream, np.uint8).copy() if stream else np.zeros(1, np.uint8)
    out = np.zeros(n_vals, dtype)
    st = lib().mdz_ref_rlev1_lane_decode(_ptr(src), len(stream), out.dtype.itemsize, _ptr(out), n_vals)
    return st, out


class RleV1Container:
    """Lane-transposed RLE v1 container: payload + col_len[32n] + blk_off[n+1]."""

    def __init__(self, payload, col_len, blk_off, n_chunks, chunk, elem_bytes, line, raw_bytes):
        self.payload, self.col_len, self.blk_off = payload, col_len, blk_off
        self.n_chunks, self.chunk, self.elem_bytes, self.line = n_chunks, chunk, elem_bytes, line
        self.raw_bytes = raw_bytes


def _take(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))), shape=(n,)).copy()
    lib().mdz_free(C.cast(ptr, C.c_void_p))
    return arr


def rlev1_encode(data: np.ndarray, chunk: int, line: int, threads: int | None = None) -> RleV1Container:
    data = np.ascontiguousarray(data)
    raw = data.view(np.uint8).reshape(-1)
    pay, col, off = u8p(), u64p(), u64p()
    pb, nc = C.c_uint64(), C.c_uint64()
    rc = lib().mdz_ref_rlev1_encode(_ptr(raw), raw.size, chunk, data.dtype.itemsize, line,
                                    C.byref(pay), C.byref(pb), C.byref(col), C.byref(off),
                                    C.byref(nc), threads or default_threads())
    if rc != 0:
        raise ValueError(f"rlev1 encode rejected parameters (status {rc})")
    n = nc.value
    return RleV1Container(_take(pay, pb.value, np.uint8), _take(col, 32 * n, np.uint64),
                          _take(off, n + 1, np.uint64), n, chunk, data.dtype.itemsize, line, raw.size)


def rlev1_decode(c: RleV1Container, out_bytes: int | None = None, threads: int | None = None):
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    pay = c.payload if c.payload.size else np.zeros(1, np.uint8)
    rc = lib().mdz_ref_rlev1_decode(_ptr(pay), _ptr(c.blk_off), _ptr(c.col_len), c.n_chunks, c.chunk,
                                    c.elem_bytes, c.line, _ptr(out), out_bytes, _ptr(status),
                                    threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]


# ----------------------------------------------------------------- RLE v2 --
def rlev2_lane_encode(vals) -> bytes:
    v = np.ascontiguousarray(np.asarray(vals, dtype=np.uint64))
    cap = len(v) * 10 + 64
    out = np.zeros(cap, np.uint8)
    n = lib().mdz_ref_rlev2_lane_encode(_ptr(v) if v.size else None, len(v), _ptr(out), cap)
    return out[:n].tobytes()


def rlev2_lane_decode(stream: bytes, n_vals: int, elem_bytes: int = 8) -> tuple[int, np.ndarray]:
    src = np.frombuffer(stream, np.uint8).copy() if stream else np.zeros(1, np.uint8)
    out = np.zeros(max(n_vals, 1), np.uint64)
    st = lib().mdz_ref_rlev2_lane_decode(_ptr(src), len(stream), elem_bytes, _ptr(out), n_vals)
    return st, out[:n_vals]


class RleV2Container:
    """Lane-transposed RLE v2: payload, col_len[32n], blk_off[n+1] (blk_off[n] = raw size)."""

    def __init__(self, payload, col_len, blk_off, n_chunks, chunk, elem_bytes, read_unit, raw_bytes,
                 payload_bytes):
        self.payload, self.col_len, self.blk_off = payload, col_len, blk_off
        self.n_chunks, self.chunk, self.elem_bytes, self.read_unit = n_chunks, chunk, elem_bytes, read_unit
        self.raw_bytes, self.payload_bytes = raw_bytes, payload_bytes


def rlev2_encode(data: np.ndarray, chunk: int, read_unit: int, threads: int | None = None) -> RleV2Container:
    data = np.ascontiguousarray(data)
    raw = data.view(np.uint8).reshape(-1)
    pay, col, off = u8p(), u64p(), u64p()
    pb, nc = C.c_uint64(), C.c_uint64()
    rc = lib().mdz_ref_rlev2_encode(_ptr(raw), raw.size, chunk, data.dtype.itemsize, read_unit,
                                    C.byref(pay), C.byref(pb), C.byref(col), C.byref(off), C.byref(nc),
                                    threads or default_threads())
    if rc != 0:
        raise ValueError(f"rlev2 encode rejected parameters (status {rc})")
    n = nc.value
    return RleV2Container(_take(pay, pb.value, np.uint8), _take(col, 32 * n, np.uint64),
                          _take(off, n + 1, np.uint64), n, chunk, data.dtype.itemsize, read_unit,
                          raw.size, pb.value)


def rlev2_decode(c: RleV2Container, out_bytes: int | None = None, threads: int | None = None):
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    pay = c.payload if c.payload.size else np.zeros(1, np.uint8)
    rc = lib().mdz_ref_rlev2_decode(_ptr(pay), _ptr(c.blk_off), _ptr(c.col_len), c.n_chunks, c.chunk,
                                    c.elem_bytes, c.read_unit, _ptr(out), out_bytes, _ptr(status),
                                    threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]


def gen_rlev2_i64(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().mdz_gen_rlev2_i64(_ptr(out), n, seed)
    return out


# --------------------------------------------------------------- byte RLE --
def brle_chunk_encode(data: bytes) -> bytes:
    src = np.frombuffer(data, np.uint8).copy() if data else np.zeros(1, np.uint8)
    out = np.zeros(len(data) + len(data) // 128 + 2, np.uint8)
    n = lib().mdz_ref_brle_chunk_encode(_ptr(src), len(data), _ptr(out))
    return out[:n].tobytes()


def brle_chunk_decode(stream: bytes, cap: int) -> tuple[int, bytes]:
    src = np.frombuffer(stream, np.uint8).copy() if stream else np.zeros(1, np.uint8)
    out = np.zeros(max(cap, 1), np.uint8)
    got = C.c_uint64()
    st = lib().mdz_ref_brle_chunk_decode(_ptr(src), len(stream), _ptr(out), cap, C.byref(got))
    return st, out[:got.value].tobytes()


class BrleContainer:
    def __init__(self, data, ptr, n_chunks, chunk, raw_bytes):
        self.data, self.ptr, self.n_chunks, self.chunk, self.raw_bytes = data, ptr, n_chunks, chunk, raw_bytes

    def to_bytes(self) -> bytes:
        """Reference file layout [u32 n+1][u64 ptr[n+1]][u64 out_bytes][data] (brle.h:278-292)."""
        return (np.uint32(self.n_chunks + 1).tobytes() + self.ptr.astype('<u8').tobytes()
                + np.uint64(self.raw_bytes).tobytes() + self.data.tobytes())


def brle_encode(data: np.ndarray, chunk: int, threads: int | None = None) -> BrleContainer:
    raw = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    d, p, nc = u8p(), u64p(), C.c_uint64()
    rc = lib().mdz_ref_brle_encode(_ptr(raw), raw.size, chunk, C.byref(d), C.byref(p), C.byref(nc),
                                   threads or default_threads())
    if rc != 0:
        raise ValueError(f"brle encode rejected parameters (status {rc})")
    n = nc.value
    ptr = _take(p, n + 1, np.uint64)
    return BrleContainer(_take(d, int(ptr[-1]), np.uint8), ptr, n, chunk, raw.size)


def brle_decode(c: BrleContainer, out_bytes: int | None = None, threads: int | None = None):
    out_bytes = c.raw_bytes if out_bytes is None else out_bytes
    out = np.zeros(max(out_bytes, 1), np.uint8)
    status = np.zeros(max(c.n_chunks, 1), np.int32)
    data = c.data if c.data.size else np.zeros(1, np.uint8)
    rc = lib().mdz_ref_brle_decode(_ptr(data), _ptr(c.ptr), c.n_chunks, c.chunk, _ptr(out), out_bytes,
                                   _ptr(status), threads or default_threads())
    return rc, out[:out_bytes], status[:c.n_chunks]


# ------------------------------------------------------------------- LZSS --
def lzss_lane_encode(lane: bytes, naive: bool = False) -> bytes:
    a = np.frombuffer(bytes(lane), np.uint8) if len(lane) else np.zeros(1, np.uint8)
    cap = len(lane) + len(lane) // 8 + 16
    out = np.zeros(cap, np.uint8)
    f = lib().mdz_ref_lzss_lane_encode_naive if naive else lib().mdz_ref_lzss_lane_encode
    n = f(_ptr(a), len(lane), _ptr(out), cap)
    assert n <= cap
    return out[:n].tobytes()


def lzss_lane_decode(stream: bytes, out_len: int) -> tuple[int, bytes]:
    a = np.frombuffer(bytes(stream) + b"\0", np.uint8)
    out = np.zeros(max(out_len, 1), np.uint8)
    st = lib().mdz_ref_lzss_lane_decode(_ptr(a), len(stream), _ptr(out), out_len)
    return st, out[:out_len].tobytes()


class LzssContainer:
    """payload + blk_off[n+1] + col_len[32n] (sorted) + col_map[32n]."""

    def __init__(self, payload, col_len, blk_off, col_map, n_chunks, chunk):
        self.payload, self.col_len, self.blk_off, self.col_map = payload, col_len, blk_off, col_map
        self.n_chunks, self.chunk = n_chunks, chunk
        self.raw_bytes = n_chunks * chunk

    def meta_bytes(self) -> bytes:
        """.meta side file: [u64 4][u64 lens[4]][u64 chunk][blk_off][col_len][col_map]
        (src/lzss/lzss_gpu.cu:81-103, include/lzss/lzss_gpu2.h:2090-2108)."""
        blobs = [np.uint64(self.chunk).tobytes(), self.blk_off.astype("<u8").tobytes(),
                 self.col_len.astype("<u8").tobytes(), self.col_map.astype(np.uint8).tobytes()]
        return (np.uint64(len(blobs)).tobytes() + np.array([len(b) for b in blobs], "<u8").tobytes()
                + b"".join(blobs))


def lzss_encode(data: np.ndarray, chunk: int, threads: int | None = None) -> LzssContainer:
    raw = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    if raw.size == 0:
        raw = np.zeros(1, np.uint8)[:0]
    src = raw if raw.size else np.zeros(1, np.uint8)
    pay, pb, cl, bo, cm, nc = u8p(), C.c_uint64(), u64p(), u64p(), u8p(), C.c_uint64()
    rc = lib().mdz_ref_lzss_encode(_ptr(src), raw.size, chunk, C.byref(pay), C.byref(pb), C.byref(cl),
                                   C.byref(bo), C.byref(cm), C.byref(nc), threads or default_threads())
    if rc != 0:
        raise ValueError(f"lzss encode rejected parameters (status {rc})")
    n = nc.value
    return LzssContainer(_take(pay, pb.value, np.uint8), _take(cl, 32 * n, np.uint64),
                         _take(bo, n + 1, np.uint64), _take(cm, 32 * n, np.uint8), n, chunk)


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


def gen_lzss_markov(n: int, seed: int) -> np.ndarray:
    out = np.empty(max(n, 1), np.uint8)
    lib().mdz_gen_lzss_markov(_ptr(out), n, seed)
    return out[:n]


# ---------------------------------------------------------------- Deflate --
def zlib_version() -> str:
    return lib().mdz_zlib_version().decode()


def inflate(stream: bytes, cap: int, zlib_hdr: bool = True) -> tuple[int, bytes]:
    a = np.frombuffer(bytes(stream) + b"\0", np.uint8)
    out = np.zeros(max(cap, 1), np.uint8)
    w = C.c_uint64()
    st = lib().mdz_ref_inflate(_ptr(a), len(stream), _ptr(out), cap, C.byref(w), int(zlib_hdr))
    return st, out[:w.value].tobytes()


class DeflateBatch:
    """Concatenated zlib streams (16-byte aligned starts) and their raw data."""

    def __init__(self, src, src_off, src_len, raw, raw_off, raw_len, mode):
        self.src, self.src_off, self.src_len = src, src_off, src_len
        self.raw, self.raw_off, self.raw_len, self.mode = raw, raw_off, raw_len, mode
        self.count = len(src_len)

    def stream(self, i) -> bytes:
        o = int(self.src_off[i])
        return self.src[o:o + int(self.src_len[i])].tobytes()

    def raw_stream(self, i) -> bytes:
        o = int(self.raw_off[i])
        return self.raw[o:o + int(self.raw_len[i])].tobytes()


def gen_deflate_batch(count: int, seed: int, min_size: int = 4096, max_size: int = 4 << 20,
                      threads: int | None = None) -> DeflateBatch:
    s, so, sl, r, ro, rl, md = u8p(), u64p(), u64p(), u8p(), u64p(), u64p(), u8p()
    rc = lib().mdz_gen_deflate_batch(count, seed, min_size, max_size, C.byref(s), C.byref(so),
                                     C.byref(sl), C.byref(r), C.byref(ro), C.byref(rl), C.byref(md),
                                     threads or default_threads())
    if rc != 0:
        raise ValueError("bad deflate batch parameters")
    so_a = _take(so, count + 1, np.uint64)
    ro_a = _take(ro, count + 1, np.uint64)
    return DeflateBatch(_take(s, int(so_a[-1]) + 16, np.uint8), so_a[:-1].copy(), _take(sl, count, np.uint64),
                        _take(r, int(ro_a[-1]) + 1, np.uint8)[:int(ro_a[-1])], ro_a[:-1].copy(),
                        _take(rl, count, np.uint64), _take(md, count, np.uint8))


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


# ------------------------------------------------------------- generators --
def gen_rlev1_u32(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    lib().mdz_gen_rlev1_u32(_ptr(out), n, seed)
    return out


def gen_brle_u8(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.uint8)
    lib().mdz_gen_brle_u8(_ptr(out), n, seed)
    return out
