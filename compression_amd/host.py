"""ctypes binding of the mdz host library (libmdz_host.so): the container
encoders (the CPU compress paths that produce the reference's formats) and
the seeded synthetic workloads.  No GPU is involved.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MDZ_HOST_LIB names another build (the sanitizer build of oracle/asan.sh)
_LIB_PATH = os.environ.get("MDZ_HOST_LIB") or os.path.join(_HERE, "libmdz_host.so")
_lib = None

u8p = C.POINTER(C.c_uint8)
u64p = C.POINTER(C.c_uint64)


def lib():
    """The host library; built on first use if missing (gcc, zlib)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            from .build import build_host
            build_host()
        L = C.CDLL(_LIB_PATH)
        L.mdz_rng_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.mdz_rng_next.restype = C.c_uint64
        L.mdz_rng_next.argtypes = [C.c_void_p]
        L.mdz_host_cores.restype = C.c_int
        L.mdz_free.argtypes = [C.c_void_p]
        L.mdz_rlev1_lane_encode.restype = C.c_uint64
        L.mdz_rlev1_lane_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64]
        L.mdz_rlev1_encode.restype = C.c_int
        L.mdz_rlev1_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_uint32,
                                       C.POINTER(u8p), u64p, C.POINTER(u64p), C.POINTER(u64p), u64p, C.c_int]
        L.mdz_rlev2_lane_encode.restype = C.c_uint64
        L.mdz_rlev2_lane_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_rlev2_plain_chunk_encode.restype = C.c_uint64
        L.mdz_rlev2_plain_chunk_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_rlev2_plain_encode.restype = C.c_int
        L.mdz_rlev2_plain_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(u8p),
                                             C.POINTER(C.c_uint64), C.c_int]
        L.mdz_rlev2_encode.restype = C.c_int
        L.mdz_rlev2_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_uint32,
                                       C.POINTER(u8p), u64p, C.POINTER(u64p), C.POINTER(u64p), u64p, C.c_int]
        L.mdz_brle_chunk_encode.restype = C.c_uint64
        L.mdz_brle_chunk_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.mdz_brle_encode.restype = C.c_int
        L.mdz_brle_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(u8p),
                                      C.POINTER(u64p), u64p, C.c_int]
        L.mdz_lzss_lane_encode.restype = C.c_uint64
        L.mdz_lzss_lane_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_lzss_encode.restype = C.c_int
        L.mdz_lzss_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(u8p), u64p,
                                      C.POINTER(u64p), C.POINTER(u64p), C.POINTER(u8p), u64p, C.c_int]
        L.mdz_brle_trans_lane_encode.restype = C.c_uint64
        L.mdz_brle_trans_lane_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        L.mdz_brle_trans_encode.restype = C.c_int
        L.mdz_brle_trans_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(u8p), u64p,
                                            C.POINTER(u64p), C.POINTER(u64p), C.POINTER(u8p), u64p, C.c_int]
        L.mdz_lzss_plain_encode.restype = C.c_int
        L.mdz_lzss_plain_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(u8p),
                                            C.POINTER(u64p), u64p, C.c_int]
        L.mdz_gen_rlev1_u32.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.mdz_gen_rlev2_i64.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.mdz_gen_brle_u8.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.mdz_gen_lzss_markov.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.mdz_markov_table.restype = C.c_void_p
        L.mdz_markov_table.argtypes = [C.c_uint64]
        L.mdz_markov_gen.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_double]
        L.mdz_gen_deflate_batch.restype = C.c_int
        L.mdz_gen_deflate_batch.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                            C.POINTER(u8p), C.POINTER(u64p), C.POINTER(u64p),
                                            C.POINTER(u8p), C.POINTER(u64p), C.POINTER(u64p),
                                            C.POINTER(u8p), C.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)

def host_cores() -> int:
    return int(lib().mdz_host_cores())

def default_threads() -> int:
    return max(1, min(host_cores(), 16))

def rng_stream(seed: int, n: int) -> np.ndarray:
    state = (C.c_uint64 * 4)()
    lib().mdz_rng_seed(state, seed)
    return np.array([lib().mdz_rng_next(state) for _ in range(n)], dtype=np.uint64)

def rlev1_lane_encode(vals: np.ndarray) -> bytes:
    vals = np.ascontiguousarray(vals)
    eb = vals.dtype.itemsize
    cap = len(vals) * 11 + 16
    out = np.zeros(cap, np.uint8)
    n = lib().mdz_rlev1_lane_encode(_ptr(vals), len(vals), eb, _ptr(out), cap)
    return out[:n].tobytes()

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
    rc = lib().mdz_rlev1_encode(_ptr(raw), raw.size, chunk, data.dtype.itemsize, line,
                                    C.byref(pay), C.byref(pb), C.byref(col), C.byref(off),
                                    C.byref(nc), threads or default_threads())
    if rc != 0:
        raise ValueError(f"rlev1 encode rejected parameters (status {rc})")
    n = nc.value
    return RleV1Container(_take(pay, pb.value, np.uint8), _take(col, 32 * n, np.uint64),
                          _take(off, n + 1, np.uint64), n, chunk, data.dtype.itemsize, line, raw.size)

def rlev2_lane_encode(vals) -> bytes:
    v = np.ascontiguousarray(np.asarray(vals, dtype=np.uint64))
    cap = len(v) * 10 + 64
    out = np.zeros(cap, np.uint8)
    n = lib().mdz_rlev2_lane_encode(_ptr(v) if v.size else None, len(v), _ptr(out), cap)
    return out[:n].tobytes()

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
    rc = lib().mdz_rlev2_encode(_ptr(raw), raw.size, chunk, data.dtype.itemsize, read_unit,
                                    C.byref(pay), C.byref(pb), C.byref(col), C.byref(off), C.byref(nc),
                                    threads or default_threads())
    if rc != 0:
        raise ValueError(f"rlev2 encode rejected parameters (status {rc})")
    n = nc.value
    return RleV2Container(_take(pay, pb.value, np.uint8), _take(col, 32 * n, np.uint64),
                          _take(off, n + 1, np.uint64), n, chunk, data.dtype.itemsize, read_unit,
                          raw.size, pb.value)

def rlev2_plain_chunk_encode(vals) -> bytes:
    """One chunk's plain RLE v2 run stream (int64 values in [0, 2^63))."""
    v = np.ascontiguousarray(np.asarray(vals, dtype=np.int64))
    cap = len(v) * 10 + 64
    out = np.zeros(cap, np.uint8)
    n = lib().mdz_rlev2_plain_chunk_encode(_ptr(v) if v.size else None, len(v), _ptr(out), cap)
    return out[:n].tobytes()

class RleV2PlainContainer:
    """Plain RLE v2 container bytes [u32 n+1][u64 ptr[n+1]][u64 raw_bytes][streams],
    with its parts parsed out: ptr (n+1 offsets into data), data (the
    streams), raw_bytes."""

    def __init__(self, blob: np.ndarray, chunk: int):
        self.blob = np.ascontiguousarray(blob, np.uint8)
        self.chunk = chunk
        np_ = int(self.blob[:4].view(np.uint32)[0])
        self.n_chunks = np_ - 1
        self.ptr = self.blob[4:4 + 8 * np_].view(np.uint64).copy()
        self.raw_bytes = int(self.blob[4 + 8 * np_:12 + 8 * np_].view(np.uint64)[0])
        self.data_off = 12 + 8 * np_
        self.data = self.blob[self.data_off:]

def rlev2_plain_encode(data: np.ndarray, chunk: int = 32768, threads: int | None = None) -> RleV2PlainContainer:
    raw = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    out, ob = u8p(), C.c_uint64()
    rc = lib().mdz_rlev2_plain_encode(_ptr(raw), raw.size, chunk, C.byref(out), C.byref(ob),
                                      threads or default_threads())
    if rc != 0:
        raise ValueError(f"plain rlev2 encode rejected its input (status {rc})")
    return RleV2PlainContainer(_take(out, ob.value, np.uint8), chunk)

def gen_rlev2_i64(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().mdz_gen_rlev2_i64(_ptr(out), n, seed)
    return out

def brle_chunk_encode(data: bytes) -> bytes:
    src = np.frombuffer(data, np.uint8).copy() if data else np.zeros(1, np.uint8)
    out = np.zeros(len(data) + len(data) // 128 + 2, np.uint8)
    n = lib().mdz_brle_chunk_encode(_ptr(src), len(data), _ptr(out))
    return out[:n].tobytes()

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
    rc = lib().mdz_brle_encode(_ptr(raw), raw.size, chunk, C.byref(d), C.byref(p), C.byref(nc),
                                   threads or default_threads())
    if rc != 0:
        raise ValueError(f"brle encode rejected parameters (status {rc})")
    n = nc.value
    ptr = _take(p, n + 1, np.uint64)
    return BrleContainer(_take(d, int(ptr[-1]), np.uint8), ptr, n, chunk, raw.size)

def lzss_lane_encode(lane: bytes) -> bytes:
    a = np.frombuffer(bytes(lane), np.uint8) if len(lane) else np.zeros(1, np.uint8)
    cap = len(lane) + len(lane) // 8 + 16
    out = np.zeros(cap, np.uint8)
    n = lib().mdz_lzss_lane_encode(_ptr(a), len(lane), _ptr(out), cap)
    assert n <= cap
    return out[:n].tobytes()

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
    rc = lib().mdz_lzss_encode(_ptr(src), raw.size, chunk, C.byref(pay), C.byref(pb), C.byref(cl),
                                   C.byref(bo), C.byref(cm), C.byref(nc), threads or default_threads())
    if rc != 0:
        raise ValueError(f"lzss encode rejected parameters (status {rc})")
    n = nc.value
    return LzssContainer(_take(pay, pb.value, np.uint8), _take(cl, 32 * n, np.uint64),
                         _take(bo, n + 1, np.uint64), _take(cm, 32 * n, np.uint8), n, chunk)

def brle_trans_lane_encode(lane: bytes) -> bytes:
    """One lane of brle_trans::compress_func (include/brle/brle_trans.h:401-562)."""
    a = np.frombuffer(bytes(lane), np.uint8) if len(lane) else np.zeros(1, np.uint8)
    cap = len(lane) + len(lane) // 8 + 16
    out = np.zeros(cap, np.uint8)
    n = lib().mdz_brle_trans_lane_encode(_ptr(a), len(lane), _ptr(out), cap)
    assert n <= cap
    return out[:n].tobytes()

def brle_trans_encode(data: np.ndarray, chunk: int, threads: int | None = None) -> LzssContainer:
    """Lane-transposed byte RLE (brle_trans.h:565-669) in the LZSS lane container."""
    raw = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    src = raw if raw.size else np.zeros(1, np.uint8)
    pay, pb, cl, bo, cm, nc = u8p(), C.c_uint64(), u64p(), u64p(), u8p(), C.c_uint64()
    rc = lib().mdz_brle_trans_encode(_ptr(src), raw.size, chunk, C.byref(pay), C.byref(pb), C.byref(cl),
                                     C.byref(bo), C.byref(cm), C.byref(nc), threads or default_threads())
    if rc != 0:
        raise ValueError(f"brle_trans encode rejected parameters (status {rc})")
    n = nc.value
    return LzssContainer(_take(pay, pb.value, np.uint8), _take(cl, 32 * n, np.uint64),
                         _take(bo, n + 1, np.uint64), _take(cm, 32 * n, np.uint8), n, chunk)

class LzssPlainContainer:
    """Plain LZSS (lzss_gpu2.h:1450-1566,718-792): chunk streams packed back to
    back in `data`, lens[c] the inclusive prefix sum of their lengths."""

    def __init__(self, data, lens, n_chunks, chunk, raw_bytes):
        self.data, self.lens, self.n_chunks, self.chunk, self.raw_bytes = data, lens, n_chunks, chunk, raw_bytes

    def stream(self, c: int) -> bytes:
        b = int(self.lens[c - 1]) if c else 0
        return self.data[b:int(self.lens[c])].tobytes()

def lzss_plain_encode(data: np.ndarray, chunk: int, threads: int | None = None) -> LzssPlainContainer:
    raw = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    src = raw if raw.size else np.zeros(1, np.uint8)
    d, ln, nc = u8p(), u64p(), C.c_uint64()
    rc = lib().mdz_lzss_plain_encode(_ptr(src), raw.size, chunk, C.byref(d), C.byref(ln), C.byref(nc),
                                     threads or default_threads())
    if rc != 0:
        raise ValueError(f"plain lzss encode rejected parameters (status {rc})")
    n = nc.value
    lens = _take(ln, n, np.uint64)
    return LzssPlainContainer(_take(d, int(lens[-1]) if n else 0, np.uint8), lens, n, chunk, raw.size)

def gen_lzss_markov(n: int, seed: int) -> np.ndarray:
    out = np.empty(max(n, 1), np.uint8)
    lib().mdz_gen_lzss_markov(_ptr(out), n, seed)
    return out[:n]

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

def gen_rlev1_u32(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    lib().mdz_gen_rlev1_u32(_ptr(out), n, seed)
    return out

def gen_brle_u8(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.uint8)
    lib().mdz_gen_brle_u8(_ptr(out), n, seed)
    return out


def gen_markov(n: int, seed: int, copy_p: float) -> np.ndarray:
    """The Markov+copy source with its own transition table (seed) and copy rate."""
    cdf = lib().mdz_markov_table(seed)
    out = np.empty(max(n, 1), np.uint8)
    lib().mdz_markov_gen(cdf, _ptr(out), n, seed ^ 0x5BD1E995, float(copy_p))
    lib().mdz_free(cdf)
    return out[:n]
