"""ctypes loader for the mdz HIP library (compression_amd/libmdz.so).

The library is the product: there is no CPU fallback.  If libmdz.so is
missing this module raises at load time.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdz.so")

vp = C.c_void_p
u64 = C.c_uint64
u32 = C.c_uint32
i32 = C.c_int32


class MdzError(RuntimeError):
    def __init__(self, rc: int, what: str):
        detail = ""
        try:
            detail = (_lib.mdz_last_error() or b"").decode() if _lib is not None else ""
        except Exception:
            pass
        super().__init__(f"{what} failed with mdz error {rc}" + (f" [{detail}]" if detail else ""))
        self.rc = rc


class LanesT(C.Structure):
    """mirror of mdz_lanes_t (include/mdz.h)"""
    _fields_ = [("payload", vp), ("blk_off", vp), ("col_len", vp), ("col_map", vp),
                ("n_chunks", u64), ("chunk_bytes", u64), ("elem_bytes", u32), ("lane_line", u32),
                ("payload_bytes", u64)]


class BatchT(C.Structure):
    """mdz_batch_t (include/mdz.h)."""
    _fields_ = [("rlev2", LanesT), ("rlev2_dst", vp), ("rlev2_dst_bytes", u64), ("rlev2_status", vp),
                ("rlev2_max_chunk", u32), ("brle_max_chunk", u32), ("rlev2_max_part", u32),
                ("rlev2_window", u32), ("brle_ptr", vp), ("brle_data", vp),
                ("brle_n_chunks", u64), ("brle_chunk_bytes", u64), ("brle_dst", vp),
                ("brle_dst_bytes", u64), ("brle_status", vp), ("items", vp), ("n_items", u32),
                ("blocks", u32), ("cursor", vp)]


class LanesTuneT(C.Structure):
    _fields_ = [("ring_words", u32), ("max_chunk_bytes", u32), ("max_part_bytes", u32),
                ("max_lane_tokens", u32), ("rlev2_window", u32), ("reserved", u32 * 3)]


# name -> (restype, argtypes); every symbol declared in include/mdz.h
SIGNATURES = {
    "mdz_version": (C.c_char_p, []),
    "mdz_last_error": (C.c_char_p, []),
    "mdz_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mdz_device_cus": (C.c_int, [C.POINTER(C.c_int)]),
    "mdz_dev_malloc": (C.c_int, [C.POINTER(vp), u64]),
    "mdz_dev_free": (C.c_int, [vp]),
    "mdz_pinned_alloc": (C.c_int, [C.POINTER(vp), u64]),
    "mdz_pinned_free": (C.c_int, [vp]),
    "mdz_memcpy_htod": (C.c_int, [vp, vp, u64, vp]),
    "mdz_memcpy_dtoh": (C.c_int, [vp, vp, u64, vp]),
    "mdz_memset_d8": (C.c_int, [vp, C.c_int, u64, vp]),
    "mdz_dev_copy": (C.c_int, [vp, vp, u64, vp]),
    "mdz_stream_create": (C.c_int, [C.POINTER(vp)]),
    "mdz_stream_destroy": (C.c_int, [vp]),
    "mdz_stream_sync": (C.c_int, [vp]),
    "mdz_device_sync": (C.c_int, []),
    "mdz_event_create": (C.c_int, [C.POINTER(vp)]),
    "mdz_event_destroy": (C.c_int, [vp]),
    "mdz_event_record": (C.c_int, [vp, vp]),
    "mdz_event_sync": (C.c_int, [vp]),
    "mdz_event_query": (C.c_int, [vp]),
    "mdz_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
    "mdz_host_free": (None, [vp]),
    "mdz_rlev1_decompress": (C.c_int, [C.POINTER(LanesT), vp, u64, vp, vp]),
    "mdz_rlev1_decompress_tuned": (C.c_int, [C.POINTER(LanesT), vp, u64, vp,
                                             C.POINTER(LanesTuneT), vp]),
    "mdz_rlev1_decompress_host": (C.c_int, [vp, C.POINTER(vp), u64, C.POINTER(u64), C.c_int,
                                            u64, u32, vp, vp, u64]),
    "mdz_rle_v1_decompress_gpu": (C.c_int, [vp, C.POINTER(vp), u64, C.POINTER(u64), C.c_int,
                                            u64, u32]),
    "mdz_rlev2_decompress": (C.c_int, [C.POINTER(LanesT), vp, u64, vp, vp]),
    "mdz_rlev2_decompress_tuned": (C.c_int, [C.POINTER(LanesT), vp, u64, vp,
                                             C.POINTER(LanesTuneT), vp]),
    "mdz_rlev2_decompress_host": (C.c_int, [vp, u64, u64, vp, vp, C.POINTER(vp), C.POINTER(u64),
                                            u64, u32, u32]),
    "mdz_brle_decompress": (C.c_int, [vp, u64, u64, vp, vp, u64, vp, vp]),
    "mdz_brle_decompress_tuned": (C.c_int, [vp, u64, u64, vp, vp, u64, vp, u32, vp]),
    "mdz_brle_decompress_host": (C.c_int, [vp, C.POINTER(vp), u64, C.POINTER(u64), u64]),
    "mdz_brle_decompress_gpu": (C.c_int, [vp, C.POINTER(vp), u64, C.POINTER(u64)]),
    "mdz_lzss_decompress": (C.c_int, [C.POINTER(LanesT), vp, u64, vp, vp]),
    "mdz_lzss_decompress_host": (C.c_int, [vp, u64, C.POINTER(vp), C.POINTER(u64), u64,
                                           C.POINTER(vp), C.POINTER(u64)]),
    "mdz_meta_split": (C.c_int, [vp, u64, C.POINTER(vp), C.POINTER(u64), C.POINTER(u64)]),
    "mdz_rlev2_plain_decompress": (C.c_int, [vp, u64, vp, u64, u64, vp, u64, vp, vp]),
    "mdz_rlev2_plain_decompress_host": (C.c_int, [vp, u64, u64, C.POINTER(vp), C.POINTER(u64)]),
    "mdz_rlev2_plain_decompress_gpu": (C.c_int, [vp, u64, C.POINTER(vp), C.POINTER(u64)]),
    "mdz_inflate_batch": (C.c_int, [vp, vp, C.c_int, C.c_int, vp]),
    "mdz_inflate_batch_sched": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, u32, vp, vp, vp]),
    "mdz_inflate_batch_sched_noalloc": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, u32, vp, vp]),
    "mdz_inflate_slots": (C.c_int, [C.POINTER(u32)]),
    "mdz_inflate_live_open": (C.c_int, [u32, C.c_int, u32, u32, C.POINTER(vp)]),
    "mdz_inflate_live_submit": (C.c_int, [vp, vp, u32, C.POINTER(u32)]),
    "mdz_inflate_live_statuses": (vp, [vp]),
    "mdz_inflate_live_expired": (C.c_int, [vp]),
    "mdz_inflate_live_close": (C.c_int, [vp, vp]),
    "mdz_serve_open": (C.c_int, [u64, u32, u32, u32, C.POINTER(vp)]),
    "mdz_serve_submit": (C.c_int, [vp, vp]),
    "mdz_serve_poll": (C.c_int, [vp, vp, u32, C.POINTER(u32)]),
    "mdz_serve_release": (C.c_int, [vp, u64]),
    "mdz_serve_pending": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64)]),
    "mdz_serve_close": (C.c_int, [vp]),
    "mdz_copy_uncompressed": (C.c_int, [vp, C.c_int, vp]),
    "mdz_rlev1_compress_bound": (C.c_int, [u64, u64, C.c_uint32, C.c_uint32, C.POINTER(u64), C.POINTER(u64),
                                           C.POINTER(u64)]),
    "mdz_rlev1_compress": (C.c_int, [vp, u64, u64, C.c_uint32, C.c_uint32, vp, vp, vp, vp, vp]),
    "mdz_rle_v1_compress_gpu": (C.c_int, [vp, C.POINTER(vp), u64, C.POINTER(u64), C.c_int, u64, C.c_uint32]),
    "mdz_rlev2_compress_bound": (C.c_int, [u64, u64, C.c_uint32, C.c_uint32, C.POINTER(u64), C.POINTER(u64),
                                           C.POINTER(u64)]),
    "mdz_rlev2_compress": (C.c_int, [vp, u64, u64, C.c_uint32, C.c_uint32, vp, vp, vp, vp, vp, vp]),
    "mdz_brle_trans_decompress": (C.c_int, [C.POINTER(LanesT), vp, u64, vp, vp]),
    "mdz_lzss_plain_decompress": (C.c_int, [vp, u64, vp, u64, u64, vp, u64, vp, vp]),
    "mdz_lzss_compress_bound": (C.c_int, [u64, u64, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
    "mdz_lzss_compress": (C.c_int, [vp, u64, u64, vp, vp, vp, vp, vp, vp]),
    "mdz_lzss_compress_gpu": (C.c_int, [vp, u64, C.POINTER(vp), C.POINTER(u64), C.POINTER(vp), C.POINTER(vp),
                                        C.POINTER(u64)]),
    "mdz_brle_compress_bound": (C.c_int, [u64, u64, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
    "mdz_brle_compress": (C.c_int, [vp, u64, u64, vp, vp, vp, vp]),
    "mdz_brle_compress_gpu": (C.c_int, [vp, C.POINTER(vp), u64, C.POINTER(u64)]),
    "mdz_file_write": (C.c_int, [C.c_char_p, vp, vp, vp, vp, vp, vp]),
    "mdz_file_open": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
    "mdz_file_info": (C.c_int, [vp, vp]),
    "mdz_file_raw_len": (C.c_int, [vp, vp]),
    "mdz_file_decode": (C.c_int, [vp, vp, u64, vp, u64, vp]),
    "mdz_file_close": (C.c_int, [vp]),
    "mdz_cost_features": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
    "mdz_batch_submit": (C.c_int, [C.POINTER(BatchT), vp]),
    "mdz_graph_begin": (C.c_int, [vp]),
    "mdz_graph_end": (C.c_int, [vp, C.POINTER(vp)]),
    "mdz_graph_launch": (C.c_int, [vp, vp]),
    "mdz_graph_destroy": (C.c_int, [vp]),
    "mdz_graph_nodes": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64)]),
    "mdz_mlp_forward": (C.c_int, [vp, vp, C.c_int, vp, vp]),
    "mdz_mlp_grad": (C.c_int, [vp, vp, vp, C.c_int, vp, vp, vp]),
    "mdz_mlp_adam": (C.c_int, [vp, vp, vp, vp, C.c_float, u32, vp]),
    "mdz_deflate_decompress_gpu": (C.c_int, [vp, C.POINTER(vp), u64, C.POINTER(u64), vp, u64, vp, u64,
                                             u64]),
    "mdz_cudf_uncompress_host": (C.c_int, [vp, u64, C.POINTER(vp), C.POINTER(u64)]),
    "mdz_cudf_uncompress": (C.c_int, [C.c_char_p, C.c_char_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m compression_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise MdzError(rc, what)


# ------------------------------------------------------------ plumbing ----
class DeviceBuffer:
    """Owned device allocation (hipMalloc through the C ABI)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = vp()
        check(lib().mdz_dev_malloc(C.byref(p), self.nbytes), "mdz_dev_malloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray, pad: int = 0, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes + pad)
        if a.nbytes:
            check(lib().mdz_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes, stream), "mdz_memcpy_htod")
        if stream is None:
            check(lib().mdz_stream_sync(None), "mdz_stream_sync")
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().mdz_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes, stream), "mdz_memcpy_htod")

    def download(self, nbytes: int | None = None, dtype=np.uint8, stream=None) -> np.ndarray:
        n = self.nbytes if nbytes is None else int(nbytes)
        out = np.empty(n, np.uint8)
        if n:
            check(lib().mdz_memcpy_dtoh(out.ctypes.data, self.ptr, n, stream), "mdz_memcpy_dtoh")
        check(lib().mdz_stream_sync(stream), "mdz_stream_sync")
        return out.view(dtype)

    def fill(self, value: int, stream=None) -> None:
        check(lib().mdz_memset_d8(self.ptr, value, self.nbytes, stream), "mdz_memset_d8")

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().mdz_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _PinnedOwner:
    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        try:
            lib().mdz_pinned_free(self.ptr)
        except Exception:
            pass


def pinned_copy(a: np.ndarray) -> np.ndarray:
    """A copy of `a` in pinned host memory (freed with the returned array):
    host-to-device copies from it are asynchronous DMA."""
    a = np.ascontiguousarray(a)
    p = vp()
    check(lib().mdz_pinned_alloc(C.byref(p), a.nbytes), "mdz_pinned_alloc")
    raw = (C.c_uint8 * max(a.nbytes, 1)).from_address(p.value)
    raw.owner = _PinnedOwner(p.value)          # frees the memory with the last view
    out = np.frombuffer(raw, np.uint8, count=a.nbytes).view(a.dtype).reshape(a.shape)
    out[...] = a
    return out


class Stream:
    def __init__(self):
        p = vp()
        check(lib().mdz_stream_create(C.byref(p)), "mdz_stream_create")
        self.ptr = p.value

    def sync(self) -> None:
        check(lib().mdz_stream_sync(self.ptr), "mdz_stream_sync")

    def __del__(self):
        try:
            lib().mdz_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = vp()
        check(lib().mdz_event_create(C.byref(p)), "mdz_event_create")
        self.ptr = p.value

    def record(self, stream=None) -> None:
        check(lib().mdz_event_record(self.ptr, stream.ptr if isinstance(stream, Stream) else stream),
              "mdz_event_record")

    def sync(self) -> None:
        check(lib().mdz_event_sync(self.ptr), "mdz_event_sync")

    def query(self) -> bool:
        rc = lib().mdz_event_query(self.ptr)
        if rc < 0:
            check(rc, "mdz_event_query")
        return rc == 1

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().mdz_event_elapsed_ms(self.ptr, end.ptr, C.byref(ms)), "mdz_event_elapsed_ms")
        return float(ms.value)

    def __del__(self):
        try:
            lib().mdz_event_destroy(self.ptr)
        except Exception:
            pass


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().mdz_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def device_cus() -> int:
    """Compute units of the current device."""
    n = C.c_int(0)
    check(lib().mdz_device_cus(C.byref(n)), "mdz_device_cus")
    return n.value
