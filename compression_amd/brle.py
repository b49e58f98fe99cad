"""Byte RLE decode -- host-side mirror of the reference's ``brle`` namespace.

``decompress_gpu`` takes the reference container
[u32 n_ptr][u64 ptr[n_ptr]][u64 out_bytes][data] (include/brle/brle.h:305-342)
and returns the decoded bytes; ``decompress`` is the device-resident form.
``decompress_trans`` decodes the lane-transposed variant (``brle_trans``,
include/brle/brle_trans.h:72-202), whose container is the LZSS lane container.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceBuffer, MdzError, check, lib
from .lanes import as_ptr


class DeviceBrle:
    """ptr[n+1] and the compressed bytes, resident in HBM."""

    def __init__(self, data: np.ndarray, ptr: np.ndarray, n_chunks: int, chunk_bytes: int, stream=None):
        data = np.ascontiguousarray(data, np.uint8)
        self.data = DeviceBuffer.from_array(data, pad=16, stream=stream)
        self.ptr = DeviceBuffer.from_array(np.ascontiguousarray(ptr, np.uint64), stream=stream)
        self.n_chunks, self.chunk_bytes = int(n_chunks), int(chunk_bytes)
        p = np.asarray(ptr, np.int64)
        self.max_chunk = int(np.diff(p).max()) if len(p) > 1 else 0
        self.chunk_sizes = np.diff(p) if len(p) > 1 else np.zeros(0, np.int64)
        self.data_bytes = data.nbytes


def decompress_trans(lanes, dst, dst_bytes: int | None = None, status=None, stream=None) -> None:
    """Enqueue the decode of a brle_trans container held as ``lanes.DeviceLanes``
    (with ``col_map``)."""
    if lanes.col_map is None:
        raise ValueError("brle_trans containers carry col_map")
    s = lanes.struct()
    n = lanes.out_bytes if dst_bytes is None else int(dst_bytes)
    check(lib().mdz_brle_trans_decompress(C.byref(s), as_ptr(dst), n, as_ptr(status), as_ptr(stream)),
          "mdz_brle_trans_decompress")


def decompress(b: DeviceBrle, dst, dst_bytes: int, status=None, stream=None, tuned: bool = True) -> None:
    if tuned:
        rc = lib().mdz_brle_decompress_tuned(b.ptr.ptr, b.n_chunks, b.chunk_bytes, b.data.ptr, as_ptr(dst),
                                             int(dst_bytes), as_ptr(status), b.max_chunk, as_ptr(stream))
    else:
        rc = lib().mdz_brle_decompress(b.ptr.ptr, b.n_chunks, b.chunk_bytes, b.data.ptr, as_ptr(dst),
                                       int(dst_bytes), as_ptr(status), as_ptr(stream))
    check(rc, "mdz_brle_decompress")


def decompress_gpu(container: bytes, chunk_bytes: int = 256 * 1024) -> bytes:
    src = np.frombuffer(container, np.uint8)
    out = C.c_void_p()
    n = C.c_uint64()
    rc = lib().mdz_brle_decompress_host(src.ctypes.data, C.byref(out), src.size, C.byref(n), chunk_bytes)
    if rc not in (0, -5):
        raise MdzError(rc, "mdz_brle_decompress_host")
    data = C.string_at(out.value, n.value) if n.value else b""
    lib().mdz_host_free(out)
    if rc != 0:
        raise MdzError(rc, "mdz_brle_decompress_host (corrupt chunk)")
    return data


def compress_bound(in_bytes: int, chunk_bytes: int = 256 * 1024) -> tuple[int, int, int]:
    """(n_chunks, work_bytes, data_cap) for ``compress``."""
    n, w, d = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(lib().mdz_brle_compress_bound(in_bytes, chunk_bytes, C.byref(n), C.byref(w), C.byref(d)),
          "mdz_brle_compress_bound")
    return n.value, w.value, d.value


def compress(src, in_bytes: int, data, ptr, work, chunk_bytes: int = 256 * 1024, stream=None) -> None:
    """Device-resident encode: ``src`` (in_bytes) -> ``data`` + ``ptr`` (u64[n+1]),
    with ``work`` sized by ``compress_bound``."""
    check(lib().mdz_brle_compress(as_ptr(src), in_bytes, chunk_bytes, as_ptr(work), as_ptr(data),
                                  as_ptr(ptr), as_ptr(stream)), "mdz_brle_compress")


def compress_gpu(data: bytes) -> bytes:
    """brle::compress_gpu: host bytes -> the reference container (256 KiB chunks)."""
    src = np.frombuffer(data, np.uint8)
    out = C.c_void_p()
    n = C.c_uint64()
    check(lib().mdz_brle_compress_gpu(src.ctypes.data, C.byref(out), src.size, C.byref(n)),
          "mdz_brle_compress_gpu")
    blob = C.string_at(out.value, n.value)
    lib().mdz_host_free(out)
    return blob
