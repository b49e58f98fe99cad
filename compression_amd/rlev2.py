"""RLE v2 decode -- host-side mirror of the reference's ``rlev2`` namespace.

``decompress_gpu`` keeps the shape of ``rlev2::decompress_gpu``
(include/rlev2/decoder_trans.h:615-681): payload, n_chunks, blk_off (whose
last entry is the uncompressed size), col_len -> host bytes.  ``decompress``
is the device-resident, stream-ordered form.  ``decompress_plain`` and
``decompress_plain_gpu`` are the same two forms for the plain format (one
run stream per chunk, include/rlev2/decoder.h:190-262).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LanesTuneT, MdzError, check, lib
from .lanes import DeviceLanes, as_ptr


def decompress(lanes: DeviceLanes, dst, dst_bytes: int | None = None, status=None, stream=None,
               max_chunk_bytes: int = 0, quarters: bool = False, window: int = 0) -> None:
    """Enqueue the decode of every chunk; ``lanes.lane_line`` is READ_UNIT.

    ``quarters`` decodes each chunk's lane quarters with their own workgroups
    (LDS sized by the largest quarter, which the host knows from col_len);
    by default one workgroup decodes each whole chunk, which measures faster
    on the 1 GiB column (0.79 ms against 1.08 ms: a quarter's staging and
    setup cost nearly what a whole chunk's do).  ``window`` is the number of
    run records a consumer wavefront claims at once (8, 16, 32 or 64; 0 = 16)."""
    s = lanes.struct()
    n = lanes.out_bytes if dst_bytes is None else int(dst_bytes)
    part = lanes.max_part if quarters else 0
    if max_chunk_bytes or part or window:
        tune = LanesTuneT(0, max_chunk_bytes, part, 0, window)
        rc = lib().mdz_rlev2_decompress_tuned(C.byref(s), as_ptr(dst), n, as_ptr(status), C.byref(tune),
                                              as_ptr(stream))
    else:
        rc = lib().mdz_rlev2_decompress(C.byref(s), as_ptr(dst), n, as_ptr(status), as_ptr(stream))
    check(rc, "mdz_rlev2_decompress")


def decompress_gpu(payload: bytes, n_chunks: int, blk_off: np.ndarray, col_len: np.ndarray,
                   chunk_size: int, read_unit: int, elem_bytes: int) -> bytes:
    src = np.frombuffer(payload, np.uint8)
    off = np.ascontiguousarray(blk_off, np.uint64)
    col = np.ascontiguousarray(col_len, np.uint64)
    out = C.c_void_p()
    n = C.c_uint64()
    rc = lib().mdz_rlev2_decompress_host(src.ctypes.data, src.size, n_chunks, off.ctypes.data,
                                         col.ctypes.data, C.byref(out), C.byref(n), chunk_size,
                                         read_unit, elem_bytes)
    if rc not in (0, -5):
        raise MdzError(rc, "mdz_rlev2_decompress_host")
    data = C.string_at(out.value, n.value) if n.value else b""
    lib().mdz_host_free(out)
    if rc != 0:
        raise MdzError(rc, "mdz_rlev2_decompress_host (corrupt chunk)")
    return data


def decompress_plain(data, data_bytes: int, ptr, n_chunks: int, chunk_bytes: int, dst, dst_bytes: int,
                     status=None, stream=None) -> None:
    """Enqueue the plain-format decode (one run stream per chunk of int64
    values, include/rlev2/decoder.h:190-226): chunk c's stream is
    ``data[ptr[c]:ptr[c+1]]`` (device pointers; ``ptr`` the container's
    offset table) and decodes to ``dst[c*chunk_bytes:]``, clipped to
    ``dst_bytes``."""
    check(lib().mdz_rlev2_plain_decompress(as_ptr(data), int(data_bytes), as_ptr(ptr), int(n_chunks),
                                           int(chunk_bytes), as_ptr(dst), int(dst_bytes), as_ptr(status),
                                           as_ptr(stream)), "mdz_rlev2_plain_decompress")


def decompress_plain_gpu(container: bytes, chunk_size: int = 32 * 1024) -> bytes:
    """``rlev2::decompress_gpu`` of the plain container (decoder.h:227-262):
    [u32 n+1][u64 ptr[n+1]][u64 raw_bytes][streams] -> raw bytes."""
    src = np.frombuffer(container, np.uint8)
    out = C.c_void_p()
    n = C.c_uint64()
    rc = lib().mdz_rlev2_plain_decompress_host(src.ctypes.data, src.size, chunk_size, C.byref(out), C.byref(n))
    if rc not in (0, -5):
        raise MdzError(rc, "mdz_rlev2_plain_decompress_host")
    data = C.string_at(out.value, n.value) if n.value else b""
    lib().mdz_host_free(out)
    if rc != 0:
        raise MdzError(rc, "mdz_rlev2_plain_decompress_host (corrupt chunk)")
    return data


def compress_bound(in_bytes: int, chunk_bytes: int, elem_bytes: int, read_unit: int) -> tuple[int, int, int]:
    """(n_chunks, work_bytes, payload_cap) for ``compress``."""
    n, w, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(lib().mdz_rlev2_compress_bound(in_bytes, chunk_bytes, elem_bytes, read_unit, C.byref(n), C.byref(w),
                                         C.byref(p)), "mdz_rlev2_compress_bound")
    return n.value, w.value, p.value


def compress(src, in_bytes: int, chunk_bytes: int, elem_bytes: int, read_unit: int, work, payload, col_len,
             blk_off, payload_bytes=None, stream=None) -> None:
    """Device-resident encode (rlev2::compress_gpu_transpose): ``src`` -> ``payload``,
    ``col_len`` (u64[32n]), ``blk_off`` (u64[n+1], last = in_bytes) and the
    payload length in ``payload_bytes`` (device u64)."""
    check(lib().mdz_rlev2_compress(as_ptr(src), in_bytes, chunk_bytes, elem_bytes, read_unit, as_ptr(work),
                                   as_ptr(payload), as_ptr(col_len), as_ptr(blk_off), as_ptr(payload_bytes),
                                   as_ptr(stream)), "mdz_rlev2_compress")
