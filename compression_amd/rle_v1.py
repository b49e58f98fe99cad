"""RLE v1 decode -- host-side mirror of the reference's ``rle_v1`` namespace.

``decompress_gpu`` keeps the shape of ``rle_v1::decompress_gpu``
(include/rle_v1/rlev1.h:99-199): host bytes in, host bytes out.
``decompress`` is the device-resident, stream-ordered form used by the
benchmark and the batch scheduler.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LanesTuneT, MdzError, check, lib
from .lanes import DeviceLanes, as_ptr


def decompress(lanes: DeviceLanes, dst, dst_bytes: int | None = None, status=None, stream=None,
               ring_words: int = 0, max_chunk_bytes: int = 0, max_lane_tokens: int = 0) -> None:
    """Enqueue the decode of every chunk of ``lanes`` into device buffer ``dst``.

    ``max_chunk_bytes`` (largest chunk payload, if the caller knows it) sizes
    the per-chunk LDS area; ``ring_words`` forces the streaming path and
    ``max_lane_tokens`` bounds the per-lane token lists (a chunk with a longer
    lane takes the checkpoint decode)."""
    s = lanes.struct()
    n = lanes.out_bytes if dst_bytes is None else int(dst_bytes)
    if ring_words or max_chunk_bytes or max_lane_tokens:
        tune = LanesTuneT(ring_words, max_chunk_bytes, 0, max_lane_tokens)
        rc = lib().mdz_rlev1_decompress_tuned(C.byref(s), as_ptr(dst), n, as_ptr(status), C.byref(tune),
                                              as_ptr(stream))
    else:
        rc = lib().mdz_rlev1_decompress(C.byref(s), as_ptr(dst), n, as_ptr(status), as_ptr(stream))
    check(rc, "mdz_rlev1_decompress")


def decompress_gpu(file_bytes: bytes, comp_col_len: int, chunk_size: int, elem_bytes: int,
                   col_len: np.ndarray, blk_off: np.ndarray) -> bytes:
    """Reference-shaped host entry: ``file_bytes`` = [u64 raw_bytes][chunks...]."""
    src = np.frombuffer(file_bytes, np.uint8)
    col = np.ascontiguousarray(col_len, np.uint64)
    off = np.ascontiguousarray(blk_off, np.uint64)
    out = C.c_void_p()
    n = C.c_uint64()
    rc = lib().mdz_rlev1_decompress_host(src.ctypes.data, C.byref(out), src.size, C.byref(n),
                                         comp_col_len, chunk_size, elem_bytes, col.ctypes.data,
                                         off.ctypes.data, len(off) - 1)
    if rc not in (0, -5):
        raise MdzError(rc, "mdz_rlev1_decompress_host")
    data = C.string_at(out.value, n.value) if n.value else b""
    lib().mdz_host_free(out)
    if rc != 0:
        raise MdzError(rc, "mdz_rlev1_decompress_host (corrupt chunk)")
    return data


def compress_bound(in_bytes: int, chunk_bytes: int, elem_bytes: int, comp_col_len: int) -> tuple[int, int, int]:
    """(n_chunks, work_bytes, payload_cap) for ``compress``."""
    n, w, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(lib().mdz_rlev1_compress_bound(in_bytes, chunk_bytes, elem_bytes, comp_col_len, C.byref(n),
                                         C.byref(w), C.byref(p)), "mdz_rlev1_compress_bound")
    return n.value, w.value, p.value


def compress(src, in_bytes: int, chunk_bytes: int, elem_bytes: int, comp_col_len: int, work, payload,
             col_len, blk_off, stream=None) -> None:
    """Device-resident encode (rle_v1::compress_gpu): ``src`` -> ``payload``,
    ``col_len`` (u64[32n]) and ``blk_off`` (u64[n+1])."""
    check(lib().mdz_rlev1_compress(as_ptr(src), in_bytes, chunk_bytes, elem_bytes, comp_col_len, as_ptr(work),
                                   as_ptr(payload), as_ptr(col_len), as_ptr(blk_off), as_ptr(stream)),
          "mdz_rlev1_compress")


def compress_gpu(data: bytes, comp_col_len: int, chunk_size: int, elem_bytes: int) -> bytes:
    """rle_v1::compress_gpu: returns [u64 size][payload] and writes col_len /
    blk_off to ./input_data/rle_v1_col_len.bin and rle_v1_blk_offset.bin."""
    src = np.frombuffer(data, np.uint8)
    out = C.c_void_p()
    n = C.c_uint64()
    check(lib().mdz_rle_v1_compress_gpu(src.ctypes.data, C.byref(out), src.size, C.byref(n), comp_col_len,
                                        chunk_size, elem_bytes), "mdz_rle_v1_compress_gpu")
    blob = C.string_at(out.value, n.value)
    lib().mdz_host_free(out)
    return blob
