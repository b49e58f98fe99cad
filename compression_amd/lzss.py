"""LZSS decode -- host-side mirror of the reference's ``lzss`` namespace.

``decompress_gpu`` keeps the shape of ``lzss::decompress_gpu``
(include/lzss/lzss_gpu2.h:2129-2314): payload plus the four ``.meta`` blobs
{chunk_size, blk_off[n+1], col_len[32n] (sorted), col_map[32n]} -> host
bytes.  ``decompress`` is the device-resident, stream-ordered form, and
``split_meta`` reads a ``.meta`` side file image (src/lzss/lzss_gpu.cu:109-133).
``decompress_plain`` is the plain per-chunk format's device decode
(``lzss::kernel_decompress`` with ``lens``, include/lzss/lzss_gpu2.h:1908-1913,
718-792).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import MdzError, check, lib
from .lanes import DeviceLanes, as_ptr


def decompress(lanes: DeviceLanes, dst, dst_bytes: int | None = None, status=None,
               stream=None) -> None:
    """Enqueue the decode of every chunk (``lanes.col_map`` is required)."""
    if lanes.col_map is None:
        raise ValueError("LZSS containers carry col_map")
    s = lanes.struct()
    n = lanes.out_bytes if dst_bytes is None else int(dst_bytes)
    check(lib().mdz_lzss_decompress(C.byref(s), as_ptr(dst), n, as_ptr(status), as_ptr(stream)),
          "mdz_lzss_decompress")


def decompress_plain(data, data_bytes: int, lens, n_chunks: int, chunk_bytes: int, dst, dst_bytes: int,
                     status=None, stream=None) -> None:
    """Enqueue the plain-format decode: chunk c's stream is
    ``data[lens[c-1]:lens[c]]`` (device pointers; ``lens`` the inclusive
    prefix sums) and decodes to ``dst[c*chunk_bytes:]``, clipped to
    ``dst_bytes``."""
    check(lib().mdz_lzss_plain_decompress(as_ptr(data), int(data_bytes), as_ptr(lens), int(n_chunks),
                                          int(chunk_bytes), as_ptr(dst), int(dst_bytes), as_ptr(status),
                                          as_ptr(stream)), "mdz_lzss_plain_decompress")


def split_meta(meta: bytes) -> list[bytes]:
    """``[u64 num][u64 len[num]][blob...]`` -> blobs (include/utils.h:46-118)."""
    buf = np.frombuffer(meta, np.uint8)
    num = C.c_uint64(0)
    rc = lib().mdz_meta_split(buf.ctypes.data, buf.size, None, None, C.byref(num))
    if rc not in (0, -1):
        raise MdzError(rc, "mdz_meta_split")
    k = num.value
    blobs = (C.c_void_p * max(k, 1))()
    lens = (C.c_uint64 * max(k, 1))()
    check(lib().mdz_meta_split(buf.ctypes.data, buf.size, blobs, lens, C.byref(num)), "mdz_meta_split")
    base = buf.ctypes.data
    return [bytes(buf[blobs[i] - base: blobs[i] - base + lens[i]]) for i in range(k)]


def decompress_gpu(payload: bytes, metadata: list[bytes]) -> bytes:
    src = np.frombuffer(payload, np.uint8) if len(payload) else np.zeros(1, np.uint8)
    arrs = [np.frombuffer(b, np.uint8) if len(b) else np.zeros(1, np.uint8) for b in metadata]
    ptrs = (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])
    lens = (C.c_uint64 * max(len(arrs), 1))(*[len(b) for b in metadata])
    out = C.c_void_p()
    n = C.c_uint64()
    rc = lib().mdz_lzss_decompress_host(src.ctypes.data, len(payload), ptrs, lens, len(metadata),
                                        C.byref(out), C.byref(n))
    if rc not in (0, -5):
        raise MdzError(rc, "mdz_lzss_decompress_host")
    data = C.string_at(out.value, n.value) if n.value else b""
    lib().mdz_host_free(out)
    if rc != 0:
        raise MdzError(rc, "mdz_lzss_decompress_host (corrupt chunk)")
    return data


def compress_bound(in_bytes: int, chunk_bytes: int) -> tuple[int, int, int]:
    """(n_chunks, work_bytes, payload_cap) for ``compress``."""
    n, w, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(lib().mdz_lzss_compress_bound(in_bytes, chunk_bytes, C.byref(n), C.byref(w), C.byref(p)),
          "mdz_lzss_compress_bound")
    return n.value, w.value, p.value


def compress(src, in_bytes: int, chunk_bytes: int, work, payload, col_len, col_map, blk_off, stream=None) -> None:
    """Device-resident encode (lzss::compress_gpu): ``src`` -> ``payload``,
    sorted ``col_len`` (u64[32n]), ``col_map`` (u8[32n]) and ``blk_off`` (u64[n+1])."""
    check(lib().mdz_lzss_compress(as_ptr(src), in_bytes, chunk_bytes, as_ptr(work), as_ptr(payload),
                                  as_ptr(col_len), as_ptr(col_map), as_ptr(blk_off), as_ptr(stream)),
          "mdz_lzss_compress")


def compress_gpu(data: bytes) -> tuple[bytes, list[bytes]]:
    """lzss::compress_gpu: (payload, [chunk, blk_off, col_len, col_map] blobs)."""
    src = np.frombuffer(data, np.uint8)
    out, md, lens = C.c_void_p(), C.c_void_p(), C.c_void_p()
    n, num = C.c_uint64(), C.c_uint64()
    check(lib().mdz_lzss_compress_gpu(src.ctypes.data, src.size, C.byref(out), C.byref(n), C.byref(md),
                                      C.byref(lens), C.byref(num)), "mdz_lzss_compress_gpu")
    payload = C.string_at(out.value, n.value) if n.value else b""
    lib().mdz_host_free(out)
    blobs_p = C.cast(md, C.POINTER(C.c_void_p))
    lens_p = C.cast(lens, C.POINTER(C.c_uint64))
    blobs = []
    for k in range(num.value):
        blobs.append(C.string_at(blobs_p[k], lens_p[k]))
        lib().mdz_host_free(C.c_void_p(blobs_p[k]))
    lib().mdz_host_free(md)
    lib().mdz_host_free(lens)
    return payload, blobs
