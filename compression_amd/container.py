"""Self-describing container files (MDZF) and their staged reader.

A file carries one container of any codec with its metadata (blk_off,
col_len, col_map, Deflate stream sizes), so a decode needs nothing else: it
replaces the reference's working-directory side files (include/rle_v1/
rlev1.h:81-87,110-173) and the LZSS ``.meta`` blob (include/utils.h:46-118).
``File.decode`` moves the payload to HBM in pieces through pinned staging
buffers and decodes each piece as soon as it lands.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceBuffer, MdzError, check, lib
from .lanes import as_ptr

CODEC = {"rlev1": 1, "rlev2": 2, "brle": 3, "lzss": 4, "deflate": 5}


class FileInfo(C.Structure):
    _fields_ = [("codec", C.c_uint32), ("elem_bytes", C.c_uint32), ("lane_line", C.c_uint32),
                ("framing", C.c_uint32), ("n_chunks", C.c_uint64), ("chunk_bytes", C.c_uint64),
                ("raw_bytes", C.c_uint64), ("payload_bytes", C.c_uint64)]


def _arr(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype)


def write(path: str, codec: str, payload, blk_off, n_chunks: int, raw_bytes: int, chunk_bytes: int = 0,
          elem_bytes: int = 0, lane_line: int = 0, col_len=None, col_map=None, raw_len=None,
          framing: int = 0) -> None:
    pay = np.ascontiguousarray(np.frombuffer(payload, np.uint8) if isinstance(payload, (bytes, bytearray))
                               else payload, np.uint8).reshape(-1)
    info = FileInfo(CODEC[codec], elem_bytes, lane_line, framing, n_chunks, chunk_bytes, raw_bytes, pay.size)
    off, col, cmap, rl = _arr(blk_off, np.uint64), _arr(col_len, np.uint64), _arr(col_map, np.uint8), \
        _arr(raw_len, np.uint64)
    ptr = lambda a: None if a is None else a.ctypes.data
    check(lib().mdz_file_write(path.encode(), C.byref(info), ptr(pay) if pay.size else None, ptr(off),
                               ptr(col), ptr(cmap), ptr(rl)), "mdz_file_write")


def save(path: str, c) -> None:
    """Writes a host-library container (RLE v1 / RLE v2 / byte RLE / LZSS)."""
    kind = type(c).__name__
    if kind == "RleV1Container":
        write(path, "rlev1", c.payload, c.blk_off, c.n_chunks, c.raw_bytes, c.chunk, c.elem_bytes,
              c.line, col_len=c.col_len)
    elif kind == "RleV2Container":
        write(path, "rlev2", c.payload[:c.payload_bytes], c.blk_off, c.n_chunks, c.raw_bytes, c.chunk,
              c.elem_bytes, c.read_unit, col_len=c.col_len)
    elif kind == "BrleContainer":
        write(path, "brle", c.data, c.ptr, c.n_chunks, c.raw_bytes, c.chunk)
    elif kind == "LzssContainer":
        write(path, "lzss", c.payload, c.blk_off, c.n_chunks, c.raw_bytes, c.chunk, col_len=c.col_len,
              col_map=c.col_map)
    else:
        raise TypeError(f"not a container: {kind}")


def save_deflate(path: str, streams, raw_lens, framing: int = 1) -> None:
    """Deflate streams packed back to back; stream i spans blk_off[i]..blk_off[i+1]."""
    lens = np.array([len(s) for s in streams], np.uint64)
    off = np.zeros(len(streams) + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    rl = np.ascontiguousarray(raw_lens, np.uint64)
    write(path, "deflate", b"".join(streams), off, len(streams), int(rl.sum()), raw_len=rl,
          framing=framing)


class File:
    """An open container file (read-only mapping; device copies made on the first decode)."""

    def __init__(self, path: str):
        self._h = C.c_void_p()
        check(lib().mdz_file_open(path.encode(), C.byref(self._h)), "mdz_file_open")
        self.info = FileInfo()
        check(lib().mdz_file_info(self._h, C.byref(self.info)), "mdz_file_info")

    @property
    def codec(self) -> str:
        return {v: k for k, v in CODEC.items()}[self.info.codec]

    def raw_len(self) -> np.ndarray:
        out = np.zeros(self.info.n_chunks, np.uint64)
        check(lib().mdz_file_raw_len(self._h, out.ctypes.data), "mdz_file_raw_len")
        return out

    def decode(self, dst: DeviceBuffer, status: DeviceBuffer | None = None, stage_bytes: int = 0,
               stream=None) -> None:
        check(lib().mdz_file_decode(self._h, dst.ptr, dst.nbytes, status.ptr if status else None,
                                    stage_bytes, as_ptr(stream)), "mdz_file_decode")

    def close(self) -> None:
        if self._h:
            lib().mdz_file_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["CODEC", "File", "FileInfo", "MdzError", "save", "save_deflate", "write"]
