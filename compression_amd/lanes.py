"""Device-resident lane-transposed containers (RLE v1, RLE v2, LZSS).

The container shape is the reference's: chunk c's payload occupies
[blk_off[c], blk_off[c+1]) and holds 32 lane streams of col_len[32c+t] bytes
interleaved in 4-byte words (include/rle_v1/encoder.h:540-588,
include/rlev2/encoder_trans.h:291-349, include/lzss/lzss_gpu2.h:497-509).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceBuffer, LanesT


class DeviceLanes:
    """Payload + metadata uploaded once and kept resident in HBM."""

    def __init__(self, payload: np.ndarray, blk_off: np.ndarray, col_len: np.ndarray,
                 n_chunks: int, chunk_bytes: int, elem_bytes: int = 0, lane_line: int = 0,
                 col_map: np.ndarray | None = None, stream=None):
        payload = np.ascontiguousarray(payload, np.uint8)
        self.payload = DeviceBuffer.from_array(payload, pad=16, stream=stream)
        self.blk_off = DeviceBuffer.from_array(np.ascontiguousarray(blk_off, np.uint64), stream=stream)
        self.col_len = DeviceBuffer.from_array(np.ascontiguousarray(col_len, np.uint64), stream=stream)
        self.col_map = (DeviceBuffer.from_array(np.ascontiguousarray(col_map, np.uint8), stream=stream)
                        if col_map is not None else None)
        self.n_chunks, self.chunk_bytes = int(n_chunks), int(chunk_bytes)
        self.elem_bytes, self.lane_line = int(elem_bytes), int(lane_line)
        self.payload_bytes = payload.nbytes
        # bytes of each chunk's lane quarters (lanes 8q..8q+7) once staged,
        # every lane rounded up to 16: sizes the per-quarter LDS area
        cl = np.asarray(col_len, np.int64)[:32 * self.n_chunks]
        if cl.size == 32 * self.n_chunks and self.n_chunks:
            self.part_bytes = ((cl + 15) & ~15).reshape(self.n_chunks, 4, 8).sum(axis=2)
            self.max_part = int(self.part_bytes.max())
        else:
            self.part_bytes = np.zeros((self.n_chunks, 4), np.int64)
            self.max_part = 0

    def struct(self) -> LanesT:
        return LanesT(self.payload.ptr, self.blk_off.ptr, self.col_len.ptr,
                      self.col_map.ptr if self.col_map is not None else None,
                      self.n_chunks, self.chunk_bytes, self.elem_bytes, self.lane_line,
                      self.payload_bytes)

    def meta_bytes(self) -> int:
        """Metadata bytes a decode reads: col_len, blk_off (+ col_map)."""
        return 8 * 32 * self.n_chunks + 8 * (self.n_chunks + 1) + (32 * self.n_chunks if self.col_map else 0)

    @property
    def out_bytes(self) -> int:
        return self.n_chunks * self.chunk_bytes


def as_ptr(x) -> int | None:
    if x is None:
        return None
    if isinstance(x, DeviceBuffer):
        return x.ptr
    return int(x)


def byref(s):
    return C.byref(s)
