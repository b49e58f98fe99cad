"""Batched persistent decode of an RLE v2 container and a byte-RLE container
in one launch (mdz_batch_submit), with the scheduler's work-item order.

Work items are ``codec << 31 | chunk`` (codec 0 = RLE v2, 1 = byte RLE).
``size_order`` is the compressed-size heuristic: items longest first by
compressed bytes, each codec's bytes weighted by a per-codec cost factor.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BatchT, DeviceBuffer, check, lib
from .lanes import as_ptr

RLEV2, BRLE = 0, 1


def item(codec: int, chunk: np.ndarray) -> np.ndarray:
    return (np.uint32(codec) << np.uint32(31)) | np.asarray(chunk, np.uint32)


class Batch:
    """Keeps the launch description, the cursor and an optional item order."""

    def __init__(self, rlev2=None, rlev2_dst=None, rlev2_dst_bytes=0, rlev2_status=None,
                 rlev2_max_chunk=0, brle=None, brle_dst=None, brle_dst_bytes=0, brle_status=None):
        self.rlev2, self.brle = rlev2, brle
        self.cursor = DeviceBuffer(4)
        self.items = None
        self.n_items = 0
        self.blocks = 0
        b = BatchT()
        if rlev2 is not None:
            b.rlev2 = rlev2.struct()
            b.rlev2_dst = as_ptr(rlev2_dst)
            b.rlev2_dst_bytes = int(rlev2_dst_bytes)
            b.rlev2_status = as_ptr(rlev2_status)
            b.rlev2_max_chunk = int(rlev2_max_chunk)
        if brle is not None:
            b.brle_ptr = brle.ptr.ptr
            b.brle_data = brle.data.ptr
            b.brle_n_chunks = brle.n_chunks
            b.brle_chunk_bytes = brle.chunk_bytes
            b.brle_dst = as_ptr(brle_dst)
            b.brle_dst_bytes = int(brle_dst_bytes)
            b.brle_status = as_ptr(brle_status)
            b.brle_max_chunk = brle.max_chunk
        b.cursor = self.cursor.ptr
        self.desc = b

    def set_order(self, items: np.ndarray | None) -> None:
        if items is None:
            self.items, self.n_items = None, 0
        else:
            items = np.ascontiguousarray(items, np.uint32)
            self.items = DeviceBuffer.from_array(items)
            self.n_items = items.size
        self.desc.items = self.items.ptr if self.items is not None else None
        self.desc.n_items = self.n_items

    def submit(self, stream=None, blocks: int | None = None) -> None:
        self.desc.blocks = self.blocks if blocks is None else int(blocks)
        check(lib().mdz_batch_submit(C.byref(self.desc), as_ptr(stream)), "mdz_batch_submit")


def size_order(rlev2_chunk_bytes: np.ndarray, brle_chunk_bytes: np.ndarray,
               brle_weight: float = 1.0) -> np.ndarray:
    """Items longest-first by (weighted) compressed chunk size."""
    n1, n2 = len(rlev2_chunk_bytes), len(brle_chunk_bytes)
    cost = np.concatenate([np.asarray(rlev2_chunk_bytes, np.float64),
                           brle_weight * np.asarray(brle_chunk_bytes, np.float64)])
    codes = np.concatenate([item(RLEV2, np.arange(n1)), item(BRLE, np.arange(n2))])
    return codes[np.lexsort((np.arange(n1 + n2), -cost))]
