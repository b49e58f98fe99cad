"""Batched persistent decode of an RLE v2 container and a byte-RLE container
in one launch (mdz_batch_submit), with the scheduler's work-item order.

Work items are ``codec << 31 | part << 29 | chunk`` (codec 0 = RLE v2, 1 =
byte RLE; part is the RLE v2 lane quarter when the batch decodes RLE v2 in
quarters).  ``Batch.size_order`` is the compressed-size heuristic: items
longest first by compressed bytes, byte RLE weighted by a cost factor.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BatchT, DeviceBuffer, check, lib
from .lanes import as_ptr

RLEV2, BRLE = 0, 1


def item(codec: int, chunk: np.ndarray, part=0) -> np.ndarray:
    """codec << 31 | part << 29 | chunk (part: RLE v2 lane quarter, 0..3)."""
    return ((np.uint32(codec) << np.uint32(31)) | (np.asarray(part, np.uint32) << np.uint32(29)) |
            np.asarray(chunk, np.uint32))


class Batch:
    """Keeps the launch description, the cursor and an optional item order.

    ``quarters`` decodes RLE v2 in lane quarters (four 4-wave workgroups per
    chunk instead of one 8-wave workgroup); ``window`` is the RLE v2 consumer
    window (run records claimed at once, 0 = 16)."""

    def __init__(self, rlev2=None, rlev2_dst=None, rlev2_dst_bytes=0, rlev2_status=None,
                 rlev2_max_chunk=0, brle=None, brle_dst=None, brle_dst_bytes=0, brle_status=None,
                 quarters: bool = False, window: int = 0):
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
            b.rlev2_max_part = rlev2.max_part if quarters else 0
            b.rlev2_window = int(window)
        self.parts = 4 if (rlev2 is not None and quarters and rlev2.max_part) else 1
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

    def all_items(self) -> np.ndarray:
        """Every work item, in index order (RLE v2 units, then byte-RLE chunks)."""
        n1 = self.rlev2.n_chunks if self.rlev2 is not None else 0
        n2 = self.brle.n_chunks if self.brle is not None else 0
        k = np.arange(n1 * self.parts)
        return np.concatenate([item(RLEV2, k // self.parts, k % self.parts), item(BRLE, np.arange(n2))])

    def size_order(self, brle_weight: float = 1.0) -> np.ndarray:
        """Items longest-first by compressed bytes: RLE v2 units by their lanes'
        bytes (quarters or whole chunks), byte-RLE chunks by theirs."""
        c = self.rlev2
        if c is not None and self.parts > 1:
            r2 = c.part_bytes.reshape(-1)
        elif c is not None:
            r2 = c.part_bytes.sum(axis=1)
        else:
            r2 = np.zeros(0, np.int64)
        b2 = self.brle.chunk_sizes if self.brle is not None else np.zeros(0, np.int64)
        cost = np.concatenate([r2.astype(np.float64), brle_weight * b2.astype(np.float64)])
        return self.all_items()[np.lexsort((np.arange(cost.size), -cost))]

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
