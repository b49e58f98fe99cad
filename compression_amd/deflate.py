"""Deflate / zlib inflate -- host-side mirror of cuDF's ``gpuinflate``
(include/cudf_deflate.cuh:1399-1407) and the repo's ``deflate::decompress_gpu``
(include/zlib/zlib.h:966-1082).

``DeviceBatch`` keeps a batch of streams and their output slots resident in
HBM with the descriptor array the kernels read; ``inflate`` is the plain
per-stream launch and ``inflate_sched`` the persistent work-queue launch;
``LiveInflate`` keeps one persistent launch running and takes streams while
it runs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceBuffer, MdzError, check, lib
from .lanes import as_ptr

DESC = np.dtype([("src", "<u8"), ("src_bytes", "<u8"), ("dst", "<u8"), ("dst_bytes", "<u8")])
STATUS = np.dtype([("bytes_written", "<u8"), ("status", "<i4"), ("reserved", "<u4")])
COST = np.dtype([("cycles", "<u8"), ("symbols", "<u4"), ("matches", "<u4"), ("stored_blocks", "<u4"),
                 ("fixed_blocks", "<u4"), ("dynamic_blocks", "<u4"), ("reserved", "<u4")])


class DeviceBatch:
    """Streams src[src_off[i] : src_off[i] + src_len[i]] and output slots of
    dst_caps[i] bytes, all device-resident."""

    def __init__(self, src: np.ndarray, src_off, src_len, dst_caps, stream=None):
        src = np.ascontiguousarray(src, np.uint8)
        self.count = len(src_len)
        self.src = DeviceBuffer.from_array(src, pad=16, stream=stream)
        caps = np.ascontiguousarray(dst_caps, np.uint64)
        self.dst_off = np.zeros(self.count, np.uint64)
        if self.count:
            self.dst_off[1:] = np.cumsum((caps + 15) // 16 * 16)[:-1]
        total = int(self.dst_off[-1] + caps[-1]) if self.count else 0
        self.dst = DeviceBuffer(max(total, 1))
        self.caps = caps
        d = np.zeros(self.count, DESC)
        d["src"] = self.src.ptr + np.asarray(src_off, np.uint64)
        d["src_bytes"] = np.asarray(src_len, np.uint64)
        d["dst"] = self.dst.ptr + self.dst_off
        d["dst_bytes"] = caps
        self.host_desc = d
        self.desc = DeviceBuffer.from_array(d, stream=stream)
        self.status = DeviceBuffer(max(self.count, 1) * STATUS.itemsize)
        self.cursor = DeviceBuffer(4)
        self.order = None
        self.cost = None

    def enable_cost_log(self) -> None:
        self.cost = DeviceBuffer(max(self.count, 1) * COST.itemsize)

    def costs(self) -> np.ndarray:
        return self.cost.download(self.count * COST.itemsize).view(COST)

    def set_order(self, order) -> None:
        self.order = DeviceBuffer.from_array(np.ascontiguousarray(order, np.uint32))

    def statuses(self) -> np.ndarray:
        return self.status.download(self.count * STATUS.itemsize).view(STATUS)

    def output(self, i: int) -> bytes:
        o = int(self.dst_off[i])
        return self.dst.download(o + int(self.caps[i]))[o:].tobytes()

    def outputs(self) -> np.ndarray:
        return self.dst.download()


# parse_hdr values (the framing around each stream's Deflate data)
HDR_RAW, HDR_ZLIB, HDR_GZIP, HDR_VERIFY = 0, 1, 2, 0x100
ST_BAD_CHECKSUM = -13


def inflate(b: DeviceBatch, parse_hdr: int = HDR_ZLIB, stream=None) -> None:
    """cuDF ``gpuinflate`` over the batch.  ``parse_hdr`` is HDR_RAW, HDR_ZLIB
    or HDR_GZIP (cuDF's nonzero parse_hdr), optionally | HDR_VERIFY to check
    the Adler-32 / CRC-32 trailer; True means HDR_ZLIB."""
    check(lib().mdz_inflate_batch(b.desc.ptr, b.status.ptr, b.count, int(parse_hdr), as_ptr(stream)),
          "mdz_inflate_batch")


def inflate_sched(b: DeviceBatch, parse_hdr: int = HDR_ZLIB, blocks: int = 0, stream=None) -> None:
    check(lib().mdz_inflate_batch_sched(b.desc.ptr, b.status.ptr, b.count, int(parse_hdr),
                                        b.order.ptr if b.order is not None else None, blocks,
                                        b.cursor.ptr, b.cost.ptr if b.cost is not None else None,
                                        as_ptr(stream)), "mdz_inflate_batch_sched")


ST_LIVE_PENDING = 0x7FFFFFFF


class LiveInflate:
    """A persistent inflate launch that takes streams while it runs -- no
    counterpart in the reference; for serving loops that fold Deflate
    requests into a decode already on the GPU.

    ``submit`` publishes descriptors (``DESC`` records of device pointers) and
    returns the first one's index; ``statuses`` reads the live status array,
    where an item shows ``ST_LIVE_PENDING`` until its output is in memory;
    ``close`` ends the queue and returns the final statuses.  Raw or zlib
    framing.  The workers end on their own after ``idle_ms`` (0: 8 s) without
    their next item; once one has, ``expired`` is True: items already taken
    still finish, later ones may stay pending.
    While a queue is open, hipFree and hipDeviceSynchronize on this device
    wait for its running kernel: allocate first, free after ``close``."""

    def __init__(self, capacity: int, parse_hdr: int = HDR_ZLIB, blocks: int = 0, idle_ms: int = 0):
        q = C.c_void_p()
        check(lib().mdz_inflate_live_open(int(capacity), int(parse_hdr), int(blocks), int(idle_ms),
                                          C.byref(q)), "mdz_inflate_live_open")
        self._q = q
        self.capacity = int(capacity)
        self.published = 0
        ptr = lib().mdz_inflate_live_statuses(q)
        raw = (C.c_uint8 * (self.capacity * STATUS.itemsize)).from_address(ptr)
        self._st = np.frombuffer(raw, np.uint8).view(STATUS)
        self._keep = []                   # buffers the submitted descriptors point into

    def submit(self, desc: np.ndarray, keep=None) -> int:
        d = np.ascontiguousarray(desc, DESC)
        first = C.c_uint32()
        check(lib().mdz_inflate_live_submit(self._q, d.ctypes.data, len(d), C.byref(first)),
              "mdz_inflate_live_submit")
        self.published += len(d)
        if keep is not None:
            self._keep.append(keep)
        return first.value

    def statuses(self) -> np.ndarray:
        """Snapshot of the first ``published`` statuses: each status is read
        before its byte count, so a finished item's count is final."""
        n = self.published
        st = self._st["status"][:n].copy()
        out = np.zeros(n, STATUS)
        out["status"] = st
        out["bytes_written"] = self._st["bytes_written"][:n]
        out["reserved"] = self._st["reserved"][:n]
        return out

    @property
    def expired(self) -> bool:
        return bool(lib().mdz_inflate_live_expired(self._q))

    def close(self) -> np.ndarray:
        if self._q is None:
            raise MdzError(-1, "mdz_inflate_live_close: queue already closed")
        out = np.zeros(max(self.published, 1), STATUS)
        q, self._q = self._q, None
        self._st = None
        rc = lib().mdz_inflate_live_close(q, out.ctypes.data)
        self._keep = []
        check(rc, "mdz_inflate_live_close")
        return out[:self.published]

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if self._q is not None:
            self.close()

    def __del__(self):
        if getattr(self, "_q", None) is not None:
            try:
                self.close()
            except Exception:
                pass


def copy_uncompressed(b: DeviceBatch, stream=None) -> None:
    """cuDF ``gpu_copy_uncompressed_blocks``: the first min(src, dst) bytes of
    each stream go to its output slot unchanged."""
    check(lib().mdz_copy_uncompressed(b.desc.ptr, b.count, as_ptr(stream)), "mdz_copy_uncompressed")


def slots() -> int:
    v = C.c_uint32()
    check(lib().mdz_inflate_slots(C.byref(v)), "mdz_inflate_slots")
    return v.value


def decompress_gpu(data: bytes, col_len: np.ndarray, blk_offset: np.ndarray, chunk_size: int) -> bytes:
    """The repo container: zlib streams, their padded sizes, [size, 0, offsets...]."""
    src = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
    cl = np.ascontiguousarray(col_len, np.uint64)
    bo = np.ascontiguousarray(blk_offset, np.uint64)
    out = C.c_void_p()
    n = C.c_uint64()
    rc = lib().mdz_deflate_decompress_gpu(src.ctypes.data, C.byref(out), len(data), C.byref(n),
                                          cl.ctypes.data, cl.nbytes, bo.ctypes.data, bo.nbytes,
                                          chunk_size)
    if rc not in (0, -5):
        raise MdzError(rc, "mdz_deflate_decompress_gpu")
    res = C.string_at(out.value, n.value) if n.value else b""
    lib().mdz_host_free(out)
    if rc != 0:
        raise MdzError(rc, "mdz_deflate_decompress_gpu (corrupt stream)")
    return res


def cudf_uncompress(data: bytes) -> bytes:
    """The cuDF container (``[u64 chunk][u64 n][u64 size[n]][zlib streams]``,
    ``rapids_compressor.py:222-275``) read as ``base_deflate.h:20-130`` reads
    it: 2-byte zlib header skipped, raw inflate into chunk slots, output
    (n-1)*chunk + the last stream's bytes."""
    src = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
    out = C.c_void_p()
    n = C.c_uint64()
    rc = lib().mdz_cudf_uncompress_host(src.ctypes.data, len(data), C.byref(out), C.byref(n))
    if rc not in (0, -5):
        raise MdzError(rc, "mdz_cudf_uncompress_host")
    res = C.string_at(out.value, n.value) if n.value else b""
    lib().mdz_host_free(out)
    if rc != 0:
        raise MdzError(rc, "mdz_cudf_uncompress_host (corrupt stream)")
    return res
