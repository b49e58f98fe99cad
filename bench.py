#!/usr/bin/env python3
"""mdz benchmark: decompressed GB/s and fraction of the HBM roofline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

The default workload, ``all``, measures configs 1-4 in one process (RLE v1
64 MiB; RLE v2 1 GiB + byte RLE 256 MiB through the batched persistent
kernel; LZSS 1 GiB; the 4,096-stream uneven Deflate batch) and then config 5
(mixed-codec serving, p50/p99 at 10-80% of saturation).  A step is one
decode pass over each workload's whole synthetic input, which is resident in
HBM before the timed region starts.  K steps are enqueued back-to-back on
one stream between two HIP events; the device is synchronised on both
sides.  Prints ONE JSON line: the headline over configs 1-4 and one record
per config, each with the roofline of its dominant kernel and its CPU
baseline (all of the process's host cores up to 16, and one core).

Algorithmic bytes (each byte touched once):
  lane-transposed codecs: payload + 256*n (col_len) + 8*(n+1) (blk_off)
                          [+ 32*n col_map for LZSS] + out
  byte RLE:               payload + 8*(n+1) + 12 + out
  Deflate batch:          sum(src) + sum(out) + 48*count
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

# HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default).
# Serving keeps a persistent inflate kernel running on its own stream (the
# live queue); a stream that shared its hardware queue would wait behind it.
# The workloads before it leave a few streams alive, so the process asks for
# more queues before anything starts HIP.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "decompressed GB/s + % HBM roofline per codec; batched vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of each workload's kernel, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel calibration; the
# summaries these come from are under profiles/.
TRAFFIC: dict = {
    "batch_kernel<8> (rle_v2+byte_rle)": 1790353203,   # profiles/config2_rlev2_brle_batched.md
    "rlev2_i64_1GiB_plain": 1422589952,                # profiles/rlev2_i64_1GiB_plain.md
    "brle_u8_256MiB_plain": 536279348,                 # profiles/brle_u8_256MiB_plain.md
    "rlev1_u32_64MiB_plain": 105728819,                # profiles/rlev1_u32_64MiB_plain.md
    "lzss_markov_1GiB_64KiB": 1861173555,              # profiles/lzss_markov_1GiB_64KiB.md
    "lzss_plain_markov_1GiB_64KiB": 1199976756,        # profiles/lzss_plain_markov_1GiB_64KiB.md
    "brle_trans_u8_256MiB": 436797235,                 # profiles/brle_trans_u8_256MiB.md
    "rlev2_plain_i64_1GiB_32KiB": 1616674816,          # profiles/rlev2_plain_i64_1GiB_32KiB.md
    "deflate_4096_streams_uneven": 2818533785,         # profiles/inflate_config4_4096_streams.md
}


def _cpu_time(fn, min_reps=3, max_reps=7, budget_s=20.0):
    ts = []
    t_end = time.perf_counter() + budget_s
    for i in range(max_reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
        if i + 1 >= min_reps and time.perf_counter() > t_end:
            break
    return statistics.median(ts)


# ------------------------------------------------------------- workloads --
class RleV1Workload:
    """RLE v1, plain method: 64 MiB u32 column, 128 KiB chunks, one chunk per
    1024-thread workgroup (16 waves: a lane walker and unit decoders),
    COMP_COL_LEN = line bytes."""
    name = "rlev1_u32_64MiB_plain"
    dtype = "u32"

    def __init__(self, line=128, ring=0, n_vals=1 << 24, seed=0x4D445A01, **_):
        from compression_amd import host as H
        from compression_amd.lanes import DeviceLanes
        import compression_amd as mdz
        self.mdz = mdz
        self.line, self.ring = line, ring
        self.data = H.gen_rlev1_u32(n_vals, seed)
        self.c = H.rlev1_encode(self.data, 131072, line)
        c = self.c
        self.lanes = DeviceLanes(c.payload, c.blk_off, c.col_len, c.n_chunks, 131072, 4, line)
        self.dst = mdz.DeviceBuffer(c.raw_bytes)
        self.status = mdz.DeviceBuffer(4 * c.n_chunks)
        self.out_bytes = c.raw_bytes
        self.max_chunk = int(np.diff(c.blk_off.astype(np.int64)).max())
        n = c.n_chunks
        self.alg_bytes = int(c.blk_off[-1]) + 256 * n + 8 * (n + 1) + c.raw_bytes
        self.config = {"workload": self.name, "codec": "rle_v1", "elem": "u32",
                       "out_bytes": c.raw_bytes, "payload_bytes": int(c.blk_off[-1]),
                       "n_chunks": n, "chunk_bytes": 131072, "comp_col_len": line,
                       "max_chunk_payload": self.max_chunk,
                       "ratio": round(int(c.blk_off[-1]) / c.raw_bytes, 4)}
        if ring:
            self.config["ring_words"] = ring

    def step(self, stream):
        from compression_amd import rle_v1
        rle_v1.decompress(self.lanes, self.dst, self.out_bytes, self.status, stream=stream,
                          ring_words=self.ring, max_chunk_bytes=self.max_chunk)

    def verify(self):
        got = self.dst.download(self.out_bytes)
        st = self.status.download(dtype=np.int32)
        return bool((st == 0).all() and np.array_equal(got, self.data.view(np.uint8)))

    def cpu(self, threads):
        from oracle import pyoracle as O   # the CPU baseline leg
        c = self.c
        out = np.empty(c.raw_bytes, np.uint8)
        t = _cpu_time(lambda: O.rlev1_decode(c, threads=threads))
        return {"value": round(c.raw_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({c.raw_bytes >> 20} MiB out, {c.n_chunks} chunks), "
                          f"oracle C decoder, median of >=3 runs"}


class RleV2Workload:
    """RLE v2, plain method: 1 GiB int64 column (SHORT_REPEAT / DELTA / DIRECT /
    PATCHED_BASE segments 15/35/35/15), 128 KiB chunks, READ_UNIT elements per
    lane line; one workgroup (4 waves) per chunk, each wave decoding whole lane
    streams cooperatively."""
    name = "rlev2_i64_1GiB_plain"
    dtype = "u64"

    def __init__(self, line=512, n_vals=1 << 27, seed=0x4D445A02, **_):
        from compression_amd import host as H
        from compression_amd.lanes import DeviceLanes
        import compression_amd as mdz
        self.mdz = mdz
        self.ru = line
        self.data = H.gen_rlev2_i64(n_vals, seed)
        self.c = c = H.rlev2_encode(self.data, 131072, line)
        self.lanes = DeviceLanes(c.payload, c.blk_off, c.col_len, c.n_chunks, 131072, 8, line)
        self.dst = mdz.DeviceBuffer(c.raw_bytes)
        self.status = mdz.DeviceBuffer(4 * c.n_chunks)
        self.out_bytes = c.raw_bytes
        ends = np.append(c.blk_off[1:-1].astype(np.int64), c.payload_bytes)
        self.max_chunk = int((ends - c.blk_off[:-1].astype(np.int64)).max())
        n = c.n_chunks
        self.alg_bytes = c.payload_bytes + 256 * n + 8 * (n + 1) + c.raw_bytes
        self.config = {"workload": self.name, "codec": "rle_v2", "elem": "u64",
                       "out_bytes": c.raw_bytes, "payload_bytes": c.payload_bytes,
                       "n_chunks": n, "chunk_bytes": 131072, "read_unit": line,
                       "max_chunk_payload": self.max_chunk,
                       "ratio": round(c.payload_bytes / c.raw_bytes, 4)}

    def step(self, stream):
        from compression_amd import rlev2
        rlev2.decompress(self.lanes, self.dst, self.out_bytes, self.status, stream=stream,
                         max_chunk_bytes=self.max_chunk)

    def verify(self):
        got = self.dst.download(self.out_bytes)
        st = self.status.download(dtype=np.int32)
        return bool((st == 0).all() and np.array_equal(got, self.data.view(np.uint8)))

    def cpu(self, threads):
        from oracle import pyoracle as O   # the CPU baseline leg
        c = self.c
        t = _cpu_time(lambda: O.rlev2_decode(c, threads=threads))
        return {"value": round(c.raw_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({c.raw_bytes >> 20} MiB out, {c.n_chunks} chunks), "
                          f"oracle C decoder, median of >=3 runs"}


class RleV2PlainWorkload:
    """RLE v2, plain per-chunk format (include/rlev2/decoder.h:190-262): the
    1 GiB int64 column of the RLE v2 workload with its values masked to
    [0, 2^63) (the plain format's encoder and decoder agree only on those),
    one run stream per 32 KiB chunk (the reference's CHUNK_SIZE); one wave
    per chunk, four per 256-thread workgroup."""
    name = "rlev2_plain_i64_1GiB_32KiB"
    dtype = "int64"

    def __init__(self, n_vals=1 << 27, chunk=32768, seed=0x4D445A02, **_):
        from compression_amd import host as H
        import compression_amd as mdz
        self.mdz = mdz
        self.data = H.gen_rlev2_i64(n_vals, seed).view(np.int64) & np.int64(0x7FFFFFFFFFFFFFFF)
        self.c = c = H.rlev2_plain_encode(self.data, chunk)
        self.chunk = chunk
        self.d_data = mdz.DeviceBuffer.from_array(c.data, pad=16)
        self.d_ptr = mdz.DeviceBuffer.from_array(c.ptr)
        self.dst = mdz.DeviceBuffer(c.raw_bytes)
        self.status = mdz.DeviceBuffer(4 * c.n_chunks)
        self.out_bytes = c.raw_bytes
        nc = c.n_chunks
        self.alg_bytes = int(c.data.size) + 8 * (nc + 1) + c.raw_bytes
        self.config = {"workload": self.name, "codec": "rle_v2", "format": "plain", "elem": "int64",
                       "out_bytes": c.raw_bytes, "payload_bytes": int(c.data.size), "n_chunks": nc,
                       "chunk_bytes": chunk, "ratio": round(c.data.size / c.raw_bytes, 4)}

    def step(self, stream):
        from compression_amd import rlev2
        rlev2.decompress_plain(self.d_data, self.c.data.size, self.d_ptr, self.c.n_chunks, self.chunk,
                               self.dst, self.out_bytes, self.status, stream=stream)

    def verify(self):
        got = self.dst.download(self.out_bytes)
        st = self.status.download(dtype=np.int32)
        return bool((st == 0).all() and np.array_equal(got, self.data.view(np.uint8)))

    def cpu(self, threads):
        from oracle import pyoracle as O   # the CPU baseline leg
        c = self.c
        t = _cpu_time(lambda: O.rlev2_plain_decode(c.blob, c.chunk, threads=threads))
        return {"value": round(c.raw_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({c.raw_bytes >> 20} MiB out, {c.n_chunks} chunks), "
                          f"oracle C decoder (the reference's plain block_decode), median of >=3 runs"}


class BrleWorkload:
    """Byte RLE, plain method: 256 MiB u8 stream of runs (fresh byte, length
    Geom(1/12)), 128 KiB chunks; one 256-thread workgroup per chunk."""
    name = "brle_u8_256MiB_plain"
    dtype = "u8"

    def __init__(self, n=1 << 28, chunk=131072, seed=0x4D445A03, **_):
        from compression_amd import host as H
        from compression_amd import brle
        import compression_amd as mdz
        self.mdz = mdz
        self.data = H.gen_brle_u8(n, seed)
        self.c = c = H.brle_encode(self.data, chunk)
        self.dev = brle.DeviceBrle(c.data, c.ptr, c.n_chunks, chunk)
        self.dst = mdz.DeviceBuffer(c.raw_bytes)
        self.status = mdz.DeviceBuffer(4 * c.n_chunks)
        self.out_bytes = c.raw_bytes
        self.alg_bytes = int(c.ptr[-1]) + 8 * (c.n_chunks + 1) + 12 + c.raw_bytes
        self.config = {"workload": self.name, "codec": "byte_rle", "out_bytes": c.raw_bytes,
                       "payload_bytes": int(c.ptr[-1]), "n_chunks": c.n_chunks, "chunk_bytes": chunk,
                       "max_chunk_payload": self.dev.max_chunk,
                       "ratio": round(int(c.ptr[-1]) / c.raw_bytes, 4)}

    def step(self, stream):
        from compression_amd import brle
        brle.decompress(self.dev, self.dst, self.out_bytes, self.status, stream=stream)

    def verify(self):
        got = self.dst.download(self.out_bytes)
        st = self.status.download(dtype=np.int32)
        return bool((st == 0).all() and np.array_equal(got, self.data))

    def cpu(self, threads):
        from oracle import pyoracle as O   # the CPU baseline leg
        c = self.c
        t = _cpu_time(lambda: O.brle_decode(c, threads=threads))
        return {"value": round(c.raw_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({c.raw_bytes >> 20} MiB out, {c.n_chunks} chunks), "
                          f"oracle C decoder, median of >=3 runs"}


class LzssWorkload:
    """LZSS, lane-transposed: 1 GiB of 64-symbol order-2 Markov text with 40%
    copy events, 64 KiB chunks (2 KiB lanes); a workgroup decodes 4 lanes of a
    chunk, one wave per lane."""
    name = "lzss_markov_1GiB_64KiB"
    dtype = "u8"

    def __init__(self, n=1 << 30, chunk=65536, seed=0x4D445A03, **_):
        from compression_amd import host as H
        from compression_amd.lanes import DeviceLanes
        import compression_amd as mdz
        self.mdz = mdz
        self.data = H.gen_lzss_markov(n, seed)
        self.c = c = H.lzss_encode(self.data, chunk)
        self.lanes = DeviceLanes(c.payload, c.blk_off, c.col_len, c.n_chunks, chunk, 1, 4,
                                 col_map=c.col_map)
        self.dst = mdz.DeviceBuffer(c.raw_bytes)
        self.status = mdz.DeviceBuffer(4 * c.n_chunks)
        self.out_bytes = c.raw_bytes
        nc = c.n_chunks
        self.alg_bytes = c.payload.size + 256 * nc + 8 * (nc + 1) + 32 * nc + c.raw_bytes
        self.config = {"workload": self.name, "codec": "lzss", "out_bytes": c.raw_bytes,
                       "payload_bytes": int(c.payload.size), "n_chunks": nc, "chunk_bytes": chunk,
                       "ratio": round(c.payload.size / c.raw_bytes, 4)}

    def step(self, stream):
        from compression_amd import lzss
        lzss.decompress(self.lanes, self.dst, self.out_bytes, self.status, stream=stream)

    def verify(self):
        got = self.dst.download(self.out_bytes)
        st = self.status.download(dtype=np.int32)
        return bool((st == 0).all() and np.array_equal(got, self.data))

    def cpu(self, threads):
        from oracle import pyoracle as O   # the CPU baseline leg
        c = self.c
        t = _cpu_time(lambda: O.lzss_decode(c, threads=threads))
        return {"value": round(c.raw_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({c.raw_bytes >> 20} MiB out, {c.n_chunks} chunks), "
                          f"oracle C decoder, median of >=3 runs"}


class BrleTransWorkload:
    """Byte RLE, lane-transposed (brle_trans): the 256 MiB byte-RLE column at
    128 KiB chunks in the LZSS lane container; a workgroup decodes 4 lanes of
    a chunk, one wave per lane."""
    name = "brle_trans_u8_256MiB"
    dtype = "u8"

    def __init__(self, n=256 << 20, chunk=131072, seed=0x4D445A02, **_):
        from compression_amd import host as H
        from compression_amd.lanes import DeviceLanes
        import compression_amd as mdz
        self.mdz = mdz
        self.data = H.gen_brle_u8(n, seed)
        self.c = c = H.brle_trans_encode(self.data, chunk)
        self.lanes = DeviceLanes(c.payload, c.blk_off, c.col_len, c.n_chunks, chunk, 1, 4,
                                 col_map=c.col_map)
        self.dst = mdz.DeviceBuffer(c.raw_bytes)
        self.status = mdz.DeviceBuffer(4 * c.n_chunks)
        self.out_bytes = c.raw_bytes
        nc = c.n_chunks
        self.alg_bytes = c.payload.size + 256 * nc + 8 * (nc + 1) + 32 * nc + c.raw_bytes
        self.config = {"workload": self.name, "codec": "brle_trans", "out_bytes": c.raw_bytes,
                       "payload_bytes": int(c.payload.size), "n_chunks": nc, "chunk_bytes": chunk,
                       "ratio": round(c.payload.size / c.raw_bytes, 4)}

    def step(self, stream):
        from compression_amd import brle
        brle.decompress_trans(self.lanes, self.dst, self.out_bytes, self.status, stream=stream)

    def verify(self):
        got = self.dst.download(self.out_bytes)
        st = self.status.download(dtype=np.int32)
        return bool((st == 0).all() and np.array_equal(got, self.data))

    def cpu(self, threads):
        from oracle import pyoracle as O   # the CPU baseline leg
        c = self.c
        t = _cpu_time(lambda: O.brle_trans_decode(c, threads=threads))
        return {"value": round(c.raw_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({c.raw_bytes >> 20} MiB out, {c.n_chunks} chunks), "
                          f"oracle C decoder, median of >=3 runs"}


class LzssPlainWorkload:
    """LZSS, plain per-chunk format: the same 1 GiB Markov+copy stream, each
    64 KiB chunk one token stream (back-references up to 2 KiB kept whole);
    one 512-thread workgroup per chunk."""
    name = "lzss_plain_markov_1GiB_64KiB"
    dtype = "u8"

    def __init__(self, n=1 << 30, chunk=65536, seed=0x4D445A03, **_):
        from compression_amd import host as H
        import compression_amd as mdz
        self.mdz = mdz
        self.data = H.gen_lzss_markov(n, seed)
        self.c = c = H.lzss_plain_encode(self.data, chunk)
        self.chunk = chunk
        self.d_data = mdz.DeviceBuffer.from_array(c.data, pad=16)
        self.d_lens = mdz.DeviceBuffer.from_array(c.lens)
        self.dst = mdz.DeviceBuffer(c.raw_bytes)
        self.status = mdz.DeviceBuffer(4 * c.n_chunks)
        self.out_bytes = c.raw_bytes
        nc = c.n_chunks
        self.alg_bytes = int(c.data.size) + 8 * nc + c.raw_bytes
        self.config = {"workload": self.name, "codec": "lzss", "format": "plain", "out_bytes": c.raw_bytes,
                       "payload_bytes": int(c.data.size), "n_chunks": nc, "chunk_bytes": chunk,
                       "ratio": round(c.data.size / c.raw_bytes, 4)}

    def step(self, stream):
        from compression_amd import lzss
        lzss.decompress_plain(self.d_data, self.c.data.size, self.d_lens, self.c.n_chunks, self.chunk,
                              self.dst, self.out_bytes, self.status, stream=stream)

    def verify(self):
        got = self.dst.download(self.out_bytes)
        st = self.status.download(dtype=np.int32)
        return bool((st == 0).all() and np.array_equal(got, self.data))

    def cpu(self, threads):
        from oracle import pyoracle as O   # the CPU baseline leg
        c = self.c
        t = _cpu_time(lambda: O.lzss_plain_decode(c, threads=threads))
        return {"value": round(c.raw_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({c.raw_bytes >> 20} MiB out, {c.n_chunks} chunks), "
                          f"oracle C decoder (the reference's plain decompress_func loop), median of >=3 runs"}


class Config2Workload:
    """RLE v2 + byte RLE: the 1 GiB int64 column and the 256 MiB byte-RLE
    stream (128 KiB chunks).  Batched (default): both containers' 10,240
    chunks decoded by ONE persistent launch that drains a device work queue
    ordered longest-compressed-chunk first.  Plain: one per-chunk launch per
    codec."""
    dtype = "u64+u8"

    def __init__(self, line=512, batched=True, tuned=True, graphed=True, **_):
        import compression_amd as mdz
        from compression_amd import batch, graph, tune
        self.parts = [RleV2Workload(line=line), BrleWorkload()]
        r, b = self.parts
        self.batched = batched
        self.name = ("rlev2_i64_1GiB+brle_u8_256MiB_" + ("batched" if batched else "plain"))
        self.out_bytes = sum(p.out_bytes for p in self.parts)
        self.alg_bytes = sum(p.alg_bytes for p in self.parts)
        self.tuned = None
        if batched and tuned:
            # launch shape from the autotuner (untimed here; each candidate's
            # decoded bytes are checked against the first's)
            self.batch, res = tune.tune_batch(
                rlev2=r.lanes, rlev2_dst=r.dst, rlev2_dst_bytes=r.out_bytes, rlev2_status=r.status,
                rlev2_max_chunk=r.max_chunk, brle=b.dev, brle_dst=b.dst, brle_dst_bytes=b.out_bytes,
                brle_status=b.status, reps=5)
            ok_ms = [t for _, t in res.candidates if t is not None]
            self.tuned = {"choice": res.choice, "ms": round(res.ms, 5), "from_db": res.from_db,
                          "candidates": len(res.candidates),
                          "slowest_candidate_ms": round(max(ok_ms), 5) if ok_ms else None}
        else:
            self.batch = batch.Batch(r.lanes, r.dst, r.out_bytes, r.status, r.max_chunk,
                                     b.dev, b.dst, b.out_bytes, b.status)
            self.batch.set_order(self.batch.size_order())
        self.graph = None
        if batched and graphed:
            # the pipeline (cursor reset, persistent kernel, status passes)
            # captured once and replayed with one launch per step
            self.gstream = mdz.Stream()
            with graph.capture(self.gstream) as g:
                self.batch.submit(self.gstream.ptr)
            self.graph = g
        launch = "per-codec per-chunk"
        if batched:
            launch = ("hipGraph of the " if self.graph else "") + "persistent work queue" + (
                ", autotuned shape" if self.tuned else " (size order)")
        self.config = {"workload": self.name, "codecs": ["rle_v2", "byte_rle"], "launch": launch,
                       "out_bytes": self.out_bytes, "parts": [p.config for p in self.parts]}
        if self.tuned:
            self.config["tuned"] = self.tuned

    def kernels(self):
        if self.batched:
            return [("batch_kernel<8> (rle_v2+byte_rle)", self.step, self.alg_bytes)]
        return [(p.name, p.step, p.alg_bytes) for p in self.parts]

    def plain_step(self, stream):
        for p in self.parts:
            p.step(stream)

    def step(self, stream):
        if self.graph is not None:
            self.graph.launch(stream)
        elif self.batched:
            self.batch.submit(stream)
        else:
            self.plain_step(stream)

    def verify(self):
        return all(p.verify() for p in self.parts)

    def cpu(self, threads):
        res = [p.cpu(threads) for p in self.parts]
        t = sum(p.out_bytes / (r["value"] * 1e9) for p, r in zip(self.parts, res))
        return {"value": round(self.out_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": "whole workload: oracle C decoders for both codecs, "
                          "median of >=3 runs each; " + "; ".join(r["sample"] for r in res)}


class DeflateWorkload:
    """Deflate: config 4's uneven batch -- 4096 zlib streams, sizes log-uniform
    4 KiB..4 MiB (80% level 6, 10% fixed Huffman, 10% stored), one output slot
    each.  The step is ONE persistent work-queue launch over all streams in
    longest-predicted-cost-first order, the cost from the learned model
    (trained here on a separately generated batch the decoder logged).  The
    plain per-stream launch and the compressed-size order are timed beside
    it."""
    dtype = "u8"
    name = "deflate_4096_streams_uneven"

    def __init__(self, count=4096, **_):
        from compression_amd import deflate, host as H, sched
        self.b = b = H.gen_deflate_batch(count, 0x4D445A04, 4096, 4 << 20, threads=16)
        self.db = deflate.DeviceBatch(b.src, b.src_off, b.src_len, b.raw_len)
        self.out_bytes = int(b.raw_len.sum())
        self.alg_bytes = int(b.src_len.sum()) + self.out_bytes + 48 * count
        # learned order: train on a different batch's decoder-logged costs
        tb = H.gen_deflate_batch(1024, 0x4D445A44, 4096, 4 << 20, threads=16)
        tdb = deflate.DeviceBatch(tb.src, tb.src_off, tb.src_len, tb.raw_len)
        tdb.enable_cost_log()
        deflate.inflate_sched(tdb)
        self.model = sched.train_on_log(tdb, tdb.costs()["cycles"], steps=2000)
        pred = self.model.predict(sched.features(self.db), count)
        self.order_model = sched.lpt_order(pred)
        self.order_size = sched.lpt_order(b.src_len.astype(np.float64))
        self.db.set_order(self.order_size)
        self.size_db = self.db.order
        # order (learned cost / compressed size / index) and grid from the
        # autotuner; the chosen order stays set on the batch
        from compression_amd import graph, tune
        import compression_amd as mdz
        self.blocks, res = tune.tune_inflate(
            self.db, {"learned": self.order_model, "size": self.order_size, "index": None}, reps=2)
        self.order_db = self.db.order
        self.tuned = {"choice": res.choice, "ms": round(res.ms, 3), "from_db": res.from_db,
                      "candidates": [[c, None if t is None else round(t, 3)] for c, t in res.candidates]}
        self.gstream = mdz.Stream()
        with graph.capture(self.gstream) as g:
            deflate.inflate_sched(self.db, blocks=self.blocks, stream=self.gstream.ptr)
        self.graph = g
        self.config = {"workload": self.name, "codec": "deflate", "streams": count,
                       "out_bytes": self.out_bytes, "src_bytes": int(b.src_len.sum()),
                       "modes": {"level6": int((b.mode == 6).sum()), "fixed": int((b.mode == 1).sum()),
                                 "stored": int((b.mode == 0).sum())},
                       "launch": "hipGraph of the persistent work queue; order and grid autotuned "
                                 "over learned-cost LPT, compressed-size LPT and index order",
                       "tuned": self.tuned, "slots": deflate.slots()}

    def step(self, stream):
        self.graph.launch(stream)

    def plain_step(self, stream):
        from compression_amd import deflate
        deflate.inflate(self.db, stream=stream)

    def size_step(self, stream):
        from compression_amd import deflate
        self.db.order = self.size_db
        deflate.inflate_sched(self.db, blocks=self.blocks, stream=stream)
        self.db.order = self.order_db

    batched = True

    def verify(self):
        st = self.db.statuses()
        if not ((st["status"] == 0).all() and (st["bytes_written"] == self.b.raw_len).all()):
            return False
        out, b = self.db.outputs(), self.b
        for i in range(b.count):                      # every stream, byte for byte
            o, ro, n = int(self.db.dst_off[i]), int(b.raw_off[i]), int(b.raw_len[i])
            if not np.array_equal(out[o:o + n], b.raw[ro:ro + n]):
                return False
        return True

    def cpu(self, threads):
        from oracle import pyoracle as O   # the CPU baseline leg: zlib inflate
        b = self.b
        t = _cpu_time(lambda: O.inflate_batch(b, threads=threads))
        return {"value": round(self.out_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "reference",
                "sample": f"whole batch ({self.out_bytes >> 20} MiB out, {b.count} streams), "
                          f"zlib {O.zlib_version()} inflate per stream, median of >=3 runs"}


class CopyWorkload:
    """Calibration: a plain 16-byte-vector device copy of 512 MiB (reads and
    writes each byte once).  Its FETCH_SIZE/WRITE_SIZE per byte scale the
    decoders' counters to HBM bytes."""
    name = "copy_512MiB"
    dtype = "u8"

    def __init__(self, nbytes=512 << 20, **_):
        import compression_amd as mdz
        self.mdz = mdz
        self.src = mdz.DeviceBuffer(nbytes)
        self.src.fill(0x5A)
        self.dst = mdz.DeviceBuffer(nbytes)
        self.out_bytes = nbytes
        self.alg_bytes = 2 * nbytes
        self.config = {"workload": self.name, "bytes": nbytes}

    def step(self, stream):
        self.mdz._lib.check(self.mdz.lib().mdz_dev_copy(self.dst.ptr, self.src.ptr, self.out_bytes,
                                                        stream), "mdz_dev_copy")

    def verify(self):
        return bool((self.dst.download(4096) == 0x5A).all())

    def cpu(self, threads):
        return None


class BrleEncodeWorkload:
    """Byte RLE encode on the GPU (the byte-RLE workload's 256 MiB u8 stream,
    256 KiB chunks as brle::compress_gpu uses): encode, scan, pack."""
    name = "brle_u8_256MiB_encode"
    dtype = "u8"

    def __init__(self, n=1 << 28, chunk=262144, seed=0x4D445A03, **_):
        from compression_amd import brle
        from compression_amd import host as H
        import compression_amd as mdz
        self.brle, self.H, self.chunk = brle, H, chunk
        self.data = H.gen_brle_u8(n, seed)
        self.n, wb, cap = brle.compress_bound(n, chunk)
        self.src = mdz.DeviceBuffer.from_array(self.data)
        self.work, self.dst = mdz.DeviceBuffer(wb), mdz.DeviceBuffer(cap)
        self.ptr = mdz.DeviceBuffer(8 * (self.n + 1))
        self.out_bytes = n                      # input bytes encoded per step
        self.alg_bytes = None
        self.config = {"workload": self.name, "codec": "byte_rle", "direction": "encode",
                       "in_bytes": n, "chunk_bytes": chunk, "n_chunks": self.n}

    def step(self, stream):
        self.brle.compress(self.src, self.out_bytes, self.dst, self.ptr, self.work, self.chunk, stream)

    def verify(self):
        ptr = self.ptr.download(8 * (self.n + 1)).view(np.uint64)
        ref = self.H.brle_encode(self.data, self.chunk)
        self.alg_bytes = int(self.out_bytes + 2 * ptr[-1] + 16 * self.n)
        self.config["out_bytes"] = int(ptr[-1])
        return bool(np.array_equal(ptr, ref.ptr) and
                    self.dst.download(int(ptr[-1])).tobytes() == ref.data.tobytes())

    def kernels(self):
        return [(self.name, self.step, self.alg_bytes)]

    def cpu(self, threads):
        t = _cpu_time(lambda: self.H.brle_encode(self.data, self.chunk, threads=threads))
        return {"value": round(self.out_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({self.out_bytes >> 20} MiB in, {self.n} chunks), host "
                          f"C encoder (the reference encoder's algorithm), median of >=3 runs"}


class RleV1EncodeWorkload:
    """RLE v1 encode on the GPU: the rlev1 workload's 64 MiB u32 column,
    128 KiB chunks, 128-byte lines (lane encode, scan, interleave)."""
    name = "rlev1_u32_64MiB_encode"
    dtype = "u32"

    def __init__(self, n_vals=1 << 24, chunk=131072, line=128, seed=0x4D445A01, **_):
        from compression_amd import host as H
        from compression_amd import rle_v1
        import compression_amd as mdz
        self.rle_v1, self.H, self.chunk, self.line = rle_v1, H, chunk, line
        self.data = H.gen_rlev1_u32(n_vals, seed)
        raw = self.data.view(np.uint8)
        self.n, wb, cap = rle_v1.compress_bound(raw.size, chunk, 4, line)
        self.src = mdz.DeviceBuffer.from_array(raw)
        self.work, self.pay = mdz.DeviceBuffer(wb), mdz.DeviceBuffer(cap)
        self.col, self.off = mdz.DeviceBuffer(256 * self.n), mdz.DeviceBuffer(8 * (self.n + 1))
        self.out_bytes = raw.size               # input bytes encoded per step
        self.alg_bytes = None
        self.config = {"workload": self.name, "codec": "rle_v1", "direction": "encode", "elem": "u32",
                       "in_bytes": raw.size, "chunk_bytes": chunk, "line": line, "n_chunks": self.n}

    def step(self, stream):
        self.rle_v1.compress(self.src, self.out_bytes, self.chunk, 4, self.line, self.work, self.pay,
                             self.col, self.off, stream)

    def verify(self):
        off = self.off.download(8 * (self.n + 1)).view(np.uint64)
        ref = self.H.rlev1_encode(self.data, self.chunk, self.line)
        self.alg_bytes = int(self.out_bytes + 2 * off[-1] + 256 * self.n + 8 * (self.n + 1))
        self.config["payload_bytes"] = int(off[-1])
        return bool(np.array_equal(off, ref.blk_off) and
                    np.array_equal(self.col.download(256 * self.n).view(np.uint64), ref.col_len) and
                    self.pay.download(int(off[-1])).tobytes() == ref.payload.tobytes())

    def kernels(self):
        return [(self.name, self.step, self.alg_bytes)]

    def cpu(self, threads):
        t = _cpu_time(lambda: self.H.rlev1_encode(self.data, self.chunk, self.line, threads=threads))
        return {"value": round(self.out_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({self.out_bytes >> 20} MiB in, {self.n} chunks), host C "
                          f"encoder (the reference encoder's algorithm), median of >=3 runs"}


class RleV2EncodeWorkload:
    """RLE v2 encode on the GPU: the rlev2 workload's 1 GiB int64 column,
    128 KiB chunks, READ_UNIT 512 (lane encode, scan, interleave)."""
    name = "rlev2_i64_1GiB_encode"
    dtype = "u64"

    def __init__(self, line=512, n_vals=1 << 27, chunk=131072, seed=0x4D445A02, **_):
        from compression_amd import host as H
        from compression_amd import rlev2
        import compression_amd as mdz
        self.rlev2, self.H, self.chunk, self.ru = rlev2, H, chunk, line
        self.data = H.gen_rlev2_i64(n_vals, seed)
        raw = self.data.view(np.uint8)
        self.n, wb, cap = rlev2.compress_bound(raw.size, chunk, 8, line)
        self.src = mdz.DeviceBuffer.from_array(raw)
        self.work, self.pay, self.pb = mdz.DeviceBuffer(wb), mdz.DeviceBuffer(cap), mdz.DeviceBuffer(8)
        self.col, self.off = mdz.DeviceBuffer(256 * self.n), mdz.DeviceBuffer(8 * (self.n + 1))
        self.out_bytes = raw.size               # input bytes encoded per step
        self.alg_bytes = None
        self.config = {"workload": self.name, "codec": "rle_v2", "direction": "encode", "elem": "u64",
                       "in_bytes": raw.size, "chunk_bytes": chunk, "read_unit": line, "n_chunks": self.n}

    def step(self, stream):
        self.rlev2.compress(self.src, self.out_bytes, self.chunk, 8, self.ru, self.work, self.pay, self.col,
                            self.off, self.pb, stream)

    def verify(self):
        nbytes = int(self.pb.download(8).view(np.uint64)[0])
        ref = self.H.rlev2_encode(self.data, self.chunk, self.ru)
        self.alg_bytes = int(self.out_bytes + 2 * nbytes + 256 * self.n + 8 * (self.n + 1))
        self.config["payload_bytes"] = nbytes
        return bool(nbytes == ref.payload_bytes and
                    np.array_equal(self.off.download(8 * (self.n + 1)).view(np.uint64), ref.blk_off) and
                    np.array_equal(self.col.download(256 * self.n).view(np.uint64), ref.col_len) and
                    self.pay.download(nbytes).tobytes() == ref.payload[:nbytes].tobytes())

    def kernels(self):
        return [(self.name, self.step, self.alg_bytes)]

    def cpu(self, threads):
        t = _cpu_time(lambda: self.H.rlev2_encode(self.data, self.chunk, self.ru, threads=threads),
                      min_reps=2, max_reps=3)
        return {"value": round(self.out_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({self.out_bytes >> 20} MiB in, {self.n} chunks), host C "
                          f"encoder (the reference encoder's algorithm), median of 2-3 runs"}


class LzssEncodeWorkload:
    """LZSS encode on the GPU: the lzss workload's 1 GiB Markov+copy stream,
    64 KiB chunks (lane matchers, lane sort, scan, interleave)."""
    name = "lzss_1GiB_encode"
    dtype = "u8"

    def __init__(self, n=1 << 30, chunk=65536, seed=0x4D445A03, **_):
        from compression_amd import host as H
        from compression_amd import lzss
        import compression_amd as mdz
        self.lzss, self.H, self.chunk = lzss, H, chunk
        self.data = H.gen_lzss_markov(n, seed)
        self.n, wb, cap = lzss.compress_bound(n, chunk)
        self.src = mdz.DeviceBuffer.from_array(self.data)
        self.work, self.pay = mdz.DeviceBuffer(wb), mdz.DeviceBuffer(cap)
        self.col, self.cmap = mdz.DeviceBuffer(256 * self.n), mdz.DeviceBuffer(32 * self.n)
        self.off = mdz.DeviceBuffer(8 * (self.n + 1))
        self.out_bytes = n                      # input bytes encoded per step
        self.alg_bytes = None
        self.config = {"workload": self.name, "codec": "lzss", "direction": "encode", "in_bytes": n,
                       "chunk_bytes": chunk, "n_chunks": self.n}

    def step(self, stream):
        self.lzss.compress(self.src, self.out_bytes, self.chunk, self.work, self.pay, self.col, self.cmap,
                           self.off, stream)

    def verify(self):
        off = self.off.download(8 * (self.n + 1)).view(np.uint64)
        ref = self.H.lzss_encode(self.data, self.chunk)
        self.alg_bytes = int(self.out_bytes + 2 * off[-1] + 288 * self.n)
        self.config["payload_bytes"] = int(off[-1])
        return bool(np.array_equal(off, ref.blk_off) and
                    np.array_equal(self.col.download(256 * self.n).view(np.uint64), ref.col_len) and
                    np.array_equal(self.cmap.download(32 * self.n), ref.col_map) and
                    self.pay.download(int(off[-1])).tobytes() == ref.payload[:int(off[-1])].tobytes())

    def kernels(self):
        return [(self.name, self.step, self.alg_bytes)]

    def cpu(self, threads):
        t = _cpu_time(lambda: self.H.lzss_encode(self.data, self.chunk, threads=threads), min_reps=2, max_reps=3)
        return {"value": round(self.out_bytes / t / 1e9, 3), "unit": "GB/s", "cores": threads,
                "kind": "port",
                "sample": f"whole workload ({self.out_bytes >> 20} MiB in, {self.n} chunks), host C "
                          f"encoder (hash-chain matcher, same choices as the reference's scan), median of 2-3 runs"}


class InflateVerifyWorkload:
    """Trailer checks on the uneven Deflate batch (4,096 zlib streams, 4 KiB to
    4 MiB): the persistent launch in longest-first order with the Adler-32
    check (MDZ_HDR_VERIFY), timed beside the same launch without it."""
    name = "deflate_4096_streams_verify"
    dtype = "u8"

    def __init__(self, count=4096, **_):
        from compression_amd import deflate
        from compression_amd import host as H
        self.deflate = deflate
        self.b = b = H.gen_deflate_batch(count, 0x4D445A04, 4096, 4 << 20, threads=16)
        self.db = deflate.DeviceBatch(b.src, b.src_off, b.src_len, b.raw_len)
        self.db.set_order(np.argsort(-b.src_len.astype(np.int64), kind="stable").astype(np.uint32))
        self.out_bytes = int(b.raw_len.sum())
        self.alg_bytes = int(b.src_len.sum()) + 2 * self.out_bytes + 48 * count
        self.config = {"workload": self.name, "codec": "deflate", "streams": count, "framing": "zlib+verify",
                       "out_bytes": self.out_bytes, "order": "compressed size, longest first"}

    def step(self, stream):
        self.deflate.inflate_sched(self.db, parse_hdr=self.deflate.HDR_ZLIB | self.deflate.HDR_VERIFY,
                                   stream=stream)

    def plain_framing_step(self, stream):       # the same launch without the check
        self.deflate.inflate_sched(self.db, parse_hdr=self.deflate.HDR_ZLIB, stream=stream)

    def kernels(self):
        return [("inflate_kernel, zlib + Adler-32 check", self.step, self.alg_bytes),
                ("inflate_kernel, zlib, unchecked", self.plain_framing_step, self.alg_bytes)]

    def verify(self):
        st = self.db.statuses()
        return bool((st["status"] == 0).all() and (st["bytes_written"] == self.b.raw_len).all())

    def cpu(self, threads):
        return None


class CopyUncompressedWorkload:
    """cuDF's stored-stream copy: 4,096 streams of 4 KiB to 1 MiB (log-uniform,
    about 744 MB), source offsets 1 byte off the destination's alignment in a
    quarter of them; one launch of mdz_copy_uncompressed."""
    name = "copy_uncompressed_4096_streams"
    dtype = "u8"

    def __init__(self, count=4096, **_):
        from compression_amd import deflate
        self.deflate = deflate
        rng = np.random.default_rng(0x4D445A06)
        sizes = np.exp(rng.uniform(np.log(4096), np.log(1 << 20), count)).astype(np.int64)
        skew = np.where(rng.random(count) < 0.25, 1, 0)
        offs = np.zeros(count, np.int64)
        pos = 0
        for i in range(count):
            pos += (-pos) % 16 + int(skew[i])
            offs[i] = pos
            pos += int(sizes[i])
        src = rng.integers(0, 256, pos + 16).astype(np.uint8)
        self.db = deflate.DeviceBatch(src, offs, sizes, sizes)
        self.src, self.offs, self.sizes = src, offs, sizes
        self.out_bytes = int(sizes.sum())
        self.alg_bytes = 2 * self.out_bytes + 32 * count
        self.config = {"workload": self.name, "streams": count, "bytes": self.out_bytes,
                       "misaligned_streams": int(skew.sum())}

    def step(self, stream):
        self.deflate.copy_uncompressed(self.db, stream=stream)

    def verify(self):
        out = self.db.outputs()
        for i in range(0, len(self.sizes), 97):
            o, n = int(self.db.dst_off[i]), int(self.sizes[i])
            if out[o:o + n].tobytes() != self.src[self.offs[i]:self.offs[i] + n].tobytes():
                return False
        return True

    def cpu(self, threads):
        return None


class FileWorkload:
    """The RLE v2 column of the rlev2 workload (1 GiB int64, 128 KiB chunks)
    decoded from a container file: each step reads the payload from the file
    through pinned staging pieces, copies them over PCIe and decodes each piece
    once it has landed.  The rate includes the file read and PCIe; the file
    sits in the page cache after the first step."""
    name = "rlev2_i64_1GiB_from_file"
    dtype = "u64"

    def __init__(self, line=512, n_vals=1 << 27, seed=0x4D445A02, stage=64 << 20, **_):
        import os
        import tempfile
        from compression_amd import container
        from compression_amd import host as H
        import compression_amd as mdz
        self.data = H.gen_rlev2_i64(n_vals, seed)
        c = H.rlev2_encode(self.data, 131072, line)
        fd, self.path = tempfile.mkstemp(suffix=".mdz")
        os.close(fd)
        container.save(self.path, c)
        self.file = container.File(self.path)
        self.dst = mdz.DeviceBuffer(c.raw_bytes)
        self.status = mdz.DeviceBuffer(4 * c.n_chunks)
        self.stage = stage
        self.out_bytes = c.raw_bytes
        n = c.n_chunks
        self.alg_bytes = c.payload_bytes + 256 * n + 8 * (n + 1) + c.raw_bytes
        self.config = {"workload": self.name, "codec": "rle_v2", "elem": "u64",
                       "out_bytes": c.raw_bytes, "file_payload_bytes": c.payload_bytes,
                       "n_chunks": n, "chunk_bytes": 131072, "read_unit": line,
                       "stage_bytes": stage, "includes": "file read + PCIe H2D + decode"}

    def step(self, stream):
        self.file.decode(self.dst, self.status, stage_bytes=self.stage, stream=stream)

    def verify(self):
        got = self.dst.download(self.out_bytes)
        st = self.status.download(dtype=np.int32)
        return bool((st == 0).all() and np.array_equal(got, self.data.view(np.uint8)))

    def cpu(self, threads):
        return None

    def __del__(self):
        import os
        try:
            self.file.close()
            os.unlink(self.path)
        except Exception:
            pass


WORKLOADS = {"config2": Config2Workload,
             "config2-plain": lambda **kw: Config2Workload(batched=False, **kw),
             "rlev1": RleV1Workload, "rlev2": RleV2Workload,
             "brle": BrleWorkload, "lzss": LzssWorkload, "lzss-plain": LzssPlainWorkload,
             "brle-trans": BrleTransWorkload, "rlev2-plain": RleV2PlainWorkload,
             "deflate": DeflateWorkload,
             "copy": CopyWorkload, "file": FileWorkload, "brle-encode": BrleEncodeWorkload,
             "rlev1-encode": RleV1EncodeWorkload, "rlev2-encode": RleV2EncodeWorkload,
             "lzss-encode": LzssEncodeWorkload, "inflate-verify": InflateVerifyWorkload,
             "copy-uncompressed": CopyUncompressedWorkload}
DEFAULT = "all"


def _log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_protocol(wl, threads: int) -> dict | None:
    """The CPU baseline: the workload's host decoder over `threads` cores (one
    chunk or stream per task, dynamic scheduling), then over one core, with
    the host's affinity core count, CPU model and zlib version."""
    from compression_amd import host as H
    from oracle import pyoracle as O   # the CPU baseline leg
    cb = wl.cpu(threads)
    if cb is None:
        return None
    one = wl.cpu(1)
    cb["one_core"] = {"value": one["value"], "unit": one["unit"], "sample": one["sample"]}
    cb["affinity_cores"] = H.host_cores()
    cb["cpu_model"] = _cpu_model()
    cb["zlib_version"] = O.zlib_version()
    return cb


def _time_steps(fn, stream, steps):
    import compression_amd as mdz
    e0, e1 = mdz.Event(), mdz.Event()
    mdz.lib().mdz_device_sync()
    e0.record(stream)
    for _ in range(steps):
        fn(stream.ptr)
    e1.record(stream)
    e1.sync()
    mdz.lib().mdz_device_sync()
    return e0.elapsed_ms(e1) / steps


def measure(wl, args, stream, threads: int) -> dict:
    """One workload: W warm-up steps, verification, K timed steps between HIP
    events on `stream`, each kernel's own launch time, batched vs plain where
    the workload has both, and the CPU baseline."""
    for _ in range(args.warmup):
        wl.step(stream.ptr)
    stream.sync()
    ok = wl.verify()
    ms = _time_steps(wl.step, stream, args.steps)
    gbs = wl.out_bytes / (ms * 1e-3) / 1e9
    # dominant kernel: each codec's launch timed alone with HIP events on the
    # stream it is enqueued on
    kernels = wl.kernels() if hasattr(wl, "kernels") else [(wl.name, wl.step, wl.alg_bytes)]
    per_kernel = []
    for kname, fn, kbytes in kernels:
        kms = _time_steps(fn, stream, args.steps)
        per_kernel.append({"kernel": kname, "ms": round(kms, 5), "alg_bytes": kbytes,
                           "GBps": round(kbytes / (kms * 1e-3) / 1e9, 1) if kbytes else None})
    dom = max(per_kernel, key=lambda k: k["ms"])
    ach = dom["GBps"]
    rec = {"value": round(gbs, 2), "unit": "GB/s", "ms_per_step": round(ms, 5), "dtype": wl.dtype,
           "config": wl.config}
    if ach is not None:
        rec["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": TRAFFIC.get(dom["kernel"]),
                           "kernel": dom["kernel"], "alg_bytes_per_launch": dom["alg_bytes"]}
    rec["kernels"] = per_kernel
    if getattr(wl, "batched", False):               # batched vs plain, same data, same run
        plain_ms = _time_steps(wl.plain_step, stream, args.steps)
        bvp = {"batched_ms": round(ms, 5), "plain_ms": round(plain_ms, 5),
               "speedup": round(plain_ms / ms, 3)}
        if hasattr(wl, "size_step"):                # compressed-size order, same launch
            bvp["size_order_ms"] = round(_time_steps(wl.size_step, stream, args.steps), 5)
        rec["batched_vs_plain"] = bvp
    rec["bit_exact"] = ok
    if not args.no_cpu:
        cb = cpu_protocol(wl, threads)
        if cb is not None:
            rec["cpu_baseline"] = cb
    return rec


def _header(args, rec: dict) -> dict:
    return {"metric": METRIC, "value": rec["value"], "unit": rec["unit"], "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": rec.get("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": rec["dtype"],
            "data": "synthetic (seeded generators, encoded by the host library)"}


def _threads() -> int:
    from compression_amd import host as H
    return H.default_threads()


def run(args):
    import compression_amd as mdz

    if mdz.device_count() < 1:
        raise SystemExit("bench.py needs a GPU")
    kw = {}
    if args.line:
        kw["line"] = args.line
    if args.ring:
        kw["ring"] = args.ring
    if args.no_tune:
        kw["tuned"] = False
    if args.no_graph:
        kw["graphed"] = False
    wl = WORKLOADS[args.workload](**kw)
    stream = mdz.Stream()
    rec = measure(wl, args, stream, _threads())
    line = _header(args, rec)
    line.update({k: v for k, v in rec.items() if k not in line})
    print(json.dumps(line), flush=True)


def rlev1_line_sweep(args, stream, lines=(32, 128, 512)) -> list:
    """Config 1 at COMP_COL_LEN 32, 128 and 512: kernel time per launch and
    bit-exactness at each line length."""
    out = []
    for ln in lines:
        wl = RleV1Workload(line=ln)
        wl.step(stream.ptr)
        stream.sync()
        ok = wl.verify()
        ms = _time_steps(wl.step, stream, args.steps)
        out.append({"comp_col_len": ln, "ms": round(ms, 5), "payload_bytes": wl.config["payload_bytes"],
                    "alg_GBps": round(wl.alg_bytes / (ms * 1e-3) / 1e9, 1),
                    "frac": round(wl.alg_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "bit_exact": ok})
        del wl
    return out


ALL_CONFIGS = [("config1", "rlev1", "RLE v1, 64 MiB u32 column, plain"),
               ("config2", "config2", "RLE v2 1 GiB int64 + byte RLE 256 MiB, batched persistent"),
               ("config3", "lzss", "LZSS, 1 GiB Markov+copy, 64 KiB chunks, lane-transposed"),
               ("config2-rlev2-plain", "rlev2-plain", "RLE v2 1 GiB int64, 32 KiB chunks, plain format"),
               ("config2-brle-trans", "brle-trans", "byte RLE 256 MiB, 128 KiB chunks, lane-transposed"),
               ("config3-plain", "lzss-plain", "LZSS, 1 GiB Markov+copy, 64 KiB chunks, plain format"),
               ("config4", "deflate", "Deflate, 4,096 uneven zlib streams, scheduled persistent")]


def run_all(args):
    """Default: every config in one process.  Configs 1-4 are decoded with
    their inputs resident in HBM; a step is one decode pass over each of the
    four workloads, so the headline value is their total decompressed bytes
    over the sum of their per-step times.  Config 5 (serving) follows with
    its p50/p99 at 10-80% of saturation.  Each config keeps its own record
    (value, roofline of its dominant kernel, CPU baseline)."""
    import gc

    import compression_amd as mdz

    if mdz.device_count() < 1:
        raise SystemExit("bench.py needs a GPU")
    threads = _threads()
    stream = mdz.Stream()
    recs = {}
    for key, wname, desc in ALL_CONFIGS:
        _log(f"{key}: building {wname} ({desc})")
        wl = WORKLOADS[wname]()
        _log(f"{key}: measuring")
        rec = measure(wl, args, stream, threads)
        rec["description"] = desc
        recs[key] = rec
        _log(f"{key}: {rec['ms_per_step']} ms/step, {rec['value']} GB/s, bit_exact={rec['bit_exact']}")
        del wl
        gc.collect()
    if not args.no_sweep:
        _log("config1: COMP_COL_LEN sweep 32/128/512")
        recs["config1"]["comp_col_len_sweep"] = rlev1_line_sweep(args, stream)
    if not args.no_serve:
        _log(f"config5: serving, {args.requests} requests")
        recs["config5"] = serve_record(args)
        _log(f"config5: p99 at 80% load {recs['config5']['p99_ms_at_80']} ms")
    parts = [recs[k] for k, _, _ in ALL_CONFIGS]
    out_bytes = sum(r["config"]["out_bytes"] for r in parts)
    ms = sum(r["ms_per_step"] for r in parts)
    # headline roofline: the kernel with the most time per step over all configs
    dom = max((k for r in parts for k in r["kernels"]), key=lambda k: k["ms"])
    line = {"metric": METRIC, "value": round(out_bytes / (ms * 1e-3) / 1e9, 2), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "integer (u8/u32/u64 decode)",
            "data": "synthetic (seeded generators, encoded by the host library)",
            "config": {"workload": "configs 1-4 decoded once each per step (RLE v1 64 MiB; RLE v2 1 GiB + "
                                   "byte RLE 256 MiB batched; LZSS 1 GiB lane-transposed and 1 GiB plain; "
                                   "Deflate 4,096 streams), "
                                   "config 5 serving reported beside",
                       "out_bytes": out_bytes, "parts": [k for k, _, _ in ALL_CONFIGS]},
            "roofline": {"bound": "hbm", "achieved": dom["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(dom["GBps"] / HBM_PEAK_GBS, 4), "traffic": TRAFFIC.get(dom["kernel"]),
                         "kernel": dom["kernel"], "alg_bytes_per_launch": dom["alg_bytes"]},
            "bit_exact": all(r["bit_exact"] for r in parts)}
    if all("cpu_baseline" in r for r in parts):
        cpu_s = sum(r["config"]["out_bytes"] / (r["cpu_baseline"]["value"] * 1e9) for r in parts)
        cpu1_s = sum(r["config"]["out_bytes"] / (r["cpu_baseline"]["one_core"]["value"] * 1e9)
                     for r in parts)
        c0 = parts[0]["cpu_baseline"]
        line["cpu_baseline"] = {
            "value": round(out_bytes / cpu_s / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port", "one_core": {"value": round(out_bytes / cpu1_s / 1e9, 3), "unit": "GB/s"},
            "affinity_cores": c0["affinity_cores"], "cpu_model": c0["cpu_model"],
            "zlib_version": c0["zlib_version"],
            "sample": "each config's host decoders (the C restatements of the reference decoders; zlib "
                      "inflate for Deflate) over its whole workload, total bytes over total time"}
    line["configs"] = recs
    print(json.dumps(line), flush=True)


def serve_record(args) -> dict:
    """Config 5: mixed-codec serving.  A pool of requests (codec uniform over
    RLE v1 / RLE v2 / LZSS / Deflate, sizes log-uniform 16 KiB..64 MiB) is
    generated once; saturation is measured with every request arriving at
    t = 0; then open-loop Poisson arrivals at 10..80% of saturation are
    replayed with the same requests.  Latency = arrival -> output ready on
    the device; the 64 GB HBM budget bounds what is admitted."""
    from compression_amd import serve

    n = args.requests
    sizes = serve.request_sizes(n, 0x4D445A05)
    codecs = [serve.CODECS[i] for i in np.random.default_rng(0x4D445A05).integers(0, 4, n)]
    pool = [serve.make_request(i, codecs[i], int(sizes[i]), 0x4D445A05 + i) for i in range(n)]
    total = sum(r.out_bytes for r in pool)
    mode = getattr(args, "serve_deflate", "live")
    native = getattr(args, "serve_loop", "python") == "native" and mode == "live"
    ns = getattr(args, "serve_streams", 0) or (2 if mode == "live" else 4)
    lb = getattr(args, "serve_live_blocks", 0)
    srv = serve.Server(hbm_budget=64 << 30, n_streams=ns, verify=False, deflate_mode=mode, native=native,
                       live_blocks=lb)
    srv.run(pool[: min(8, n)], time_limit=60)            # warm-up: allocations, code objects
    for r in pool:
        r.arrival = 0.0
    sat = srv.run(pool, time_limit=120)
    sat_rps = n / max((total / (sat["GBps"] * 1e9)) if sat["GBps"] else 1.0, 1e-9)
    sweep = []
    for frac in (0.1, 0.2, 0.4, 0.6, 0.8):
        arr = serve.poisson_arrivals(n, frac * sat_rps, 0x4D445A05 + int(frac * 100))
        for r, a in zip(pool, arr):
            r.arrival = float(a)
        res = srv.run(pool, time_limit=120)
        sweep.append({"load": frac, "p50_ms": round(res["p50_ms"], 3), "p99_ms": round(res["p99_ms"], 3),
                      "GBps": round(res["GBps"], 2), "completed": res["completed"]})
    # bit-exactness of the whole pool, once, with verification on
    chk = serve.Server(hbm_budget=64 << 30, n_streams=ns, verify=True, deflate_mode=mode, native=native,
                       live_blocks=lb)
    for r in pool:
        r.arrival = 0.0
    ok = chk.run(pool, time_limit=180)["bit_exact"]
    at80 = sweep[-1]
    return {"value": at80["GBps"], "unit": "GB/s", "dtype": "mixed",
            "config": {"workload": "serve_mixed_codecs", "requests": n, "out_bytes": total,
                       "codecs": {c: codecs.count(c) for c in serve.CODECS}, "hbm_budget_GiB": 64,
                       "streams": ns, "deflate": mode, "loop": "native" if native else "python",
                       "live_blocks": lb or "quarter of the inflate slots"},
            "saturation_GBps": round(sat["GBps"], 2), "p50_ms_at_80": at80["p50_ms"],
            "p99_ms_at_80": at80["p99_ms"], "sweep": sweep, "bit_exact": ok}


def run_serve(args):
    import compression_amd as mdz

    if mdz.device_count() < 1:
        raise SystemExit("bench.py needs a GPU")
    rec = serve_record(args)
    line = {"metric": METRIC, "value": rec["value"], "unit": "GB/s", "n_gpus": 1,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "mixed",
            "data": "synthetic (seeded generators, host encoders)"}
    line.update({k: v for k, v in rec.items() if k not in line})
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS) + ["serve", "all"])
    ap.add_argument("--line", type=int, default=0, help="RLE v1 COMP_COL_LEN override")
    ap.add_argument("--ring", type=int, default=0, help="LDS ring words override")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="all: skip config 5")
    ap.add_argument("--no-sweep", action="store_true", help="all: skip the RLE v1 line sweep")
    ap.add_argument("--requests", type=int, default=64, help="request pool size for serving")
    ap.add_argument("--serve-streams", type=int, default=0,
                    help="serving: HIP streams for the per-request launches (0: 2 with the live "
                         "inflate queue, 4 without)")
    ap.add_argument("--serve-live-blocks", type=int, default=0,
                    help="serving: workgroups of the live inflate kernel (0: a quarter of the slots)")
    ap.add_argument("--serve-loop", default="python", choices=["native", "python"],
                    help="serving: the Python loop (default: measured faster) or the library's "
                         "engine (mdz_serve_*)")
    ap.add_argument("--serve-deflate", default="live", choices=["live", "launch"],
                    help="serving: Deflate requests join one running inflate kernel (live) "
                         "or a new scheduled launch per tick (launch)")
    ap.add_argument("--no-tune", action="store_true",
                    help="config2: size order and default shape instead of the autotuner's choice")
    ap.add_argument("--no-graph", action="store_true",
                    help="config2: submit the batch directly instead of replaying a hipGraph")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "mdz decodes on one GPU; chunks shard across GPUs without "
                                   "exchange, but only --gpus 1 is built and measured"}))
        raise SystemExit(2)
    if args.workload == "serve":
        run_serve(args)
    elif args.workload == "all":
        run_all(args)
    else:
        run(args)


if __name__ == "__main__":
    main()
