"""GPU: the live inflate queue -- streams published while its persistent
kernel runs decode bit-exactly (against zlib, the fixture statuses and the
batch launch on the same streams)."""
import gc
import json
import os
import time
import zlib

import numpy as np
import pytest

from tests import deflate_util as DU

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "deflate_fixtures.json")))


def _batch(streams, caps):
    from compression_amd import deflate
    offs, blob = [], bytearray()
    for s in streams:
        blob += bytes((-len(blob)) % 16)
        offs.append(len(blob))
        blob += s
    return deflate.DeviceBatch(np.frombuffer(bytes(blob) + b"\0", np.uint8), offs,
                               [len(s) for s in streams], caps)


def _sync(mdz):
    """Finishes the null stream's work and frees what earlier tests dropped:
    a hipFree while a queue is open would wait for its running kernel."""
    from compression_amd._lib import check
    gc.collect()
    check(mdz.lib().mdz_stream_sync(None), "mdz_stream_sync")


def _wait_done(q, lo, hi, deadline_s=30.0):
    """Polls the live statuses until items lo..hi-1 are no longer pending."""
    from compression_amd import deflate
    t0 = time.monotonic()
    while True:
        st = q.statuses()
        if (st["status"][lo:hi] != deflate.ST_LIVE_PENDING).all():
            return st
        assert time.monotonic() - t0 < deadline_s, f"items {lo}..{hi} still pending"
        time.sleep(0.001)


def _streams(n, seed):
    rng = np.random.default_rng(seed)
    raws, comp = [], []
    for i in range(n):
        size = int(rng.integers(0, 1 << int(rng.integers(1, 21))))
        alpha = rng.integers(2, 200)
        raw = (rng.integers(0, alpha, size) + (np.arange(size) // 97 % 7)).astype(np.uint8).tobytes()
        level = (6, 1, 9, 0)[i % 4]
        raws.append(raw)
        comp.append(zlib.compress(raw, level))
    return raws, comp


def test_streams_published_in_rounds(gpu):
    from compression_amd import deflate
    raws, comp = _streams(96, 11)
    b = _batch(comp, [len(r) for r in raws])
    b.dst.fill(0xEE)
    _sync(gpu)
    with deflate.LiveInflate(128, deflate.HDR_ZLIB, blocks=16) as q:
        assert not q.expired
        for r in range(4):
            first = q.submit(b.host_desc[24 * r:24 * (r + 1)])
            assert first == 24 * r
            st = _wait_done(q, 24 * r, 24 * (r + 1))
            assert (st["status"][:24 * (r + 1)] == 0).all()
            time.sleep(0.05)                     # the workers go back to waiting
        final = q.close()
    assert np.array_equal(final, st)
    out = b.outputs()
    for i, raw in enumerate(raws):
        o = int(b.dst_off[i])
        assert int(final["bytes_written"][i]) == len(raw)
        assert out[o:o + len(raw)].tobytes() == raw, i


def test_single_items_and_whole_batch_match_batch_launch(gpu):
    """Raw Deflate, damaged and whole streams, one submit per item for half
    of them: statuses, byte counts and bytes equal the batch launch's."""
    from compression_amd import deflate
    raws, comp = _streams(64, 12)
    rng = np.random.default_rng(3)
    streams = []
    for i, c in enumerate(comp):
        s = bytearray(c[2:-4])
        if i % 3 == 1 and len(s) > 8:
            for _ in range(4):
                s[int(rng.integers(0, len(s)))] ^= 1 << int(rng.integers(0, 8))
        streams.append(bytes(s))
    caps = [len(r) for r in raws]
    ref = _batch(streams, caps)
    ref.dst.fill(0xEE)
    deflate.inflate_sched(ref, parse_hdr=deflate.HDR_RAW)
    ref_st = ref.statuses()
    ref_out = ref.outputs()
    b = _batch(streams, caps)
    b.dst.fill(0xEE)
    _sync(gpu)
    q = deflate.LiveInflate(64, deflate.HDR_RAW)
    for i in range(32):
        q.submit(b.host_desc[i:i + 1])
    q.submit(b.host_desc[32:])
    _wait_done(q, 0, 64)
    final = q.close()
    assert np.array_equal(final["status"], ref_st["status"])
    assert np.array_equal(final["bytes_written"], ref_st["bytes_written"])
    assert (ref_st["status"] != 0).any() and (ref_st["status"] == 0).any()
    out = b.outputs()
    for i in range(64):
        o, n = int(b.dst_off[i]), int(ref_st["bytes_written"][i])
        assert np.array_equal(out[o:o + n], ref_out[o:o + n]), i
        if i % 3 != 1:                           # undamaged: the round trip holds
            assert ref_st["status"][i] == 0 and out[o:o + n].tobytes() == raws[i]


def test_fixture_statuses(gpu):
    from compression_amd import deflate
    vs, fs = FIX["valid"], FIX["faults"]
    streams = [bytes.fromhex(v["stream"]) for v in vs] + [bytes.fromhex(f["stream"]) for f in fs]
    caps = [len(v["raw"]) // 2 for v in vs] + [f["cap"] for f in fs]
    b = _batch(streams, caps)
    _sync(gpu)
    with deflate.LiveInflate(len(streams), deflate.HDR_ZLIB, blocks=4) as q:
        q.submit(b.host_desc)
        _wait_done(q, 0, len(streams))
        final = q.close()
    out = b.outputs()
    for i, v in enumerate(vs):
        o = int(b.dst_off[i])
        assert final["status"][i] == 0
        assert out[o:o + int(final["bytes_written"][i])].tobytes().hex() == v["raw"], v["name"]
    for f, s in zip(fs, final["status"][len(vs):]):
        assert s == f["status"], f["name"]


def test_queue_limits(gpu):
    from compression_amd import deflate
    with pytest.raises(gpu.MdzError):
        deflate.LiveInflate(8, deflate.HDR_GZIP)
    with pytest.raises(gpu.MdzError):
        deflate.LiveInflate(0)
    raws, comp = _streams(6, 13)
    b = _batch(comp, [len(r) for r in raws])
    _sync(gpu)
    q = deflate.LiveInflate(4, deflate.HDR_ZLIB, blocks=2)
    q.submit(b.host_desc[:3])
    with pytest.raises(gpu.MdzError):
        q.submit(b.host_desc[3:])                # 3 more past a capacity of 4
    assert q.published == 3
    q.submit(b.host_desc[3:4])
    _wait_done(q, 0, 4)
    final = q.close()
    assert (final["status"] == 0).all() and len(final) == 4
    with pytest.raises(gpu.MdzError):
        q.close()
    empty = deflate.LiveInflate(16)              # closed with nothing published
    assert len(empty.close()) == 0


def test_batch_launch_runs_beside_open_queue(gpu):
    """A full-grid batch launch on another stream completes while the queue's
    workers wait, and the queue still takes items afterwards."""
    from compression_amd import deflate
    raws, comp = _streams(40, 14)
    live = _batch(comp[:8], [len(r) for r in raws[:8]])
    other = _batch(comp[8:], [len(r) for r in raws[8:]])
    _sync(gpu)
    with deflate.LiveInflate(16, deflate.HDR_ZLIB) as q:
        q.submit(live.host_desc[:4])
        _wait_done(q, 0, 4)
        deflate.inflate_sched(other, parse_hdr=deflate.HDR_ZLIB)
        st = other.statuses()                    # waits for that launch only
        assert (st["status"] == 0).all()
        q.submit(live.host_desc[4:])
        _wait_done(q, 4, 8)
        final = q.close()
    assert (final["status"] == 0).all()
    for b, rr in ((live, raws[:8]), (other, raws[8:])):
        out = b.outputs()
        for i, raw in enumerate(rr):
            o = int(b.dst_off[i])
            assert out[o:o + len(raw)].tobytes() == raw


def test_idle_workers_expire_and_a_new_queue_takes_the_items(gpu):
    from compression_amd import deflate
    raws, comp = _streams(4, 15)
    b = _batch(comp, [len(r) for r in raws])
    _sync(gpu)
    q = deflate.LiveInflate(8, deflate.HDR_ZLIB, blocks=2, idle_ms=20)
    t0 = time.monotonic()
    while not q.expired:
        assert time.monotonic() - t0 < 10.0, "idle workers did not end"
        time.sleep(0.01)
    q.submit(b.host_desc)
    time.sleep(0.1)
    assert (q.statuses()["status"] == deflate.ST_LIVE_PENDING).all()
    assert (q.close()["status"] == deflate.ST_LIVE_PENDING).all()
    with deflate.LiveInflate(8, deflate.HDR_ZLIB, blocks=2, idle_ms=2000) as q2:
        q2.submit(b.host_desc)
        _wait_done(q2, 0, 4)
        final = q2.close()
    assert (final["status"] == 0).all()
    out = b.outputs()
    for i, raw in enumerate(raws):
        o = int(b.dst_off[i])
        assert out[o:o + len(raw)].tobytes() == raw
