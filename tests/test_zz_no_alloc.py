"""GPU: the no-allocation persistent inflate entry point
(mdz_inflate_batch_sched_noalloc, cursor in the library's own device memory)
gives mdz_inflate_batch_sched's bytes and statuses, through the C ABI, on two
calls in a row, the second on a smaller batch."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _both(gpu, b, order):
    from compression_amd import deflate
    from compression_amd._lib import lib
    ref = deflate.DeviceBatch(b.src, b.src_off, b.src_len, b.raw_len)
    ref.set_order(order)
    ref.dst.fill(0xA5)
    deflate.inflate_sched(ref)
    new = deflate.DeviceBatch(b.src, b.src_off, b.src_len, b.raw_len)
    new.set_order(order)
    new.dst.fill(0x5A)
    new.status.fill(0x7F)
    rc = lib().mdz_inflate_batch_sched_noalloc(new.desc.ptr, new.status.ptr, new.count,
                                               deflate.HDR_ZLIB, new.order.ptr, 0, None, None)
    assert rc == 0
    gpu.lib().mdz_device_sync()
    sr, sn = ref.statuses(), new.statuses()
    assert np.array_equal(sr["status"], sn["status"]) and (sn["status"] == 0).all()
    assert np.array_equal(sr["bytes_written"], sn["bytes_written"])
    assert (sn["bytes_written"] == b.raw_len).all()
    out_r, out_n = ref.outputs(), new.outputs()
    for i in range(b.count):
        o, n = int(new.dst_off[i]), int(b.raw_len[i])
        assert np.array_equal(out_n[o:o + n], out_r[int(ref.dst_off[i]):int(ref.dst_off[i]) + n])
        assert out_n[o:o + n].tobytes() == b.raw_stream(i)


def test_noalloc_matches_sched_twice(gpu):
    big = O.gen_deflate_batch(200, 0x4D445A21, 4096, 1 << 20)
    _both(gpu, big, np.argsort(-big.src_len.astype(np.int64)).astype(np.uint32))
    small = O.gen_deflate_batch(37, 0x4D445A22, 4096, 1 << 18)   # the cursor must restart at 0
    _both(gpu, small, np.arange(small.count, dtype=np.uint32)[::-1].copy())


def test_noalloc_rejects_bad_arguments(gpu):
    from compression_amd._lib import lib
    assert lib().mdz_inflate_batch_sched_noalloc(None, None, -1, 1, None, 0, None, None) == -1
    assert lib().mdz_inflate_batch_sched_noalloc(None, None, 0, 3, None, 0, None, None) == -1
    assert lib().mdz_inflate_batch_sched_noalloc(None, None, 0, 1, None, 0, None, None) == 0
