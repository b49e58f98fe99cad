"""GPU encoders at edge sizes: byte counts that cut the last value (RLE v1
pads bytewise, RLE v2 drops a partial value to zero, as their reference
encoders do), inputs shorter than one chunk, and minimal inputs; each equals
the host restatement of the reference encoder."""
import numpy as np
import pytest

from compression_amd import host

pytestmark = pytest.mark.gpu


def _dev(a):
    from compression_amd._lib import DeviceBuffer
    return DeviceBuffer.from_array(np.ascontiguousarray(a, np.uint8))


@pytest.mark.parametrize("nbytes", [8, 1001, 131072 + 3, 400_001])
def test_rlev1_odd_sizes(gpu, nbytes):
    from compression_amd import rle_v1
    from compression_amd._lib import DeviceBuffer
    raw = host.gen_rlev1_u32(nbytes // 4 + 1, 61).view(np.uint8)[:nbytes].copy()
    n, wb, cap = rle_v1.compress_bound(nbytes, 131072, 4, 128)
    work, pay, col, off = DeviceBuffer(wb), DeviceBuffer(cap), DeviceBuffer(256 * n), DeviceBuffer(8 * (n + 1))
    print(f"rlev1 encode {nbytes} bytes", flush=True)
    rle_v1.compress(_dev(raw), nbytes, 131072, 4, 128, work, pay, col, off)
    # the host encoder takes whole values: zero-extend the input bytewise
    ref = host.rlev1_encode(np.frombuffer(raw.tobytes() + bytes(-nbytes % 4), np.uint32), 131072, 128)
    got_off = off.download(8 * (n + 1)).view(np.uint64)
    assert np.array_equal(got_off, ref.blk_off)
    assert np.array_equal(col.download(256 * n).view(np.uint64), ref.col_len)
    assert pay.download(int(got_off[-1])).tobytes() == ref.payload.tobytes()


@pytest.mark.parametrize("nbytes", [16, 1003, 131072 + 5, 300_007])
def test_rlev2_odd_sizes(gpu, nbytes):
    from compression_amd import rlev2
    from compression_amd._lib import DeviceBuffer
    raw = host.gen_rlev2_i64(nbytes // 8 + 1, 62).view(np.uint8)[:nbytes].copy()
    n, wb, cap = rlev2.compress_bound(nbytes, 131072, 8, 4)
    work, pay, pb = DeviceBuffer(wb), DeviceBuffer(cap), DeviceBuffer(8)
    col, off = DeviceBuffer(256 * n), DeviceBuffer(8 * (n + 1))
    print(f"rlev2 encode {nbytes} bytes", flush=True)
    rlev2.compress(_dev(raw), nbytes, 131072, 8, 4, work, pay, col, off, pb)
    # the partial last value reads as 0, as the reference encoder gathers it
    whole = raw[: nbytes // 8 * 8].tobytes() + (bytes(8) if nbytes % 8 else b"")
    ref = host.rlev2_encode(np.frombuffer(whole, np.uint64), 131072, 4)
    got_off = off.download(8 * (n + 1)).view(np.uint64)
    assert got_off[-1] == nbytes and np.array_equal(got_off[:-1], ref.blk_off[:-1])
    assert np.array_equal(col.download(256 * n).view(np.uint64), ref.col_len)
    nb = int(pb.download(8).view(np.uint64)[0])
    assert nb == ref.payload_bytes and pay.download(nb).tobytes() == ref.payload[:nb].tobytes()


@pytest.mark.parametrize("nbytes", [1, 2, 3, 129, 262144, 262145])
def test_brle_small_sizes(gpu, nbytes):
    from compression_amd import brle
    from compression_amd._lib import DeviceBuffer
    raw = host.gen_brle_u8(nbytes, 63)
    n, wb, cap = brle.compress_bound(nbytes, 262144)
    work, dst, ptr = DeviceBuffer(wb), DeviceBuffer(cap), DeviceBuffer(8 * (n + 1))
    print(f"brle encode {nbytes} bytes", flush=True)
    brle.compress(_dev(raw), nbytes, dst, ptr, work, 262144)
    ref = host.brle_encode(raw, 262144)
    got = ptr.download(8 * (n + 1)).view(np.uint64)
    assert np.array_equal(got, ref.ptr) and dst.download(int(got[-1])).tobytes() == ref.data.tobytes()


def test_lzss_single_chunk_and_partial_tail(gpu):
    from compression_amd import lzss
    from compression_amd._lib import DeviceBuffer
    for nbytes in (65536, 65536 + 100):
        raw = host.gen_lzss_markov(nbytes, 64)
        n, wb, cap = lzss.compress_bound(nbytes, 65536)
        work, pay = DeviceBuffer(wb), DeviceBuffer(cap)
        col, cmap, off = DeviceBuffer(256 * n), DeviceBuffer(32 * n), DeviceBuffer(8 * (n + 1))
        print(f"lzss encode {nbytes} bytes", flush=True)
        lzss.compress(_dev(raw), nbytes, 65536, work, pay, col, cmap, off)
        ref = host.lzss_encode(raw, 65536)
        got = off.download(8 * (n + 1)).view(np.uint64)
        assert n == 1 and np.array_equal(got, ref.blk_off)
        assert np.array_equal(cmap.download(32 * n), ref.col_map)
        assert pay.download(int(got[-1])).tobytes() == ref.payload[:int(got[-1])].tobytes()
