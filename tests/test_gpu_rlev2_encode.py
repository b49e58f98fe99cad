"""GPU RLE v2 encoder (rlev2::compress_gpu_transpose, include/rlev2/
encoder_trans.h:428-537): payload, col_len and blk_off equal the host
restatement of the reference encoder byte for byte, and decode back."""
import numpy as np
import pytest

from compression_amd import host

pytestmark = pytest.mark.gpu


def _inputs():
    rng = np.random.default_rng(41)
    return {
        "generator_u64": host.gen_rlev2_i64(600_000, 42),
        "random_u32": rng.integers(0, 1 << 32, 300_000, dtype=np.uint64).astype(np.uint32),
        "repeats_u16": np.repeat(rng.integers(0, 60000, 30_000), rng.integers(1, 40, 30_000))[:400_000].astype(np.uint16),
        "ramp_u64": np.cumsum(rng.integers(-50, 51, 200_000)).astype(np.uint64),
        "zeros_u8": np.zeros(300_000, np.uint8),
        "outliers_u64": np.where(rng.random(250_000) < 0.05, rng.integers(0, 1 << 60, 250_000, dtype=np.uint64),
                                 rng.integers(0, 300, 250_000).astype(np.uint64)),
    }


@pytest.mark.parametrize("chunk,ru", [(131072, 512), (131072, 4), (65536, 1), (32768, 16)])
def test_matches_host_encoder(gpu, chunk, ru):
    from compression_amd import rlev2
    from compression_amd._lib import DeviceBuffer
    for name, data in _inputs().items():
        eb = data.dtype.itemsize
        if (chunk // eb // 32) % ru:
            continue
        raw = data.view(np.uint8)
        n, wb, cap = rlev2.compress_bound(raw.size, chunk, eb, ru)
        src = DeviceBuffer.from_array(raw)
        work, pay, pb = DeviceBuffer(wb), DeviceBuffer(cap), DeviceBuffer(8)
        col, off = DeviceBuffer(256 * n), DeviceBuffer(8 * (n + 1))
        print(f"rlev2 encode {name} chunk={chunk} ru={ru} n={n}", flush=True)
        rlev2.compress(src, raw.size, chunk, eb, ru, work, pay, col, off, pb)
        ref = host.rlev2_encode(data, chunk, ru)
        assert np.array_equal(off.download(8 * (n + 1)).view(np.uint64), ref.blk_off), name
        assert np.array_equal(col.download(256 * n).view(np.uint64), ref.col_len), name
        nbytes = int(pb.download(8).view(np.uint64)[0])
        assert nbytes == ref.payload_bytes, name
        assert pay.download(nbytes).tobytes() == ref.payload[:nbytes].tobytes(), name


def test_roundtrip_through_decoder(gpu):
    from compression_amd import rlev2
    from compression_amd._lib import DeviceBuffer
    from compression_amd.lanes import DeviceLanes
    data = host.gen_rlev2_i64(2_000_000, 43)
    raw = data.view(np.uint8)
    n, wb, cap = rlev2.compress_bound(raw.size, 131072, 8, 512)
    src = DeviceBuffer.from_array(raw)
    work, pay, pb = DeviceBuffer(wb), DeviceBuffer(cap), DeviceBuffer(8)
    col, off = DeviceBuffer(256 * n), DeviceBuffer(8 * (n + 1))
    rlev2.compress(src, raw.size, 131072, 8, 512, work, pay, col, off, pb)
    nbytes = int(pb.download(8).view(np.uint64)[0])
    lanes = DeviceLanes(pay.download(nbytes), off.download(8 * (n + 1)).view(np.uint64),
                        col.download(256 * n).view(np.uint64), n, 131072, 8, 512)
    dst, st = DeviceBuffer(raw.size), DeviceBuffer(4 * n)
    rlev2.decompress(lanes, dst, raw.size, st)
    assert (st.download(4 * n).view(np.int32) == 0).all()
    assert dst.download(raw.size).tobytes() == raw.tobytes()
