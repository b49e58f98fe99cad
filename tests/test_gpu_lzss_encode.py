"""GPU LZSS encoder (lzss::compress_gpu, include/lzss/lzss_gpu2.h:1952-2123):
payload, sorted col_len, col_map and blk_off equal the host restatement of
the reference encoder byte for byte, and decode back."""
import numpy as np
import pytest

from compression_amd import host

pytestmark = pytest.mark.gpu


def _inputs():
    rng = np.random.default_rng(51)
    return {
        "markov": host.gen_lzss_markov(3_000_000, 52),
        "random": rng.integers(0, 256, 1_000_000).astype(np.uint8),
        "low_alphabet": rng.integers(0, 3, 1_000_000).astype(np.uint8),
        "periodic": (np.arange(1_500_000) // 7 % 256).astype(np.uint8),
        "zeros": np.zeros(700_000, np.uint8),
    }


@pytest.mark.parametrize("chunk", [131072, 65536, 4096])
def test_matches_host_encoder(gpu, chunk):
    from compression_amd import lzss
    from compression_amd._lib import DeviceBuffer
    for name, data in _inputs().items():
        n, wb, cap = lzss.compress_bound(data.size, chunk)
        src = DeviceBuffer.from_array(data)
        work, pay = DeviceBuffer(wb), DeviceBuffer(cap)
        col, cmap, off = DeviceBuffer(256 * n), DeviceBuffer(32 * n), DeviceBuffer(8 * (n + 1))
        print(f"lzss encode {name} chunk={chunk} n={n}", flush=True)
        lzss.compress(src, data.size, chunk, work, pay, col, cmap, off)
        ref = host.lzss_encode(data, chunk)
        got_off = off.download(8 * (n + 1)).view(np.uint64)
        assert np.array_equal(got_off, ref.blk_off), name
        assert np.array_equal(col.download(256 * n).view(np.uint64), ref.col_len), name
        assert np.array_equal(cmap.download(32 * n), ref.col_map), name
        assert pay.download(int(got_off[-1])).tobytes() == ref.payload[:int(got_off[-1])].tobytes(), name


def test_host_form_roundtrip(gpu):
    from compression_amd import lzss
    data = host.gen_lzss_markov(5 * 131072 + 1000, 53)
    payload, blobs = lzss.compress_gpu(data.tobytes())
    ref = host.lzss_encode(data, 131072)
    assert payload == ref.payload[:int(ref.blk_off[-1])].tobytes()
    assert blobs[0] == np.uint64(131072).tobytes() and blobs[1] == ref.blk_off.tobytes()
    assert blobs[2] == ref.col_len.tobytes() and blobs[3] == ref.col_map.tobytes()
    assert lzss.decompress_gpu(payload, blobs) == data[:5 * 131072].tobytes()
