"""GPU byte-RLE encoder (brle::compress_gpu, include/brle/brle.h:232-303):
chunk data and ptr[] equal the host restatement of the reference encoder
(brle.h:129-230) byte for byte, and decode back to the input."""
import numpy as np
import pytest

from compression_amd import host

pytestmark = pytest.mark.gpu


def _patterns():
    rng = np.random.default_rng(21)
    out = {"random": rng.integers(0, 256, 300_000).astype(np.uint8),
           "one_byte": np.full(400_000, 7, np.uint8),
           "generator": host.gen_brle_u8(1_500_000, 22),
           "tiny": np.array([1], np.uint8),
           "two": np.array([5, 5], np.uint8),
           "three": np.array([5, 5, 5], np.uint8)}
    # runs of every length 1..140 between distinct bytes
    segs = []
    for L in range(1, 141):
        segs += [L % 251 + 1] * L + [0]
    out["run_lengths"] = np.array(segs, np.uint8)
    # literal stretches of 120..136 distinct bytes, each followed by a run of 1..5
    segs = []
    for k in range(120, 137):
        segs += [(i * 7 + k) % 255 + 1 for i in range(k)] + [0] * (k % 5 + 1)
    out["literal_limits"] = np.array(segs, np.uint8)
    # pairs and triples: "aab", "aaab" ...
    segs = []
    for i in range(20000):
        segs += [i % 200 + 1] * int(rng.integers(1, 5)) + [201 + i % 50]
    out["short_runs"] = np.array(segs, np.uint8)
    return out


@pytest.mark.parametrize("chunk", [262144, 4097, 1000, 131072])
def test_matches_host_encoder(gpu, chunk):
    from compression_amd import brle
    from compression_amd._lib import DeviceBuffer
    for name, data in _patterns().items():
        n, wb, cap = brle.compress_bound(data.size, chunk)
        src = DeviceBuffer.from_array(data)
        work, dst, ptr = DeviceBuffer(wb), DeviceBuffer(cap), DeviceBuffer(8 * (n + 1))
        dst.fill(0xEE)
        print(f"brle encode {name} chunk={chunk} n={n}", flush=True)
        brle.compress(src, data.size, dst, ptr, work, chunk_bytes=chunk)
        got_ptr = ptr.download(8 * (n + 1)).view(np.uint64)
        ref = host.brle_encode(data, chunk)
        assert np.array_equal(got_ptr, ref.ptr), name
        got = dst.download(int(got_ptr[-1]))
        assert got.tobytes() == ref.data.tobytes(), name


def test_reference_container_and_roundtrip(gpu):
    from compression_amd import brle
    data = host.gen_brle_u8(5_000_000, 23)
    blob = brle.compress_gpu(data.tobytes())
    assert blob == host.brle_encode(data, 262144).to_bytes()
    assert brle.decompress_gpu(blob) == data.tobytes()
