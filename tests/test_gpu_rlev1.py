"""GPU parity: RLE v1 decode through the C ABI vs the CPU oracle (bit-exact)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _roundtrip(mdz, data: np.ndarray, chunk: int, line: int, ring: int = 0, dst_bytes=None,
               max_lane_tokens: int = 0):
    from compression_amd import rle_v1
    from compression_amd.lanes import DeviceLanes
    c = O.rlev1_encode(data, chunk, line, threads=8)
    lanes = DeviceLanes(c.payload, c.blk_off, c.col_len, c.n_chunks, chunk, c.elem_bytes, line)
    n = c.raw_bytes if dst_bytes is None else dst_bytes
    dst = mdz.DeviceBuffer(max(n, 1))
    dst.fill(0xAB)
    st = mdz.DeviceBuffer(4 * c.n_chunks)
    rle_v1.decompress(lanes, dst, n, st, ring_words=ring, max_lane_tokens=max_lane_tokens)
    got = dst.download(n)
    status = st.download(4 * c.n_chunks, np.int32)
    rc, ref, ref_st = O.rlev1_decode(c, out_bytes=n, threads=8)
    return c, got, status, ref, ref_st


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("line", [16, 32, 128])
def test_dtypes_and_lines(gpu, dtype, line):
    x = O.gen_rlev1_u32(300_000, 11).astype(dtype)
    if dtype == np.uint64:
        x = x * np.uint64(0x10001) + np.uint64(1 << 40)
    c, got, st, ref, _ = _roundtrip(gpu, x, 32768, line)
    assert (st == 0).all()
    assert np.array_equal(got, ref) and np.array_equal(got, x.view(np.uint8))


@pytest.mark.parametrize("ring", [256, 1024, 4096, 32768])
def test_window_sliding(gpu, ring):
    # a small LDS ring forces many window slides per chunk
    x = O.gen_rlev1_u32(1 << 20, 0x4D445A01)
    c, got, st, ref, _ = _roundtrip(gpu, x, 131072, 32, ring=ring)
    assert (st == 0).all() and np.array_equal(got, x.view(np.uint8))


@pytest.mark.parametrize("line", [8, 4])
def test_narrow_lines(gpu, line):
    x = O.gen_rlev1_u32(200_000, 5)
    c, got, st, ref, _ = _roundtrip(gpu, x, 65536, line)
    assert (st == 0).all() and np.array_equal(got, x.view(np.uint8))


def test_ragged_tail_and_clipped_output(gpu):
    x = O.gen_rlev1_u32(100_003, 9)                       # not a chunk multiple
    c, got, st, ref, _ = _roundtrip(gpu, x, 16384, 32)
    assert np.array_equal(got, x.view(np.uint8))
    c, got, st, ref, _ = _roundtrip(gpu, x, 16384, 32, dst_bytes=12345)   # mid-line clip
    assert np.array_equal(got, x.view(np.uint8)[:12345])


def test_incompressible_and_constant(gpu):
    rng = np.random.default_rng(1)
    x = rng.integers(0, 1 << 32, 1 << 18, dtype=np.uint64).astype(np.uint32)  # 5-byte varints
    c, got, st, ref, _ = _roundtrip(gpu, x, 131072, 64, ring=8192)
    assert c.payload.size > x.nbytes                       # expands; still decodes
    assert (st == 0).all() and np.array_equal(got, x.view(np.uint8))
    z = np.full(1 << 18, 77, np.uint32)
    c, got, st, ref, _ = _roundtrip(gpu, z, 131072, 32)
    assert (st == 0).all() and np.array_equal(got, z.view(np.uint8))


def test_config1_full_size(gpu):
    x = O.gen_rlev1_u32(1 << 24, 0x4D445A01)               # 64 MiB, 512 chunks
    c, got, st, ref, _ = _roundtrip(gpu, x, 131072, 32)
    assert c.n_chunks == 512 and (st == 0).all()
    assert np.array_equal(got, x.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("max_tokens", [1, 6])
def test_checkpoint_path(gpu, dtype, max_tokens):
    # token lists too short for the lanes: the chunk is walked again with a
    # checkpoint per value group (1: below the batch size, so never listed)
    x = O.gen_rlev1_u32(300_000, 13).astype(dtype)
    c, got, st, ref, _ = _roundtrip(gpu, x, 32768, 32, max_lane_tokens=max_tokens)
    assert (st == 0).all() and np.array_equal(got, x.view(np.uint8))


@pytest.mark.parametrize("max_tokens", [0, 6])
def test_corrupt_streams_report_status(gpu, max_tokens):
    from compression_amd import rle_v1
    from compression_amd.lanes import DeviceLanes
    x = O.gen_rlev1_u32(1 << 16, 3)
    c = O.rlev1_encode(x, 32768, 32, threads=4)
    col = c.col_len.copy()
    col[5] = max(1, col[5] // 2)          # chunk 0 lane 5 truncated -> input overrun
    pay = c.payload.copy()
    lo = int(c.blk_off[1])
    pay[lo:lo + 4] = 0x7F                 # chunk 1: first tokens become 130-value runs
    lanes = DeviceLanes(pay, c.blk_off, col, c.n_chunks, 32768, 4, 32)
    dst = gpu.DeviceBuffer(c.raw_bytes)
    st = gpu.DeviceBuffer(4 * c.n_chunks)
    rle_v1.decompress(lanes, dst, c.raw_bytes, st, max_lane_tokens=max_tokens)
    s = st.download(4 * c.n_chunks, np.int32)
    assert s[0] != 0 and s[1] != 0 and (s[2:] == 0).all()
    ref = O.RleV1Container(payload=pay, col_len=col, blk_off=c.blk_off, n_chunks=c.n_chunks,
                           chunk=32768, elem_bytes=4, line=32, raw_bytes=c.raw_bytes)
    _, _, ref_st = O.rlev1_decode(ref, threads=4)
    assert np.array_equal(s, ref_st)          # same status codes as the oracle
    assert np.array_equal(dst.download()[2 * 32768:], x.view(np.uint8)[2 * 32768:])


@pytest.mark.parametrize("dtype", [np.uint8, np.uint32, np.uint64])
@pytest.mark.parametrize("max_tokens", [0, 6])
def test_random_corruption_matches_oracle(gpu, dtype, max_tokens):
    # byte damage in most chunks: every chunk's status equals the oracle's,
    # and undamaged chunks decode bit-exactly
    from compression_amd import rle_v1
    from compression_amd.lanes import DeviceLanes
    x = O.gen_rlev1_u32(1 << 18, 17).astype(dtype)
    chunk = 16384
    c = O.rlev1_encode(x, chunk, 32, threads=4)
    pay = c.payload.copy()
    rng = np.random.default_rng(23)
    damaged = set()
    for ch in range(c.n_chunks):
        if ch % 4 == 0:
            continue
        lo, hi = int(c.blk_off[ch]), int(c.blk_off[ch + 1])
        for _ in range(1 + ch % 3):
            pos = int(rng.integers(lo, hi))
            pay[pos] = (int(pay[pos]) ^ int(rng.integers(1, 256))) & 0xFF
        damaged.add(ch)
    lanes = DeviceLanes(pay, c.blk_off, c.col_len, c.n_chunks, chunk, c.elem_bytes, 32)
    dst = gpu.DeviceBuffer(c.raw_bytes)
    st = gpu.DeviceBuffer(4 * c.n_chunks)
    rle_v1.decompress(lanes, dst, c.raw_bytes, st, max_lane_tokens=max_tokens)
    s = st.download(4 * c.n_chunks, np.int32)
    ref = O.RleV1Container(payload=pay, col_len=c.col_len, blk_off=c.blk_off, n_chunks=c.n_chunks,
                           chunk=chunk, elem_bytes=c.elem_bytes, line=32, raw_bytes=c.raw_bytes)
    _, ref_out, ref_st = O.rlev1_decode(ref, threads=4)
    assert np.array_equal(s, ref_st)
    assert (ref_st[sorted(damaged)] != 0).sum() > len(damaged) // 2      # most damage is seen
    got = dst.download()
    for ch in range(c.n_chunks):
        if ref_st[ch] == 0:
            assert np.array_equal(got[ch * chunk:(ch + 1) * chunk], ref_out[ch * chunk:(ch + 1) * chunk]), ch


def test_host_wrappers(gpu, tmp_path):
    from compression_amd import rle_v1
    x = O.gen_rlev1_u32(70_000, 21)
    c = O.rlev1_encode(x, 32768, 32, threads=4)
    file_bytes = np.uint64(c.raw_bytes).tobytes() + c.payload.tobytes()
    out = rle_v1.decompress_gpu(file_bytes, 32, 32768, 4, c.col_len, c.blk_off)
    assert out == x.tobytes()
    # the side-file form reads ./input_data/ like the reference (rlev1.h:110-147)
    import ctypes as C
    from compression_amd._lib import lib
    os.makedirs(tmp_path / "input_data")
    c.col_len.astype("<u8").tofile(tmp_path / "input_data" / "rle_v1_col_len.bin")
    c.blk_off.astype("<u8").tofile(tmp_path / "input_data" / "rle_v1_blk_offset.bin")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        src = np.frombuffer(file_bytes, np.uint8)
        o, n = C.c_void_p(), C.c_uint64()
        rc = lib().mdz_rle_v1_decompress_gpu(src.ctypes.data, C.byref(o), src.size, C.byref(n), 32, 32768, 4)
        assert rc == 0 and C.string_at(o.value, n.value) == x.tobytes()
        lib().mdz_host_free(o)
    finally:
        os.chdir(cwd)
