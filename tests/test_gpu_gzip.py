"""GPU parity for cuDF's gzip framing (include/cudf_deflate.cuh:1179-1217,
1244-1253), the trailer checks (MDZ_HDR_VERIFY) and the batched stored-stream
copy (copy_uncompressed_kernel, :1349-1397), through the C ABI against the
oracle and zlib."""
import zlib

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.deflate_util import GZIP_FIELD_CASES, gzip_member

pytestmark = pytest.mark.gpu

ZLIB, GZIP, VERIFY = 1, 2, 0x100


def _batch(streams, caps, src_skew=0, dst_skew=0):
    """Device batch of host streams; dst_skew shifts every output slot start."""
    from compression_amd import deflate
    from compression_amd._lib import DeviceBuffer
    offs, blob = [], bytearray()
    for s in streams:
        blob += bytes((-len(blob)) % 16 + src_skew)
        offs.append(len(blob))
        blob += s
    b = deflate.DeviceBatch(np.frombuffer(bytes(blob) + b"\0", np.uint8), offs,
                            [len(s) for s in streams], [c + dst_skew for c in caps])
    if dst_skew:
        d = b.desc.download(b.count * deflate.DESC.itemsize).view(deflate.DESC).copy()
        d["dst"] += dst_skew
        d["dst_bytes"] -= dst_skew
        b.desc = DeviceBuffer.from_array(d)
        b.dst_off = b.dst_off + dst_skew
        b.caps = b.caps - dst_skew
    return b


def _run(streams, caps, hdr, sched=False, src_skew=0, dst_skew=0):
    from compression_amd import deflate
    b = _batch(streams, caps, src_skew, dst_skew)
    b.dst.fill(0xEE)
    print(f"inflate hdr={hdr:#x} sched={sched} n={len(streams)} skew={src_skew}/{dst_skew}", flush=True)
    if sched:
        b.set_order(np.arange(b.count, dtype=np.uint32)[::-1].copy())
        deflate.inflate_sched(b, parse_hdr=hdr)
    else:
        deflate.inflate(b, parse_hdr=hdr)
    st = b.statuses()
    out = b.outputs()
    return [out[int(b.dst_off[i]): int(b.dst_off[i]) + int(st["bytes_written"][i])].tobytes()
            for i in range(b.count)], st


def _datas(seed):
    rng = np.random.default_rng(seed)
    out = [b"", b"q", bytes(1023), bytes(1024), bytes(range(256)) * 4 + b"!",
           rng.integers(0, 256, 3 << 20).astype(np.uint8).tobytes()]
    out += [rng.integers(0, int(k), int(n)).astype(np.uint8).tobytes()
            for n, k in zip(rng.integers(1, 90_000, 12), rng.integers(2, 256, 12))]
    return out


@pytest.mark.parametrize("sched", [False, True])
@pytest.mark.parametrize("verify", [0, VERIFY])
def test_gzip_members(gpu, sched, verify):
    streams, datas = [], []
    for i, fields in enumerate(GZIP_FIELD_CASES):
        for d in _datas(i)[:8]:
            streams.append(gzip_member(d, level=(0, 1, 6, 9)[len(streams) % 4], **fields)[0])
            datas.append(d)
    outs, st = _run(streams, [len(d) for d in datas], GZIP | verify, sched=sched)
    assert (st["status"] == 0).all()
    for o, d in zip(outs, datas):
        assert o == d


def test_rejected_gzip_headers(gpu):
    s, _ = gzip_member(b"")
    n, _ = gzip_member(b"x" * 100, name=b"abc")
    bad = [s[:18], s[:17], b"\x1e" + s[1:], s[:2] + b"\x07" + s[3:], n[:10] + b"y" * 30]
    good = [s, n]
    outs, st = _run(bad + good, [200] * (len(bad) + len(good)), GZIP)
    for i, x in enumerate(bad + good):
        want = O.inflate(x, 200, hdr=GZIP)
        assert (st["status"][i], outs[i]) == want, i
    assert (st["status"][:len(bad)] == -1).all() and (st["bytes_written"][:len(bad)] == 0).all()


@pytest.mark.parametrize("dst_skew", [0, 1, 3])
def test_trailer_checks(gpu, dst_skew):
    streams, caps, hdrs = [], [], []
    for i, d in enumerate(_datas(40 + dst_skew)):
        level = (0, 1, 6, 9)[i % 4]
        z = zlib.compress(d, level)
        g = gzip_member(d, level)[0]
        cases = [z, z + b"\0" * 7, z[:-1]]
        for k in (1, 4):
            b = bytearray(z); b[-k] ^= 0x40; cases.append(bytes(b))
        if level == 0 and len(d) > 200:
            b = bytearray(z); b[150] ^= 0xFF; cases.append(bytes(b))
        streams += cases
        hdrs += [ZLIB] * len(cases)
        gcases = [g]
        for k in (1, 5, 8):
            b = bytearray(g); b[-k] ^= 0x01; gcases.append(bytes(b))
        streams += gcases
        hdrs += [GZIP] * len(gcases)
        caps += [len(d)] * (len(cases) + len(gcases))
    for hdr in (ZLIB, GZIP):
        idx = [i for i, h in enumerate(hdrs) if h == hdr]
        ss, cc = [streams[i] for i in idx], [caps[i] for i in idx]
        outs, st = _run(ss, cc, hdr | VERIFY, dst_skew=dst_skew)
        for j, (s, c) in enumerate(zip(ss, cc)):
            assert (st["status"][j], outs[j]) == O.inflate(s, c, hdr=hdr | VERIFY), (hdr, j)
        assert (st["status"] == -13).any() and (st["status"] == 0).any()


def test_random_corruption_with_verify(gpu):
    rng = np.random.default_rng(77)
    for hdr in (ZLIB, GZIP):
        streams, caps = [], []
        for n, lv in zip(rng.integers(50, 30_000, 30), rng.choice([0, 1, 6, 9], 30)):
            d = rng.integers(0, 12, int(n)).astype(np.uint8).tobytes()
            s = zlib.compress(d, int(lv)) if hdr == ZLIB else gzip_member(d, int(lv))[0]
            for _ in range(4):
                b = bytearray(s)
                b[int(rng.integers(len(b) // 2, len(b)))] ^= 1 << int(rng.integers(0, 8))
                streams.append(bytes(b))
                caps.append(len(d) + 64)
        _, st = _run(streams, caps, hdr | VERIFY)
        for s, c, g in zip(streams, caps, st["status"]):
            assert g == O.inflate(s, c, hdr=hdr | VERIFY)[0]


def test_bad_framing_is_rejected(gpu):
    from compression_amd import deflate
    from compression_amd._lib import MdzError
    b = _batch([zlib.compress(b"abc")], [3])
    for bad in (3, VERIFY, 0x200):
        with pytest.raises(MdzError):
            deflate.inflate(b, parse_hdr=bad)


def test_gzip_cost_features(gpu):
    from compression_amd import sched
    datas = _datas(9)[:10]
    zs = [zlib.compress(d, (0, 1, 6)[i % 3]) for i, d in enumerate(datas)]
    gs = [gzip_member(d, (0, 1, 6)[i % 3], name=b"f" * i)[0] for i, d in enumerate(datas)]
    F = sched.F
    xz = sched.features(_batch(zs, [len(d) for d in datas]), parse_hdr=ZLIB).download(
        len(datas) * 4 * F).view(np.float32).reshape(-1, F)
    xg = sched.features(_batch(gs, [len(d) for d in datas]), parse_hdr=GZIP).download(
        len(datas) * 4 * F).view(np.float32).reshape(-1, F)
    # the first block's type and what its code lengths imply (match share,
    # bits per symbol) do not depend on the framing; the size-derived
    # estimates do
    cols = [3, 4, 5, 7, 10]
    assert np.array_equal(xz[:, cols], xg[:, cols]) and (xz[:, 3:6].sum(1) == 1).all()


@pytest.mark.parametrize("src_skew,dst_skew", [(0, 0), (4, 0), (1, 3), (5, 5), (0, 7)])
def test_copy_uncompressed(gpu, src_skew, dst_skew):
    from compression_amd import deflate
    rng = np.random.default_rng(src_skew * 8 + dst_skew)
    sizes = [0, 1, 15, 16, 17, 4095, 65_537] + [int(n) for n in rng.integers(1, 3 << 20, 9)]
    blobs = [rng.integers(0, 256, n).astype(np.uint8).tobytes() for n in sizes]
    caps = [n if i % 3 else max(0, n - int(rng.integers(0, 40))) for i, n in enumerate(sizes)]
    b = _batch(blobs, caps, src_skew, dst_skew)
    b.dst.fill(0xEE)
    print(f"copy_uncompressed skew={src_skew}/{dst_skew}", flush=True)
    deflate.copy_uncompressed(b)
    out = b.outputs()
    for i, (s, c) in enumerate(zip(blobs, caps)):
        o = int(b.dst_off[i])
        n = min(len(s), c)
        assert out[o:o + n].tobytes() == s[:n], i
        assert (out[o + n:o + c] == 0xEE).all(), i
