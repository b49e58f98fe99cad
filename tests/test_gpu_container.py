"""GPU: container files decoded through the staged reader (mmap -> pinned
pieces -> HBM, decode per piece) equal the source data, for every codec and
for pieces from one chunk per piece to the whole payload at once."""
import zlib

import numpy as np
import pytest

from compression_amd import container, host
from compression_amd._lib import DeviceBuffer
from tests.deflate_util import gzip_member

pytestmark = pytest.mark.gpu


def _decode(path, stage, out_bytes, status_bytes):
    with container.File(path) as f:
        dst = DeviceBuffer(max(out_bytes, 1))
        st = DeviceBuffer(max(status_bytes, 4))
        dst.fill(0xEE)
        st.fill(0x55)
        print(f"file decode {f.codec} stage={stage}", flush=True)
        f.decode(dst, st, stage_bytes=stage)
        f.decode(dst, st, stage_bytes=stage)         # the handle's device copies are reused
        return dst.download(out_bytes), st.download(status_bytes)


def _cases():
    return {
        "rlev1": (host.gen_rlev1_u32(600_000, 11).view(np.uint8), lambda d: host.rlev1_encode(d.view(np.uint32), 131072, 128)),
        "rlev2": (host.gen_rlev2_i64(300_000, 12).view(np.uint8), lambda d: host.rlev2_encode(d.view(np.uint64), 131072, 4)),
        "brle": (host.gen_brle_u8(3_000_000, 13), lambda d: host.brle_encode(d, 262144)),
        "lzss": (host.gen_lzss_markov(2_000_000, 14), lambda d: host.lzss_encode(d, 65536)),
    }


@pytest.mark.parametrize("codec", ["rlev1", "rlev2", "brle", "lzss"])
@pytest.mark.parametrize("stage", [1, 200_000, 0])
def test_chunked_codecs(gpu, tmp_path, codec, stage):
    data, enc = _cases()[codec]
    c = enc(data)
    p = str(tmp_path / f"{codec}.mdz")
    container.save(p, c)
    out, st = _decode(p, stage, c.raw_bytes, 4 * c.n_chunks)
    assert (st.view(np.int32) == 0).all()
    assert out.tobytes() == data.tobytes()[:c.raw_bytes]


@pytest.mark.parametrize("framing", [0x101, 0x102, 1])
def test_deflate_streams(gpu, tmp_path, framing):
    from compression_amd import deflate
    rng = np.random.default_rng(framing)
    datas = [host.gen_markov(int(n), int(n) + 1, 0.4).tobytes() for n in rng.integers(1, 400_000, 40)]
    datas += [b"", bytes(70_000)]
    if framing & 3 == 2:
        streams = [gzip_member(d, (0, 1, 6, 9)[i % 4], name=b"s%d" % i)[0] for i, d in enumerate(datas)]
    else:
        streams = [zlib.compress(d, (0, 1, 6, 9)[i % 4]) for i, d in enumerate(datas)]
    p = str(tmp_path / "d.mdz")
    container.save_deflate(p, streams, [len(d) for d in datas], framing=framing)
    for stage in (1, 1 << 20, 0):
        out, st = _decode(p, stage, sum(len(d) for d in datas), deflate.STATUS.itemsize * len(datas))
        st = st.view(deflate.STATUS)
        assert (st["status"] == 0).all()
        assert list(st["bytes_written"]) == [len(d) for d in datas]
        assert out.tobytes() == b"".join(datas)
