"""Container files (MDZF) on the CPU: write, open and check without a device.
Header fields round-trip, and damaged files are refused when opened: header
CRC, magic, sections past the end, inconsistent tables."""
import os
import struct

import numpy as np
import pytest

from compression_amd import container, host
from compression_amd._lib import MdzError


@pytest.fixture
def rlev1_file(tmp_path):
    c = host.rlev1_encode(host.gen_rlev1_u32(100_000, 3), 32768, 128)
    p = str(tmp_path / "a.mdz")
    container.save(p, c)
    return p, c


def test_roundtrip_info(rlev1_file):
    p, c = rlev1_file
    with container.File(p) as f:
        i = f.info
        assert f.codec == "rlev1"
        assert (i.elem_bytes, i.lane_line, i.n_chunks, i.chunk_bytes, i.raw_bytes, i.payload_bytes) == \
            (4, 128, c.n_chunks, 32768, c.raw_bytes, c.payload.size)
    raw = open(p, "rb").read()
    assert raw[:4] == b"MDZF" and len(raw) % 4096 == c.payload.size % 4096
    pay_at = struct.unpack_from("<Q", raw, 32 + 8 * 8)[0]
    assert pay_at % 4096 == 0 and raw[pay_at:pay_at + c.payload.size] == c.payload.tobytes()


def test_every_codec_writes(tmp_path):
    data = host.gen_brle_u8(300_000, 5)
    host_cs = {"rlev2": host.rlev2_encode(host.gen_rlev2_i64(40_000, 7), 65536, 4),
               "brle": host.brle_encode(data, 65536),
               "lzss": host.lzss_encode(host.gen_lzss_markov(200_000, 9), 65536)}
    for name, c in host_cs.items():
        p = str(tmp_path / f"{name}.mdz")
        container.save(p, c)
        with container.File(p) as f:
            assert f.codec == name and f.info.n_chunks == c.n_chunks and f.info.raw_bytes == c.raw_bytes
    import zlib
    streams = [zlib.compress(bytes([k]) * (1000 * k + 1)) for k in range(1, 6)]
    p = str(tmp_path / "d.mdz")
    container.save_deflate(p, streams, [1000 * k + 1 for k in range(1, 6)], framing=0x101)
    with container.File(p) as f:
        assert f.codec == "deflate" and f.info.framing == 0x101
        assert list(f.raw_len()) == [1000 * k + 1 for k in range(1, 6)]


def _damaged(p, at, val, fmt="<B"):
    b = bytearray(open(p, "rb").read())
    struct.pack_into(fmt, b, at, val)
    q = p + ".bad"
    open(q, "wb").write(bytes(b))
    return q


def _refused(path):
    with pytest.raises(MdzError) as e:
        container.File(path)
    return e.value.rc


def test_damaged_files_are_refused(rlev1_file):
    p, c = rlev1_file
    assert _refused(_damaged(p, 0, ord("X"))) == -1             # magic
    assert _refused(_damaged(p, 40, 7, "<Q")) == -5             # chunk_bytes changed: header CRC
    raw = open(p, "rb").read()
    short = p + ".short"
    open(short, "wb").write(raw[:-1])                            # payload cut
    assert _refused(short) == -1
    # blk_off is outside the header's CRC: a decreasing table is caught by the table check
    boff_at = struct.unpack_from("<Q", raw, 32 + 4 * 8)[0]
    assert _refused(_damaged(p, boff_at + 8, 1 << 40, "<Q")) == -1
    open(short, "wb").write(raw[:100])
    assert _refused(short) == -1
    assert _refused(str(os.path.dirname(p)) + "/missing.mdz") == -1


def test_writer_checks_tables(tmp_path):
    p = str(tmp_path / "x.mdz")
    pay = np.zeros(64, np.uint8)
    with pytest.raises(MdzError):                                # blk_off past the payload
        container.write(p, "brle", pay, [0, 100], 1, 10, 256)
    with pytest.raises(MdzError):                                # raw bytes beyond n * chunk
        container.write(p, "brle", pay, [0, 64], 1, 1000, 256)
    with pytest.raises(MdzError):                                # RLE v2 last entry is the raw size
        container.write(p, "rlev2", pay, [0, 5], 1, 6, 256, 8, 1, col_len=np.zeros(32, np.uint64))
    with pytest.raises(MdzError):                                # Deflate sizes must sum to raw_bytes
        container.write(p, "deflate", pay, [0, 64], 1, 5, raw_len=[4], framing=1)
    with pytest.raises(MdzError):                                # unknown framing
        container.write(p, "deflate", pay, [0, 64], 1, 4, raw_len=[4], framing=3)
    container.write(p, "brle", pay, [0, 64], 1, 100, 256)
    assert container.File(p).info.raw_bytes == 100
