"""Deflate stream builders for tests (RFC 1951): a bit writer, fixed and
dynamic block writers, and crafted streams that hit each zlib error class."""
from __future__ import annotations

LEN_BASE = [3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83, 99,
            115, 131, 163, 195, 227, 258]
LEN_EXTRA = [0] * 8 + [1] * 4 + [2] * 4 + [3] * 4 + [4] * 4 + [5] * 4 + [0]
DIST_BASE = [1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769, 1025,
             1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577]
DIST_EXTRA = [0, 0, 0, 0] + [i // 2 for i in range(2, 28)]
CL_ORDER = [16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15]


class BitWriter:
    def __init__(self):
        self.bits: list[int] = []

    def put(self, v: int, n: int):            # fields: LSB first
        self.bits += [(v >> i) & 1 for i in range(n)]

    def code(self, c: int, n: int):           # Huffman codes: MSB first
        self.bits += [(c >> i) & 1 for i in reversed(range(n))]

    def align(self):
        self.bits += [0] * (-len(self.bits) % 8)

    def tobytes(self) -> bytes:
        b = self.bits + [0] * (-len(self.bits) % 8)
        return bytes(sum(b[i + k] << k for k in range(8)) for i in range(0, len(b), 8))


def canonical(lengths: list[int]) -> dict[int, tuple[int, int]]:
    """symbol -> (code, length) for a canonical Huffman code (RFC 1951 3.2.2)."""
    mx = max(lengths, default=0)
    cnt = [0] * (mx + 2)
    for x in lengths:
        if x:
            cnt[x] += 1
    nxt, code = [0] * (mx + 2), 0
    for b in range(1, mx + 1):
        code = (code + cnt[b - 1]) << 1
        nxt[b] = code
    out = {}
    for s, x in enumerate(lengths):
        if x:
            out[s] = (nxt[x], x)
            nxt[x] += 1
    return out


FIXED_LIT = [8] * 144 + [9] * 112 + [7] * 24 + [8] * 8
FIXED_DIST = [5] * 32


def len_sym(n: int) -> tuple[int, int, int]:
    for i in range(28, -1, -1):
        if LEN_BASE[i] <= n:
            return 257 + i, n - LEN_BASE[i], LEN_EXTRA[i]
    raise ValueError(n)


def dist_sym(d: int) -> tuple[int, int, int]:
    for i in range(29, -1, -1):
        if DIST_BASE[i] <= d:
            return i, d - DIST_BASE[i], DIST_EXTRA[i]
    raise ValueError(d)


def write_symbols(w: BitWriter, lit, dist, items):
    """items: ('L', byte) | ('M', length, distance) | ('S', raw litlen symbol) |
    ('D', length, raw distance symbol) | ('E',) end of block."""
    for it in items:
        if it[0] == "L":
            w.code(*lit[it[1]])
        elif it[0] == "S":
            w.code(*lit[it[1]])
        elif it[0] == "E":
            w.code(*lit[256])
        else:
            s, e, ne = len_sym(it[1])
            w.code(*lit[s])
            w.put(e, ne)
            if it[0] == "M":
                ds, de, dne = dist_sym(it[2])
                w.code(*dist[ds])
                w.put(de, dne)
            else:
                w.code(*dist[it[2]])


def fixed_block(w: BitWriter, items, final=True):
    w.put(int(final), 1)
    w.put(1, 2)
    write_symbols(w, canonical(FIXED_LIT), canonical(FIXED_DIST), items)


def dynamic_header(w: BitWriter, lit_lens, dist_lens, final=True, cl_lens=None, seq=None,
                   hlit=None, hdist=None):
    """Writes a dynamic block header.  The code-length sequence is written
    verbatim (literal lengths 0..15 only) unless `seq` gives CL symbols
    [(sym, extra_value)]; cl_lens overrides the code-length code."""
    w.put(int(final), 1)
    w.put(2, 2)
    nlen, ndist = len(lit_lens), len(dist_lens)
    w.put((hlit if hlit is not None else nlen - 257), 5)
    w.put((hdist if hdist is not None else ndist - 1), 5)
    if seq is None:
        seq = [(x, 0) for x in list(lit_lens) + list(dist_lens)]
    if cl_lens is None:
        used = sorted({s for s, _ in seq})
        if len(used) == 1:                      # a complete code needs two symbols
            used.append(1 if used[0] != 1 else 0)
        cl_lens = [0] * 19
        for s, x in zip(used, kraft_complete(len(used))):
            cl_lens[s] = x
    ncode = 19
    while ncode > 4 and cl_lens[CL_ORDER[ncode - 1]] == 0:
        ncode -= 1
    w.put(ncode - 4, 4)
    for i in range(ncode):
        w.put(cl_lens[CL_ORDER[i]], 3)
    cl = canonical(cl_lens)
    extra_bits = {16: 2, 17: 3, 18: 7}
    for s, e in seq:
        w.code(*cl[s])
        if s in extra_bits:
            w.put(e, extra_bits[s])


def zlib_wrap(raw: bytes) -> bytes:
    return b"\x78\x01" + raw


def kraft_complete(k: int) -> list[int]:
    """Lengths of a complete prefix code over k >= 2 symbols."""
    depth = (k - 1).bit_length()
    lens = [depth] * k
    spare = (1 << depth) - k
    for i in range(spare):
        lens[i] -= 1
    # merging: two depth-long slots become one (depth-1); repeat until complete
    while sum(2.0 ** -x for x in lens) < 1:
        for i in range(k):
            if sum(2.0 ** -x for x in lens) + 2.0 ** -lens[i] <= 1:
                lens[i] -= 1
                break
    return lens


def crafted_faults() -> list[tuple[str, bytes, int, int]]:
    """(name, zlib stream, output cap, expected status)."""
    cases = []
    cases.append(("bad_header", b"\x78\x9d\x03\x00", 16, -11))
    cases.append(("bad_method", b"\x79\x01\x03\x00", 16, -11))
    cases.append(("block_type_3", zlib_wrap(b"\x07"), 16, -1))
    cases.append(("stored_nlen", zlib_wrap(b"\x01\x05\x00\x00\x00abcde"), 16, -2))
    w = BitWriter()
    fixed_block(w, [("L", 97), ("S", 286)])
    cases.append(("fixed_sym_286", zlib_wrap(w.tobytes()), 16, -10))
    w = BitWriter()
    fixed_block(w, [("L", 97), ("D", 3, 30)])
    cases.append(("fixed_dist_30", zlib_wrap(w.tobytes()), 16, -9))
    w = BitWriter()
    fixed_block(w, [("L", 97), ("M", 3, 5), ("E",)])
    cases.append(("too_far_back", zlib_wrap(w.tobytes()), 16, -9))
    w = BitWriter()
    fixed_block(w, [("L", 97)] * 20 + [("E",)])
    s = zlib_wrap(w.tobytes())
    cases.append(("output_full", s, 10, 1))
    cases.append(("truncated", s[:6], 64, 2))
    # dynamic headers
    lit = [0] * 257
    lit[97], lit[256] = 1, 1
    w = BitWriter()
    dynamic_header(w, lit, [1], hlit=30)
    cases.append(("hlit_287", zlib_wrap(w.tobytes()), 16, -7))
    w = BitWriter()
    cl = [0] * 19
    cl[0], cl[1] = 2, 2                               # incomplete code-length code
    dynamic_header(w, lit, [1], cl_lens=cl, seq=[(1, 0)] * 258)
    cases.append(("cl_incomplete", zlib_wrap(w.tobytes()), 16, -3))
    w = BitWriter()
    cl = [0] * 19
    cl[0], cl[1], cl[16] = 1, 2, 2
    dynamic_header(w, lit, [1], cl_lens=cl, seq=[(16, 0)] + [(0, 0)] * 255)
    cases.append(("repeat_first", zlib_wrap(w.tobytes()), 16, -4))
    lit2 = [0] * 257
    lit2[97], lit2[98] = 1, 1                          # no end-of-block code
    w = BitWriter()
    dynamic_header(w, lit2, [1])
    cases.append(("no_eob", zlib_wrap(w.tobytes()), 16, -8))
    lit3 = [0] * 257
    lit3[97], lit3[98], lit3[256] = 1, 1, 1            # over-subscribed
    w = BitWriter()
    dynamic_header(w, lit3, [1])
    cases.append(("litlen_oversub", zlib_wrap(w.tobytes()), 16, -5))
    w = BitWriter()
    dynamic_header(w, lit, [1, 1, 1])
    cases.append(("dist_oversub", zlib_wrap(w.tobytes()), 16, -6))
    # single distance code of length 1: the unused pattern is invalid
    lit4 = [0] * 258
    lit4[97], lit4[256], lit4[257] = 2, 2, 1
    w = BitWriter()
    dynamic_header(w, lit4, [1])
    litc, distc = canonical(lit4), {0: (1, 1)}          # code 1 is the undefined pattern
    write_symbols(w, litc, distc, [("L", 97), ("D", 3, 0)])
    cases.append(("dist_single_unused", zlib_wrap(w.tobytes()), 16, -9))
    return cases


def gzip_member(data: bytes, level: int = 6, extra: bytes | None = None, name: bytes | None = None,
                comment: bytes | None = None, fhcrc: bool = False) -> tuple[bytes, int]:
    """An RFC 1952 member built field by field: (stream, header length).  The
    Deflate data comes from zlib (raw), the trailer from zlib's crc32."""
    import struct
    import zlib
    flg = (0x02 if fhcrc else 0) | (0x04 if extra is not None else 0) | \
          (0x08 if name is not None else 0) | (0x10 if comment is not None else 0)
    h = bytearray(b"\x1f\x8b\x08" + bytes([flg]) + struct.pack("<I", 0x5EED5EED) + b"\x00\x03")
    if extra is not None:
        h += struct.pack("<H", len(extra)) + extra
    if name is not None:
        h += name + b"\x00"
    if comment is not None:
        h += comment + b"\x00"
    if fhcrc:
        h += struct.pack("<H", zlib.crc32(bytes(h)) & 0xFFFF)
    c = zlib.compressobj(level, zlib.DEFLATED, -15)
    body = c.compress(data) + c.flush()
    return bytes(h) + body + struct.pack("<II", zlib.crc32(data), len(data) & 0xFFFFFFFF), len(h)


GZIP_FIELD_CASES = [
    {},
    {"name": b"column.bin"},
    {"extra": b"AB\x04\x00xyzw"},
    {"comment": b"a comment"},
    {"fhcrc": True},
    {"extra": b"", "name": b"n", "comment": b"", "fhcrc": True},
    {"extra": bytes(300), "name": b"long-name" * 20},
]


def orig_container(data: bytes, chunk: int = 128 * 1024):
    """The repo Deflate container as orig_compressor.py:44-125 writes it:
    the input zero-padded to a chunk multiple (:62-66), each chunk zlib
    level 9 (:87), each stream zero-padded to 8 bytes (:102-109),
    col_len = padded sizes (:91,122), blk_offset = [len(data), 0, cumsum]
    (:57-58,115-121).  Returns (payload, col_len u64, blk_offset u64)."""
    import zlib

    import numpy as np
    n = -(-len(data) // chunk)
    padded = data + bytes(n * chunk - len(data))
    streams = []
    for p in range(n):
        z = zlib.compress(padded[p * chunk:(p + 1) * chunk], 9)
        streams.append(z + bytes(-len(z) % 8))
    sizes = [len(s) for s in streams]
    blk = np.array([len(data), 0] + list(np.cumsum(sizes, dtype=np.uint64)), np.uint64)
    return b"".join(streams), np.array(sizes, np.uint64), blk


def cudf_container(data: bytes, chunk: int = 4096 * 2 * 8) -> bytes:
    """The cuDF container as rapids_compressor.py:222-275 writes it:
    [u64 chunk][u64 n][u64 len(zlib stream)[n]][streams back to back], the
    input zero-padded to a chunk multiple, level 9, no per-stream padding."""
    import struct
    import zlib
    n = -(-len(data) // chunk)
    padded = data + bytes(n * chunk - len(data))
    streams = [zlib.compress(padded[p * chunk:(p + 1) * chunk], 9) for p in range(n)]
    return struct.pack(f"<QQ{n}Q", chunk, n, *[len(s) for s in streams]) + b"".join(streams)
