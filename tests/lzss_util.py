"""Independent LZSS helpers for the tests: a token-stream decoder, a container
builder from arbitrary lane streams and a container reader, written from the
format description (include/lzss/lzss_gpu2.h:114-325,497-509,1063-1085), not
from the oracle."""
from __future__ import annotations

import numpy as np


def py_lane_decode(s: bytes, out_len: int):
    """Returns (status, bytes): 0 ok, 1 output overrun, 2 input overrun, -9 bad distance."""
    out = bytearray()
    p = 0
    while len(out) < out_len:
        if p >= len(s):
            return 2, bytes(out)
        f = s[p]
        p += 1
        for b in range(8):
            if len(out) >= out_len:
                break
            if f >> b & 1:
                if p >= len(s):
                    return 2, bytes(out)
                out.append(s[p])
                p += 1
            else:
                if p + 2 > len(s):
                    return 2, bytes(out)
                v = s[p] | s[p + 1] << 8
                p += 2
                ln, d = (v & 31) + 3, v >> 5
                if d == 0 or d > len(out):
                    return -9, bytes(out)
                for _ in range(ln):
                    if len(out) >= out_len:
                        return 1, bytes(out)
                    out.append(out[-d])
    return 0, bytes(out)


def lane_view(chunk: np.ndarray, t: int) -> bytes:
    """Lane t of a chunk: bytes 128k + 4t + j."""
    return chunk.reshape(-1, 32, 4)[:, t, :].reshape(-1).tobytes()


def sort_order(lengths):
    """Sorted position -> original lane: length descending, ties to the higher lane."""
    return sorted(range(32), key=lambda t: (-lengths[t], -t))


def build_container(chunks_streams: list[list[bytes]]):
    """Lane streams [chunk][lane] -> (payload, col_len, blk_off, col_map)."""
    payload = bytearray()
    col_len, col_map, blk_off = [], [], [0]
    for streams in chunks_streams:
        order = sort_order([len(x) for x in streams])
        sl = [streams[t] for t in order]
        words = [(len(x) + 3) // 4 for x in sl]
        body = bytearray()
        for k in range(max(words) if words else 0):
            for s in range(32):
                if words[s] > k:
                    w = sl[s][4 * k: 4 * k + 4]
                    body += w + bytes(4 - len(w))
        body += bytes((-len(body)) % 128)
        payload += body
        blk_off.append(len(payload))
        col_len += [len(x) for x in sl]
        col_map += order
    return (np.frombuffer(bytes(payload), np.uint8).copy(), np.array(col_len, np.uint64),
            np.array(blk_off, np.uint64), np.array(col_map, np.uint8))


def read_container(payload, col_len, blk_off, col_map, n_chunks):
    """Inverse of build_container: [chunk][original lane] -> stream bytes."""
    res = []
    for c in range(n_chunks):
        L = [int(x) for x in col_len[32 * c: 32 * c + 32]]
        words = [(x + 3) // 4 for x in L]
        lanes = [bytearray() for _ in range(32)]
        p = int(blk_off[c])
        for k in range(max(words)):
            for s in range(32):
                if words[s] > k:
                    lanes[s] += bytes(payload[p: p + 4])
                    p += 4
        streams = [b""] * 32
        for s in range(32):
            streams[int(col_map[32 * c + s])] = bytes(lanes[s][: L[s]])
        res.append(streams)
    return res


def assemble(lanes_out: list[list[bytes]], chunk: int) -> bytes:
    """[chunk][lane] decoded lane bytes -> chunk-major output."""
    out = bytearray()
    for lanes in lanes_out:
        buf = np.zeros((chunk // 128, 32, 4), np.uint8)
        for t in range(32):
            buf[:, t, :] = np.frombuffer(lanes[t], np.uint8).reshape(-1, 4)
        out += buf.tobytes()
    return bytes(out)
