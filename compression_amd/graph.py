"""Captured pipelines: the mdz launches issued on one stream, recorded once
into a hipGraph and replayed with one launch (mdz_graph_*).

    s = Stream()
    with capture(s) as g:
        batch.submit(s.ptr)
    g.launch(s.ptr)          # replays the whole pipeline

The graph keeps the pointers the captured calls were given, so every buffer
they name must outlive it.
"""
from __future__ import annotations

import ctypes as C
from contextlib import contextmanager

from ._lib import Stream, check, lib

vp = C.c_void_p


class Graph:
    """An instantiated hipGraph of mdz launches."""

    def __init__(self):
        self.ptr = None

    def launch(self, stream) -> None:
        check(lib().mdz_graph_launch(self.ptr, _sp(stream)), "mdz_graph_launch")

    def nodes(self) -> tuple[int, int]:
        """(all nodes, kernel nodes) of the captured graph."""
        n, k = C.c_uint64(), C.c_uint64()
        check(lib().mdz_graph_nodes(self.ptr, C.byref(n), C.byref(k)), "mdz_graph_nodes")
        return n.value, k.value

    def destroy(self) -> None:
        if self.ptr is not None:
            lib().mdz_graph_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _sp(stream):
    return stream.ptr if isinstance(stream, Stream) else stream


@contextmanager
def capture(stream):
    """Capture the mdz calls made on ``stream`` inside the block into a Graph."""
    sp = _sp(stream)
    check(lib().mdz_graph_begin(sp), "mdz_graph_begin")
    g = Graph()
    h = vp()
    try:
        yield g
    except BaseException:
        lib().mdz_graph_end(sp, C.byref(h))        # always leave capture mode
        if h.value:
            lib().mdz_graph_destroy(h)
        raise
    check(lib().mdz_graph_end(sp, C.byref(h)), "mdz_graph_end")
    g.ptr = h.value
