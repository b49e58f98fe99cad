"""Batched Deflate scheduler: a learned decode-cost model and the launch
orders it produces.

The decoder logs a cost sample per stream (mdz_inflate_cost_t: cycles and
what the stream held).  ``CostModel`` is trained on those samples with the
library's HIP forward/backward/Adam kernels and predicts log2(cycles) from
features read before decoding: stream sizes, and from the first block header
its type and the symbol, match and block counts its code lengths imply.
``lpt_order`` sorts streams longest-predicted-first for the persistent
work-queue launch; the fallback orders by compressed size.
"""
from __future__ import annotations

import ctypes as C
import heapq

import numpy as np

from ._lib import DeviceBuffer, check, lib
from .lanes import as_ptr

F, H = 11, 32
P = F * H + H + H * H + H + H + 1


def features(batch, parse_hdr: bool = True, stream=None) -> DeviceBuffer:
    """Device float32 [count][F] feature matrix for a DeviceBatch."""
    x = DeviceBuffer(max(batch.count, 1) * F * 4)
    check(lib().mdz_cost_features(batch.desc.ptr, batch.count, int(parse_hdr), x.ptr, as_ptr(stream)),
          "mdz_cost_features")
    return x


class CostModel:
    def __init__(self, params: np.ndarray):
        params = np.ascontiguousarray(params, np.float32)
        assert params.size == P
        self.p = DeviceBuffer.from_array(params)
        self.g = DeviceBuffer(P * 4)
        self.m = DeviceBuffer.from_array(np.zeros(P, np.float32))
        self.v = DeviceBuffer.from_array(np.zeros(P, np.float32))
        self.loss = DeviceBuffer(4)
        self.step = 0

    def params(self) -> np.ndarray:
        return self.p.download(P * 4, np.float32).copy()

    def grad(self, x: DeviceBuffer, t: DeviceBuffer, n: int) -> tuple[np.ndarray, float]:
        check(lib().mdz_mlp_grad(self.p.ptr, x.ptr, t.ptr, n, self.g.ptr, self.loss.ptr, None),
              "mdz_mlp_grad")
        return self.g.download(P * 4, np.float32).copy(), float(self.loss.download(4, np.float32)[0]) / max(n, 1)

    def train(self, x: DeviceBuffer, t: DeviceBuffer, n: int, steps: int = 1500, lr: float = 3e-3,
              stream=None) -> list[float]:
        hist = []
        for i in range(steps):
            check(lib().mdz_mlp_grad(self.p.ptr, x.ptr, t.ptr, n, self.g.ptr, self.loss.ptr,
                                     as_ptr(stream)), "mdz_mlp_grad")
            self.step += 1
            check(lib().mdz_mlp_adam(self.p.ptr, self.g.ptr, self.m.ptr, self.v.ptr, C.c_float(lr),
                                     self.step, as_ptr(stream)), "mdz_mlp_adam")
            if i % 100 == 0 or i == steps - 1:
                hist.append(float(self.loss.download(4, np.float32, stream)[0]) / max(n, 1))
        return hist

    def predict(self, x: DeviceBuffer, n: int, stream=None) -> np.ndarray:
        y = DeviceBuffer(max(n, 1) * 4)
        check(lib().mdz_mlp_forward(self.p.ptr, x.ptr, n, y.ptr, as_ptr(stream)), "mdz_mlp_forward")
        return y.download(n * 4, np.float32, stream).copy()


def init_params(seed: int, target_mean: float = 0.0) -> np.ndarray:
    """He-initialised weights, small positive hidden biases, output bias at
    the mean training target."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P, np.float32)
    o = 0
    p[o:o + F * H] = rng.normal(0, np.sqrt(2.0 / F), F * H); o += F * H
    p[o:o + H] = 0.1; o += H
    p[o:o + H * H] = rng.normal(0, np.sqrt(2.0 / H), H * H); o += H * H
    p[o:o + H] = 0.1; o += H
    p[o:o + H] = rng.normal(0, np.sqrt(1.0 / H), H); o += H
    p[o] = target_mean
    return p


def train_on_log(batch, cycles: np.ndarray, steps: int = 2000, seed: int = 7) -> "CostModel":
    """Fits a model to a decoder-logged batch: features of its streams ->
    log2(measured cycles)."""
    x = features(batch)
    t = np.log2(np.maximum(np.asarray(cycles, np.float64), 1.0)).astype(np.float32)
    td = DeviceBuffer.from_array(t)
    m = CostModel(init_params(seed, float(t.mean())))
    m.train(x, td, batch.count, steps=steps)
    return m


def lpt_order(cost: np.ndarray) -> np.ndarray:
    """Stream indices, largest cost first (ties by index)."""
    return np.lexsort((np.arange(cost.size), -np.asarray(cost, np.float64))).astype(np.uint32)


def makespan(order: np.ndarray, true_cost: np.ndarray, slots: int) -> float:
    """Finish time of greedy list scheduling of `order` on `slots` workers
    (what the persistent work queue does), given true per-stream costs."""
    heap = [0.0] * min(slots, max(len(order), 1))
    heapq.heapify(heap)
    for i in order:
        t = heapq.heappop(heap)
        heapq.heappush(heap, t + float(true_cost[i]))
    return max(heap) if heap else 0.0
