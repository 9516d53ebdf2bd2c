"""Ragged batches: many independent complex transforms of uneven lengths in
one buffer, run by one persistent-kernel launch (libfftjh's ragged plan),
plus the seeded batch of the ragged benchmark configuration.

The reference has no ragged (or batch) API; its callers loop over 1-D plans
(plan.jl:207-213).  PlainLoop is that loop on the GPU plans; PerSizeLoop, one
batched plan per distinct length on data laid out by length, is the stronger
plain baseline the ragged plan is also measured against.
"""
import ctypes

import numpy as np
import torch

from . import _lib, rng
from ._lib import ArgumentError, check, lib
from .api import _CPLX, _code, _PlanBase, _require, _Handle


class RaggedPlan(_PlanBase):
    """`len(sizes)` transforms; transform i is x[offsets[i] : offsets[i] + sizes[i]]."""

    def __init__(self, dtype, sizes, offsets=None, sign=-1, normalized=True, tuning=None, costs=None):
        self.dtype_code = dtype
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        if offsets is None:
            offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        if self.offsets.shape != self.sizes.shape:
            raise ArgumentError("sizes and offsets differ in length")
        self.sign = int(sign)
        self.normalized = bool(normalized)
        self.extent = int((self.offsets + self.sizes).max()) if self.sizes.size else 0
        P64 = ctypes.POINTER(ctypes.c_int64)
        h = ctypes.c_void_p()
        sz = self.sizes.ctypes.data_as(P64)
        of = self.offsets.ctypes.data_as(P64)
        if tuning is None and costs is None:
            check(lib().fftjh_plan_ragged(ctypes.byref(h), dtype, len(self.sizes), sz, of, self.sign,
                                          int(self.normalized), 0))
        else:
            t = tuning if tuning is not None else _lib.RaggedTuning()
            if costs is not None:
                self._costs = np.ascontiguousarray(costs, dtype=np.float64)
                t.costs = self._costs.ctypes.data
            check(lib().fftjh_plan_ragged_tuned(ctypes.byref(h), dtype, len(self.sizes), sz, of, self.sign,
                                                int(self.normalized), ctypes.byref(t)))
        self._h = _Handle(h.value)

    def __call__(self, y, x=None):
        if x is None:
            x, y = y, None
        _require(x, _CPLX[self.dtype_code], "input")
        if x.numel() < self.extent:
            raise _lib.DimensionMismatch(f"ragged batch spans {self.extent} elements, input has {x.numel()}")
        if y is None:
            y = torch.empty_like(x)
        _require(y, _CPLX[self.dtype_code], "output")
        if y.numel() < self.extent:
            raise _lib.DimensionMismatch(f"ragged batch spans {self.extent} elements, output has {y.numel()}")
        return self._exec(y, x)

    def __mul__(self, x):
        return self(x)


def plan_ragged(T, sizes, offsets=None, *, sign=-1, normalized=True, tuning=None, costs=None):
    return RaggedPlan(_code(T), sizes, offsets, sign, normalized, tuning, costs)


def make_tuning(groups=0, radix_pref=0, tile_max=0, pass_min=0):
    """pass_min: smooth transforms longer than this run as two tile-granular
    passes (0: the planner's default, -1: never)."""
    t = _lib.RaggedTuning()
    t.groups, t.radix_pref, t.tile_max = int(groups), int(radix_pref), int(tile_max)
    t.pass_min = int(pass_min)
    return t


class PlainLoop:
    """The reference's way: one 1-D plan per length, applied transform by
    transform (one plan execution per transform)."""

    def __init__(self, T, sizes, offsets=None, sign=-1):
        from .api import plan_bfft, plan_fft, plan_ifft
        self.sizes = np.asarray(sizes, dtype=np.int64)
        self.offsets = (np.concatenate([[0], np.cumsum(self.sizes)[:-1]]) if offsets is None
                        else np.asarray(offsets, dtype=np.int64))
        mk = plan_fft if sign < 0 else plan_ifft
        self.plans = {int(n): mk(T, int(n)) for n in np.unique(self.sizes)}

    def __call__(self, y, x):
        for n, o in zip(self.sizes.tolist(), self.offsets.tolist()):
            self.plans[n](y[o:o + n], x[o:o + n])
        return y


class PerSizeLoop:
    """The strongest plain baseline: one batched plan per distinct length
    (plan_fft(T, n, batch=count), the engine's own batch API), on the data
    laid out by length -- each length's transforms back to back -- so every
    launch is a plain batched transform with no gather.  `order[i]` is the
    index, in the ragged layout, of the i-th transform of the by-length
    layout; to_layout / from_layout move data between the two."""

    def __init__(self, T, sizes, sign=-1):
        from .api import plan_fft, plan_ifft
        self.sizes = np.asarray(sizes, dtype=np.int64)
        self.order = np.argsort(self.sizes, kind="stable")
        by_len = self.sizes[self.order]
        uniq, first, counts = np.unique(by_len, return_index=True, return_counts=True)
        offs = np.concatenate([[0], np.cumsum(by_len)[:-1]]).astype(np.int64)
        self.runs = [(int(n), int(offs[f]), int(c)) for n, f, c in zip(uniq, first, counts)]
        mk = plan_fft if sign < 0 else plan_ifft
        self.plans = {n: mk(T, n, batch=c) for n, _, c in self.runs}
        ragged_offs = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        # element gather index: by-length position -> ragged position
        self._gather = torch.from_numpy(np.concatenate(
            [np.arange(ragged_offs[i], ragged_offs[i] + self.sizes[i]) for i in self.order]))

    def to_layout(self, x):
        """The ragged-layout tensor x in the by-length layout."""
        return x[self._gather.to(x.device)]

    def from_layout(self, z):
        """The by-length-layout tensor z back in the ragged layout."""
        out = torch.empty_like(z)
        out[self._gather.to(z.device)] = z
        return out

    def __call__(self, y, x):
        for n, o, c in self.runs:
            self.plans[n](y[o:o + n * c], x[o:o + n * c])
        return y


# ------------------------------------------------------ the benchmark batch

def smooth7(lo, hi):
    """All 7-smooth integers in [lo, hi], sorted."""
    out = []
    p2 = 1
    while p2 <= hi:
        p3 = p2
        while p3 <= hi:
            p5 = p3
            while p5 <= hi:
                p7 = p5
                while p7 <= hi:
                    if p7 >= lo:
                        out.append(p7)
                    p7 *= 7
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return np.array(sorted(out), dtype=np.int64)


def primes(lo, hi):
    sieve = np.ones(hi + 1, dtype=bool)
    sieve[:2] = False
    for p in range(2, int(hi ** 0.5) + 1):
        if sieve[p]:
            sieve[p * p::p] = False
    return np.nonzero(sieve)[0][np.nonzero(sieve)[0] >= lo].astype(np.int64)


def benchmark_sizes(count=10000, prime_count=500, seed=0):
    """The ragged benchmark batch: (count - prime_count) lengths drawn
    log-uniformly over [64, 65536] and snapped (in log space) to the nearest
    7-smooth length, plus prime_count primes uniform over the primes in
    [257, 8191]; the list is then shuffled.  All draws from splitmix64(seed)."""
    cand = smooth7(64, 65536)
    logc = np.log(cand.astype(np.float64))
    ns = count - prime_count
    u = (rng.uniform(seed, ns) + 1.0) / 2.0
    t = np.log(64.0) + u * (np.log(65536.0) - np.log(64.0))
    idx = np.searchsorted(logc, t)
    idx = np.clip(idx, 1, len(cand) - 1)
    pick = np.where(np.abs(logc[idx - 1] - t) <= np.abs(logc[idx] - t), idx - 1, idx)
    smooth = cand[pick]
    pr = primes(257, 8191)
    v = (rng.uniform(seed, prime_count, first=ns) + 1.0) / 2.0
    chosen = pr[np.minimum((v * len(pr)).astype(np.int64), len(pr) - 1)]
    sizes = np.concatenate([smooth, chosen])
    # Fisher-Yates with splitmix draws
    w = (rng.uniform(seed, count, first=count) + 1.0) / 2.0
    for i in range(count - 1, 0, -1):
        j = int(w[i] * (i + 1))
        sizes[i], sizes[j] = sizes[j], sizes[i]
    return sizes
