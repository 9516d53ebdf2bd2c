"""Autotuner for ragged plans: builds the candidate plans of a small search
space (radix tower 16 or 8, largest one-tile length, persistent workgroup
count), times each with HIP events on the current stream, keeps the fastest.
An optional cost model orders the candidates and supplies the per-transform
costs that balance the workgroups."""
import itertools

import numpy as np
import torch

from .ragged import RaggedPlan, make_tuning
from .api import _code


def time_plan(plan, y, x, iters=5, warmup=1):
    """Mean milliseconds of plan(y, x) over `iters` calls (HIP events)."""
    for _ in range(warmup):
        plan(y, x)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(iters):
        plan(y, x)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / iters


def ragged_candidates(resident=1024):
    for rp, tm, g in itertools.product((16, 8), (4096, 2048), (0, resident // 2)):
        yield {"radix_pref": rp, "tile_max": tm, "groups": g}


def autotune_ragged(T, sizes, offsets=None, *, sign=-1, normalized=True, x=None, y=None, costs=None,
                    candidates=None, iters=5, budget=None):
    """Returns (best_plan, best_config, results); results lists (config, ms)."""
    code = _code(T)
    sizes = np.asarray(sizes, dtype=np.int64)
    extent = int(sizes.sum()) if offsets is None else int((np.asarray(offsets) + sizes).max())
    dt = torch.complex64 if code == 0 else torch.complex128
    if x is None:
        x = torch.randn(extent, dtype=dt, device="cuda")
    if y is None:
        y = torch.empty_like(x)
    cands = list(candidates if candidates is not None else ragged_candidates())
    if budget is not None:
        cands = cands[:budget]
    results = []
    best = (None, None, float("inf"))
    for cfg in cands:
        plan = RaggedPlan(code, sizes, offsets, sign, normalized, tuning=make_tuning(**cfg), costs=costs)
        ms = time_plan(plan, y, x, iters=iters)
        results.append((cfg, ms))
        if ms < best[2]:
            best = (plan, cfg, ms)
        else:
            del plan
    return best[0], best[1], results
