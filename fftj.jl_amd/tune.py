"""Autotuner for ragged plans: builds the candidate plans of a small search
space (radix tower 16 or 8, largest one-tile length, the length above which
transforms run as two tile-granular passes), times each with HIP events on
the current stream, keeps the fastest.
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


def ragged_candidates():
    """radix tower, largest one-tile length of the small-item launch, and the
    length above which smooth transforms run as two tile-granular passes."""
    for rp, tm, pm in itertools.product((16, 8), (4096, 2048), (16384, 4096)):
        yield {"radix_pref": rp, "tile_max": tm, "groups": 0, "pass_min": pm}


def model_costs(sizes, planner=None, tile_max=4096):
    """Per-transform costs for the ragged planner's workgroup balance from the
    learned cost model: predicted microseconds per transform of each length
    (batched plans of that length, one prediction per distinct length), plus
    the ragged kernel's fixed cost per item, which batched timings do not
    show: the planner's own per-item constants (element units: 2000 for a
    one-tile item, 6000 two-level, 3000 / 8000 resampled) converted to
    microseconds by the model's median cost per n*log2(n)."""
    from . import planner as _pl
    lp = planner if planner is not None else _pl.default_learned()
    sizes = np.asarray(sizes, dtype=np.int64)
    uniq, inv = np.unique(sizes, return_inverse=True)
    us = lp.per_transform_us(uniq)
    smooth = np.array([_pl.smooth13(int(n)) for n in uniq])
    nl = uniq * np.log2(np.maximum(uniq, 2))
    scale = float(np.median(us[smooth] / nl[smooth])) if smooth.any() else float(np.median(us / nl))
    fixed = np.where(smooth, np.where(uniq <= tile_max, 2000.0, 6000.0), np.where(uniq <= tile_max // 2, 3000.0, 8000.0))
    return (us + scale * fixed)[inv]


def autotune_ragged(T, sizes, offsets=None, *, sign=-1, normalized=True, x=None, y=None, costs=None,
                    candidates=None, iters=5, budget=None, seed_model=False):
    """Returns (best_plan, best_config, results); results lists (config, ms).
    seed_model=True adds, for every configuration, a candidate whose
    workgroup balance uses the learned cost model's per-transform costs
    (config key "costs": "model"; the others use the planner's own)."""
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
    mcosts = model_costs(sizes) if seed_model else None
    if seed_model:
        cands = [dict(c, costs="model") for c in cands] + cands
    results = []
    best = (None, None, float("inf"))
    for cfg in cands:
        knobs = {k: v for k, v in cfg.items() if k != "costs"}
        use = mcosts if cfg.get("costs") == "model" else costs
        plan = RaggedPlan(code, sizes, offsets, sign, normalized, tuning=make_tuning(**knobs), costs=use)
        ms = time_plan(plan, y, x, iters=iters)
        results.append((cfg, ms))
        if ms < best[2]:
            best = (plan, cfg, ms)
        else:
            del plan
    return best[0], best[1], results
