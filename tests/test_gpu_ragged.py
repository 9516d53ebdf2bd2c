"""GPU parity of the ragged persistent kernel: every kind of transform it
runs (one LDS tile, two-level, resampled in one tile, resampled two-level)
against the FP64 oracle, in-place use, and the 10,000-transform benchmark
batch (sampled transforms plus a whole-batch round trip)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, inputs_c, rel_l2, to_dev, to_host, tol_for, torch_dtype

pytestmark = pytest.mark.gpu

# one tile, two-level (4096 < n), resampled (primes / non-smooth) in one tile and two-level
SIZES = [1, 2, 3, 13, 64, 100, 143, 1000, 2310, 4096, 4913, 6000, 8192, 3 * 5 * 7 * 11 * 13, 65536,
         17, 257, 1009, 2053, 4099, 8191, 4097, 9973]


def _check(prec, sizes, seed=5, sign=-1):
    sizes = np.asarray(sizes, dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    x = inputs_c(seed, int(sizes.sum()), prec)
    p = F.plan_ragged(torch_dtype(prec), sizes, sign=sign)
    y = to_host(p(to_dev(x)))
    worst = 0.0
    for n, o in zip(sizes.tolist(), offs.tolist()):
        ref = O.fft(x[o:o + n].astype(np.complex128), sign=sign, scale=1.0 if sign < 0 else 1.0 / n)
        err = rel_l2(y[o:o + n], ref)
        assert err <= tol_for(prec, n), (prec, n, err)
        worst = max(worst, err)
    return p, x, y


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ragged_all_kinds(prec):
    sizes = [n for n in SIZES if prec == "f32" or n <= 32768]
    p, _, _ = _check(prec, sizes)
    d = p.describe()
    for kind in ("direct=", "two_level=", "resample=", "resample_two_level="):
        assert kind in d
    assert " two_level=0" not in d and " resample=0" not in d and "resample_two_level=0" not in d
    # lengths above 4096 points run on the large-tile launch (16384 FP32, 8192 FP64)
    assert " large_transforms=" in d and (" large_tile=16384" if prec == "f32" else " large_tile=8192") in d


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("pass_min", [4096, 16384])
def test_ragged_tile_passes(prec, pass_min):
    """Long smooth transforms as two tile-granular passes (every column tile
    of every such item in one launch, then every row tile), beside items on
    the item-granular launches, both directions, and in place."""
    sizes = np.array([64, 4096, 4608, 6000, 8192, 16384, 16807, 20000, 32768, 257, 8191, 100, 30000],
                     dtype=np.int64)
    if prec == "f32":
        sizes = np.concatenate([sizes, [65536, 50000, 65536 - 2 ** 14, 59049]])
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    x = inputs_c(11, int(sizes.sum()), prec)
    for sign in (-1, 1):
        p = F.plan_ragged(torch_dtype(prec), sizes, sign=sign, tuning=F.ragged.make_tuning(pass_min=pass_min))
        d = p.describe()
        want = int(np.sum([n > max(pass_min, 4096) for n in sizes if n not in (257, 8191)]))
        assert f" pass_transforms={want} " in d, d
        y = to_host(p(to_dev(x)))
        for n, o in zip(sizes.tolist(), offs.tolist()):
            ref = O.fft(x[o:o + n].astype(np.complex128), sign=sign, scale=1.0 if sign < 0 else 1.0 / n)
            assert rel_l2(y[o:o + n], ref) <= tol_for(prec, n), (prec, pass_min, sign, n)
        xd = to_dev(x)
        assert p(xd, xd) is xd
        assert np.array_equal(to_host(xd), y)
    # the planner's default and never: the same transforms to the tolerance
    for pm in (0, -1):
        p = F.plan_ragged(torch_dtype(prec), sizes, tuning=F.ragged.make_tuning(pass_min=pm))
        if pm < 0:
            assert " pass_transforms=" not in p.describe()
        y2 = to_host(p(to_dev(x)))
        ref = to_host(F.plan_ragged(torch_dtype(prec), sizes, tuning=F.ragged.make_tuning(pass_min=pass_min))(to_dev(x)))
        for n, o in zip(sizes.tolist(), offs.tolist()):
            assert rel_l2(y2[o:o + n], ref[o:o + n]) <= tol_for(prec, n), (prec, pm, n)


def test_ragged_inverse_and_inplace():
    sizes = [64, 4096, 8192, 257, 8191, 6000]
    p, x, y = _check("f32", sizes, sign=1)
    xd = to_dev(x)
    assert p(xd, xd) is xd
    assert np.array_equal(to_host(xd), y)


def test_ragged_offsets_with_gaps():
    sizes = np.array([100, 4096, 257], dtype=np.int64)
    offs = np.array([10, 200, 5000], dtype=np.int64)
    x = inputs_c(3, 6000, "f32")
    y0 = np.zeros_like(x)
    p = F.plan_ragged(torch.float32, sizes, offs)
    yd = to_dev(y0)
    p(yd, to_dev(x))
    y = to_host(yd)
    for n, o in zip(sizes, offs):
        assert rel_l2(y[o:o + n], O.fft(x[o:o + n].astype(np.complex128))) <= tol_for("f32", n)
    assert not y[:10].any() and not y[100 + 10:200].any()   # gaps untouched


def test_ragged_benchmark_batch():
    sizes = F.ragged.benchmark_sizes()
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    x = F.rng.uniform_complex(1, int(sizes.sum())).astype(np.complex64)
    xd = to_dev(x)
    pf = F.plan_ragged(torch.float32, sizes)
    pi = F.plan_ragged(torch.float32, sizes, sign=1)
    yd = pf(xd)
    y = to_host(yd)
    rng = np.random.default_rng(0)
    picks = list(rng.choice(len(sizes), 40, replace=False)) + [int(np.argmax(sizes)), int(np.argmin(sizes))]
    for i in picks:
        n, o = int(sizes[i]), int(offs[i])
        assert rel_l2(y[o:o + n], O.fft(x[o:o + n].astype(np.complex128))) <= tol_for("f32", n), n
    back = to_host(pi(yd))
    assert rel_l2(back, x) <= 2 * tol_for("f32", 65536)
    # the plain per-transform loop computes the same transforms
    loop = F.ragged.PlainLoop(torch.float32, sizes[:300], offs[:300])
    y2 = torch.zeros_like(xd)
    loop(y2, xd)
    e = int(offs[299] + sizes[299])
    assert rel_l2(to_host(y2)[:e], y[:e]) <= tol_for("f32", 65536)


def test_ragged_autotune_seeded_by_cost_model():
    """The learned cost model's per-transform costs seed the autotuner's
    workgroup balance; whatever it picks computes the same transforms."""
    sizes = F.ragged.benchmark_sizes()[:1500]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    mc = F.tune.model_costs(sizes)
    assert mc.shape == sizes.shape and np.all(np.isfinite(mc)) and np.all(mc > 0)
    big, small = int(np.argmax(sizes)), int(np.argmin(sizes))
    assert mc[big] > mc[small]
    x = F.rng.uniform_complex(4, int(sizes.sum())).astype(np.complex64)
    xd = to_dev(x)
    yd = torch.empty_like(xd)
    plan, cfg, results = F.tune.autotune_ragged(torch.float32, sizes, x=xd, y=yd, iters=2, budget=2,
                                                 seed_model=True)
    assert len(results) == 4 and sum(1 for c, _ in results if c.get("costs") == "model") == 2
    plan(yd, xd)
    y = to_host(yd)
    for i in list(range(0, len(sizes), 97)) + [big]:
        n, o = int(sizes[i]), int(offs[i])
        assert rel_l2(y[o:o + n], O.fft(x[o:o + n].astype(np.complex128))) <= tol_for("f32", n), (n, cfg)


def test_per_size_baseline_matches_ragged():
    # the per-size baseline (one batched plan per distinct length, data laid
    # out by length) computes the same transforms as the ragged plan
    sizes = F.ragged.benchmark_sizes(count=300, prime_count=15, seed=3)
    x = to_dev(inputs_c(8, int(sizes.sum()), "f32"))
    yr = to_host(F.plan_ragged(torch.float32, sizes)(x))
    per = F.ragged.PerSizeLoop(torch.float32, sizes)
    xs = per.to_layout(x)
    ys = torch.empty_like(xs)
    per(ys, xs)
    yp = to_host(per.from_layout(ys))
    assert len(per.runs) == len(np.unique(sizes))
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    for n, o in zip(sizes.tolist(), offs.tolist()):
        assert rel_l2(yp[o:o + n], yr[o:o + n]) <= 2 * tol_for("f32", n), n


def test_ragged_side_stream_eager_and_graph():
    """Short items (one launch), resampled primes on the large-tile launch
    and long items as two tile passes: the first two run on the plan's side
    stream beside the passes, forked from and joined back to the caller's
    stream.  Calls on another stream, and a captured graph's replays with
    eager work between them, give the same bits as the first eager call."""
    sizes = np.array([64, 100, 1000, 4096, 257, 8191, 20000, 32768, 65536, 6000, 9973], dtype=np.int64)
    x = to_dev(inputs_c(17, int(sizes.sum()), "f32"))
    p = F.plan_ragged(torch.float32, sizes)
    d = p.describe()
    assert " pass_transforms=" in d and " large_transforms=" in d and " direct=" in d
    ref = to_host(p(x))
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    xh = to_host(x)
    for n, o in zip(sizes.tolist(), offs.tolist()):
        assert rel_l2(ref[o:o + n], O.fft(xh[o:o + n].astype(np.complex128), sign=-1)) <= tol_for("f32", n), n
    y = torch.empty_like(x)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            p(y, x)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(y), ref)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        p(y, x)
    z = torch.empty_like(x)
    for i in range(3):
        y.zero_()
        p(z, x)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(to_host(y), ref), i
        assert np.array_equal(to_host(z), ref), i
