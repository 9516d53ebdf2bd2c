"""Learned planner on the GPU: the HIP MLP kernels (forward, backward, Adam)
against the float64 CPU restatement, training convergence, parity of every
candidate schedule the planners choose among, and what the trained network
gains over the heuristic default on held-out lengths."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import mlp as M
from oracle import oracle as O
from conftest import F, inputs_c, rel_l2, to_dev, to_host, tol_for

pytestmark = pytest.mark.gpu
P = F.planner


def _case(n_in, H, B, seed):
    rng = np.random.default_rng(seed)
    p = (P.init_params(n_in, H, seed) + 0.05 * rng.standard_normal(M.param_count(n_in, H))).astype(np.float32)
    X = rng.standard_normal((B, n_in)).astype(np.float32)
    t = rng.standard_normal(B).astype(np.float32)
    return p, X, t


@pytest.mark.parametrize("n_in,H,B", [(P.N_FEATURES, 64, 256), (5, 7, 1), (3, 64, 1000), (P.N_FEATURES, 64, 4096)])
def test_mlp_forward_and_gradients_match_oracle(n_in, H, B):
    p, X, t = _case(n_in, H, B, seed=B + H)
    m = P.CostModel(n_in, H)
    m.set_params(p)
    out = m.predict(X, raw=True)
    ref, _ = M.forward(p.astype(np.float64), X, n_in, H)
    assert np.max(np.abs(out - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))
    g, loss = m.gradients(X, t)
    gref, lref = M.gradients(p.astype(np.float64), X, t, n_in, H)
    assert loss == pytest.approx(lref, rel=1e-5)
    # FP32 accumulation over B rows: componentwise bound scaled by the gradient's size
    assert np.max(np.abs(g - gref)) <= 2e-5 * max(1e-3, np.max(np.abs(gref))), np.max(np.abs(g - gref))
    assert rel_l2(g, gref) <= 1e-5


def test_adam_steps_match_oracle():
    n_in, H, B = 9, 16, 64
    p, X, t = _case(n_in, H, B, seed=5)
    m = P.CostModel(n_in, H)
    m.set_params(p)
    Xd, td = to_dev(X), to_dev(t)
    pr = p.astype(np.float64)
    mm, vv = np.zeros_like(pr), np.zeros_like(pr)
    for step in range(1, 4):
        gref, _ = M.gradients(pr, X, t, n_in, H)
        pr, mm, vv = M.adam_step(pr, gref, mm, vv, step, lr=1e-2)
        m.train_step(Xd, td, 1e-2)
    got = m.get_params()
    assert np.max(np.abs(got - pr)) <= 1e-5, np.max(np.abs(got - pr))


def test_set_params_resets_and_validates():
    m = P.CostModel(4, 8)
    with pytest.raises(ValueError):
        m.set_params(np.zeros(3, np.float32))
    p = P.init_params(4, 8, 1)
    m.set_params(p)
    np.testing.assert_array_equal(m.get_params(), p)


def test_training_fits_a_known_function():
    """The HIP trainer fits y = sum(sin(x)) well below the variance of y."""
    rng = np.random.default_rng(0)
    X = rng.uniform(-2, 2, size=(4096, 6)).astype(np.float32)
    y = np.sin(X).sum(1).astype(np.float32)
    m = P.CostModel(6, 64)
    curve = m.fit(X, y, steps=3000, batch=256, lr=3e-3)
    pred = m.predict(X)
    mse = float(np.mean((pred - y) ** 2))
    assert curve[-1] < 0.1 * curve[0]
    assert mse < 0.05 * float(np.var(y)), mse


PARITY_SIZES = [(4096, 4), (2187, 4), (1 << 16, 2), (3 * 2 ** 14, 2), (1 << 20, 1), (10 ** 6, 1), (5 ** 6 * 16, 2)]


@pytest.mark.parametrize("n,batch", PARITY_SIZES)
def test_every_candidate_schedule_is_correct(n, batch):
    """Whatever the planners pick, the transform is the same: every valid
    candidate against the oracle, both directions."""
    x = inputs_c(n, n * batch, "f32")
    ref = O.fft(x.astype(np.complex128).reshape(batch, n))
    refi = O.fft(x.astype(np.complex128).reshape(batch, n), sign=1)
    xd = to_dev(x)
    yd = torch.empty_like(xd)
    vc = P.valid_candidates(n, batch)
    assert len(vc) >= 3
    for cand, desc in vc:
        for sign, r in ((-1, ref), (1, refi)):
            p = P.make_plan(n, batch, cand, sign=sign)
            assert p.describe() == P.schedule(n, batch, cand, sign=sign)
            p(yd, xd)
            err = rel_l2(to_host(yd).reshape(batch, n), r)
            assert err <= tol_for("f32", n), (cand, sign, err)


def _trained():
    d = P.DATA_DIR
    return all(os.path.exists(os.path.join(d, P.DATASET + s)) for s in (".npz", "_model.npz", "_report.json"))


@pytest.mark.skipif(not _trained(), reason="no trained planner committed")
def test_learned_planner_beats_default_on_held_out_lengths():
    """On lengths held out of training, the plan the network picks is faster
    (measured times of the committed dataset) than the heuristic default, and
    the report committed with the weights is what the HIP network gives."""
    data = P.load_dataset()
    model = P.load_model()
    pred = model.predict(data["features"])
    test = P.held_out_mask(data)
    r = P.regret_report(data, pred, test, model.margin, model.feature_range)
    assert r["sizes"] >= 100
    assert r["mean_regret_learned"] < r["mean_regret_default"], r
    # the committed ensemble: 1.05x over the default on held-out (length,
    # batch) cases, at most one of them more than 5 % slower than the default
    assert r["geomean_speedup_learned_over_default"] >= 1.045, r
    assert r["cases_learned_5pct_slower"] <= 1 and r["worst_learned_over_default"] < 1.10, r
    with open(os.path.join(P.DATA_DIR, P.DATASET + "_report.json")) as fh:
        rep = json.load(fh)
    assert rep["held_out"]["mean_regret_learned"] == pytest.approx(r["mean_regret_learned"], rel=1e-3, abs=1e-4)
    assert rep["margin"] == pytest.approx(model.margin)


@pytest.mark.skipif(not _trained(), reason="no trained planner committed")
def test_learned_choice_live():
    """Live: the learned pick for held-out lengths builds, is correct, and
    runs within 10 % of the exhaustive planner's best measured right now
    (geometric mean)."""
    data = P.load_dataset()
    lp = P.LearnedPlanner()
    sizes = sorted({int(n) for n in data["n"][P.held_out_mask(data)]})[::12][:6]
    ratios = []
    for n in sizes:
        batch = P.batch_for(n)
        x = torch.randn(n * batch, dtype=torch.complex64, device="cuda")
        y = torch.empty_like(x)
        p, cand, _ = lp.choose(n, batch)
        t_l = P.time_plan(p, y, x) * 1e3
        _, _, t_best = P.exhaustive_choose(n, batch, x=x, y=y)
        ratios.append(t_l / t_best)
        p(y, x)
        rows = min(batch, 2)
        xs = x[:rows * n].cpu().numpy().astype(np.complex128).reshape(rows, n)
        assert rel_l2(to_host(y[:rows * n]).reshape(rows, n), O.fft(xs)) <= tol_for("f32", n), cand
    assert float(np.exp(np.mean(np.log(ratios)))) <= 1.10, list(zip(sizes, ratios))


@pytest.mark.skipif(not _trained(), reason="no trained planner committed")
def test_default_planner_takes_confident_learned_pick():
    """planner=None builds the learned pick inside the model's domain (FP32,
    13-smooth, n*batch near the training cases) and the library's own
    schedule outside it (FP64, short lengths, small batches) or when asked
    for "heuristic"."""
    lp = P.default_learned()
    data = P.load_dataset()
    cases = sorted({(int(n), int(b)) for n, b in zip(data["n"], data["batch"])})[::37]
    for n, b in cases:
        assert lp.confident(n, b)
        cand = lp.pick_cached(n, b)
        assert F.plan_fft(torch.float32, n, batch=b).describe() == P.schedule(n, b, cand), (n, b, cand)
        assert F.plan_fft(torch.float32, n, batch=b, planner="heuristic").describe() == P.schedule(n, b, {})
    for T, n, b in ((torch.float64, 3 ** 12 * 4, 8), (torch.float32, 512, 4), (torch.float32, 3 ** 9, 2)):
        assert F.plan_fft(T, n, batch=b).describe() == \
            F.plan_fft(T, n, batch=b, planner="heuristic").describe(), (T, n, b)


@pytest.mark.skipif(not _trained(), reason="no trained planner committed")
@pytest.mark.parametrize("planner", ["learned", "exhaustive"])
def test_plan_fft_planner_option(planner):
    n, batch = 3 ** 4 * 2 ** 9, 3
    x = inputs_c(7, n * batch, "f32")
    for mk, sign, scale in ((F.plan_fft, -1, 1.0), (F.plan_ifft, 1, 1.0 / n), (F.plan_bfft, 1, 1.0)):
        p = mk(torch.float32, n, batch=batch, planner=planner)
        y = to_host(p * to_dev(x))
        ref = O.fft(x.astype(np.complex128).reshape(batch, n), sign=sign, scale=scale)
        assert rel_l2(y.reshape(batch, n), ref) <= tol_for("f32", n)
