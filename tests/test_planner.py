"""Learned planner, host side: the MLP oracle pinned by finite differences
and Adam's first-step closed form, candidate enumeration and schedules,
feature extraction, and the regret arithmetic.  No GPU."""
import os
import sys

import numpy as np
import pytest

from conftest import F, ROOT

sys.path.insert(0, ROOT)
from oracle import mlp as M  # noqa: E402

P = F.planner


def _params(n_in, H, seed):
    return P.init_params(n_in, H, seed).astype(np.float64) + 0.05 * np.random.default_rng(seed + 1).standard_normal(
        M.param_count(n_in, H))


def test_param_count_matches_layout():
    assert M.param_count(7, 5) == 5 * 7 + 5 + 25 + 5 + 5 + 1
    assert P.init_params(7, 5).size == M.param_count(7, 5)
    W1, b1, W2, b2, W3, b3 = M.unpack(np.arange(M.param_count(3, 2), dtype=np.float64), 3, 2)
    assert W1.shape == (2, 3) and W1[1, 0] == 3 and b1[0] == 6 and W2[0, 1] == 9 and b3[0] == 16


def test_oracle_gradient_finite_differences():
    """Hand backprop == central differences of the loss (float64), every parameter."""
    n_in, H, B = 6, 8, 13
    rng = np.random.default_rng(3)
    p = _params(n_in, H, 3)
    X = rng.standard_normal((B, n_in))
    t = rng.standard_normal(B)
    g, L = M.gradients(p, X, t, n_in, H)
    assert L == pytest.approx(M.loss(p, X, t, n_in, H), rel=1e-15)
    eps = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        a, b = p.copy(), p.copy()
        a[i] += eps
        b[i] -= eps
        fd[i] = (M.loss(a, X, t, n_in, H) - M.loss(b, X, t, n_in, H)) / (2 * eps)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_oracle_gradient_catches_a_planted_mistake():
    """A backprop with the ReLU mask taken from the wrong layer fails the same check."""
    n_in, H, B = 5, 6, 9
    rng = np.random.default_rng(4)
    p = _params(n_in, H, 4)
    X = rng.standard_normal((B, n_in))
    t = rng.standard_normal(B)
    W1, b1, W2, b2, W3, b3 = M.unpack(p, n_in, H)
    out, (X_, h1, h2) = M.forward(p, X, n_in, H)
    dout = ((out - t) / B)[:, None]
    d2 = (dout @ W3) * (h1 > 0)  # wrong mask
    gW2_bad = d2.T @ h1
    g, _ = M.gradients(p, X, t, n_in, H)
    off = H * n_in + H
    assert np.max(np.abs(gW2_bad.ravel() - g[off:off + H * H])) > 1e-6


def test_oracle_linear_network_closed_form():
    """With every pre-activation positive the net is affine: out = X A + c,
    and the gradient of the last bias is the mean residual."""
    n_in, H = 3, 4
    p = np.zeros(M.param_count(n_in, H))
    W1, b1, W2, b2, W3, b3 = M.unpack(p, n_in, H)
    W1[:] = 0.1
    b1[:] = 5.0
    W2[:] = np.eye(H) * 0.5
    b2[:] = 1.0
    W3[:] = 2.0
    b3[:] = -1.0
    X = np.array([[1.0, 2.0, 3.0], [0.0, -1.0, 0.5]])
    t = np.array([1.0, 2.0])
    out, _ = M.forward(p, X, n_in, H)
    h1 = X @ W1.T + b1
    exp_out = ((h1 @ W2.T + b2) @ W3.T + b3)[:, 0]
    np.testing.assert_allclose(out, exp_out, rtol=1e-15)
    g, _ = M.gradients(p, X, t, n_in, H)
    assert g[-1] == pytest.approx(np.mean(out - t), rel=1e-14)


def test_adam_first_step_closed_form():
    """Step 1 of bias-corrected Adam moves each parameter by lr*g/(|g|+eps)."""
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    p = np.array([1.0, 1.0, 1.0, 1.0])
    p1, m, v = M.adam_step(p, g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    np.testing.assert_allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_candidates_are_distinct_and_default_first():
    for n in (1024, 4096, 3 ** 7, 1 << 20, 10 ** 6, 1 << 24):
        c = P.candidates(n)
        assert c[0] == {}
        assert len({repr(x) for x in c}) == len(c)
        for x in c:
            if "splits" in x:
                assert int(np.prod(x["splits"])) == n
                assert all(16 <= f <= 16384 for f in x["splits"])


def test_radix_chain():
    assert P.radix_chain(4096, 16) == [16, 16, 16]
    assert P.radix_chain(4096, 8) == [8, 8, 8, 8]
    assert P.radix_chain(2 ** 7 * 3 * 13, 16) == [16, 8, 13, 3]
    assert P.radix_chain(17 * 4, 16) is None


def test_schedules_of_candidates():
    """Every valid candidate's host schedule is the one its tuning asks for."""
    n = 1 << 24
    vc = P.valid_candidates(n, 1)
    assert vc[0][1] == F._lib.describe_request(F._lib.KIND_C2C, F._lib.F32, [n], batch=1)
    for cand, desc in vc:
        passes = [ln.split() for ln in desc.splitlines() if ln.strip().startswith("pass")]
        if "splits" in cand:
            ns = [int(dict(t.split("=", 1) for t in ln if "=" in t)["n"]) for ln in passes]
            assert ns == list(reversed(cand["splits"])), (cand, ns)
        if cand.get("passes", [{}])[0].get("C"):
            kv = dict(t.split("=", 1) for t in passes[0] if "=" in t)
            assert int(kv["C"]) == cand["passes"][0]["C"]
    # radix overrides reach the generic kernel
    d = P.schedule(4096, 8, {"passes": [{"radices": [8, 8, 8, 8]}]})
    assert "radices=8,8,8,8" in d and "kernel=pow2" not in d


def test_schedule_rejects_bad_tuning():
    with pytest.raises(F.ArgumentError):
        P.schedule(4096, 1, {"splits": [64, 32]})
    with pytest.raises(F.ArgumentError):
        P.schedule(4096, 1, {"passes": [{"radices": [16, 16, 8]}]})


def test_features_from_schedule():
    desc = ("c2c f32 dims=1048576 batch=4 sign=-1 launches=2\n"
            "  pass x->y n=1024 radices=16,16,4 kernel=pow2 batch=4096 C=16 threads=64 groups=256 "
            "in_stride=1024 out_stride=1024 twiddle=w_1048576\n"
            "  pass y->y n=1024 radices=4,5,3 batch=4096 C=4 threads=256 groups=1024 in_stride=1 out_stride=1\n")
    f = P.features(desc, 1 << 20, 4)
    assert f.shape == (P.N_FEATURES,) and f.dtype == np.float32
    assert f[0] == pytest.approx(20 / 24) and f[1] == pytest.approx(2 / 24) and f[4] == pytest.approx(0.5)
    b0, b1 = 6, 6 + P.PASS_FEATURES
    assert f[b0] == 1 and f[b0 + 1] == 1 and f[b0 + 10] == 1
    assert f[b1] == 1 and f[b1 + 1] == 0 and f[b1 + 10] == 0
    assert f[b0 + 3] == pytest.approx(4 * 4 / 24) and f[b1 + 3] == pytest.approx(4 * 2 / 24)
    assert f[b1 + 9] == pytest.approx(3 / 8)  # radices 4, 5, 3 are below 8 or odd
    assert np.all(f[6 + 2 * P.PASS_FEATURES:] == 0)
    # occupancy: one-wave groups of a 139 KB tile (one per CU by LDS), then
    # four-wave groups of 35 KB (four per CU fill all 16 wave slots)
    assert f[b0 + 17] == pytest.approx(1 / 16) and f[b0 + 18] == pytest.approx(1 / 16)
    assert f[b1 + 17] == pytest.approx(4 / 16) and f[b1 + 18] == pytest.approx(1.0)
    assert f[b0 + 11] == f[b0 + 12] == f[b0 + 13] == f[b0 + 14] == f[b0 + 16] == 0


def test_training_sizes_deterministic_and_smooth():
    a = P.training_sizes(50)
    assert a == P.training_sizes(50) and len(set(a)) == 50
    assert all(P.smooth13(n) and 1024 <= n <= 1 << 24 for n in a)


def test_regret_report_arithmetic():
    """Two sizes, three candidates each: regrets by hand."""
    t = np.array([10.0, 8.0, 12.0, 5.0, 6.0, 4.0])
    data = {"size_id": np.array([0, 0, 0, 1, 1, 1]), "log_us": np.log(t).astype(np.float32),
            "is_default": np.array([1, 0, 0, 1, 0, 0]), "n": np.array([64, 64, 64, 81, 81, 81])}
    pred = np.array([3.0, 2.0, 1.0, 0.5, 0.1, 0.2])  # learned picks 12 (size 0) and 6 (size 1)
    r = P.regret_report(data, pred, np.ones(6, bool))
    assert r["sizes"] == 2
    assert r["mean_regret_learned"] == pytest.approx(((12 / 8 - 1) + (6 / 4 - 1)) / 2, rel=1e-6)
    assert r["mean_regret_default"] == pytest.approx(((10 / 8 - 1) + (5 / 4 - 1)) / 2, rel=1e-6)
    assert r["geomean_speedup_learned_over_default"] == pytest.approx(np.sqrt(10 / 12 * 5 / 6), rel=1e-6)


def _artifacts():
    d = P.DATA_DIR
    return [os.path.join(d, P.DATASET + s) for s in (".npz", "_model.npz", "_report.json")]


@pytest.mark.skipif(not all(os.path.exists(p) for p in _artifacts()), reason="no trained planner committed")
def test_committed_dataset_shape():
    data = P.load_dataset()
    n = len(data["log_us"])
    assert n >= 4000
    assert data["features"].shape == (n, P.N_FEATURES)
    assert np.all(np.isfinite(data["log_us"]))
    # one default row per (length, batch) case, two footprints per length
    cids = np.unique(data["case_id"])
    assert int(data["is_default"].sum()) == len(cids) == 2 * len(np.unique(data["size_id"]))
    with np.load(_artifacts()[1], allow_pickle=False) as z:
        k = int(z["members"])
        assert k >= 3 and all(z[f"m{i}_params"].size == M.param_count(P.N_FEATURES, P.HIDDEN) for i in range(k))
        assert 0.0 <= float(z["margin"]) <= 0.3


def test_planner_option_validated():
    import torch
    with pytest.raises(F.ArgumentError):
        F.plan_fft(torch.float32, 4096, planner="fastest")
