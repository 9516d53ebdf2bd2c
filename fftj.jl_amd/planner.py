"""Planners for 1-D complex plans: the heuristic default, an exhaustive
planner that measures every candidate, and a learned planner whose cost
model (a 3-layer MLP, hidden 64, FP32) is trained on timings the GPU plans
themselves produce, with libfftjh's own HIP forward/backward/Adam kernels.

The reference has no measuring planner: it routes by fixed host thresholds
(plan.jl:183-213) and ignores FFTW planner flags (api.jl:107-108).

Candidates for a length n: the planner's default; for a one-pass length,
other workgroup group sizes and generic radix chains (radix-8 or radix-4
towers instead of the fused power-of-two kernel); for longer lengths,
multi-pass factorizations n = F0*F1*...*F(P-1) (F0 the contiguous row pass,
run last), each also with a narrower first pass.  Features describe the
launch schedule a candidate compiles to (parsed from the plan's own
description), so the model sees exactly what runs.

Run as a script on the GPU to collect timings and train:
    python fftj.jl_amd/planner.py --out DIR [--sizes 400]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np
import torch

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import __graft_entry__
    _pkg = __graft_entry__.load_package()
    _lib = sys.modules[_pkg.__name__ + "._lib"]
    FFTPlan = sys.modules[_pkg.__name__ + ".api"].FFTPlan
    rng = sys.modules[_pkg.__name__ + ".rng"]
else:
    from . import _lib, rng
    from .api import FFTPlan

check, lib = _lib.check, _lib.lib

HERE = os.path.dirname(os.path.abspath(__file__))
DATA_DIR = os.path.join(HERE, "planner_data")
DATASET = "c2c_f32_mi355x"
MAX_PASSES = 4
PASS_FEATURES = 19
N_FEATURES = 6 + MAX_PASSES * PASS_FEATURES
HIDDEN = 64
_ONE_PASS_MAX = 16384


# --------------------------------------------------------------- candidates

def smooth13(n):
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


def _divisors(n, lo=16, hi=_ONE_PASS_MAX):
    ds = set()
    i = 1
    while i * i <= n:
        if n % i == 0:
            ds.add(i)
            ds.add(n // i)
        i += 1
    return sorted(d for d in ds if lo <= d <= hi and smooth13(d))


def _factor_lists(n, parts):
    """Ordered factorizations of n into `parts` one-pass factors (>= 16)."""
    divs = _divisors(n)
    out = []

    def rec(rem, k, cur):
        if k == 1:
            if 16 <= rem <= _ONE_PASS_MAX and smooth13(rem):
                out.append(cur + [rem])
            return
        for f in divs:
            if f > rem:
                break
            if rem % f == 0:
                rec(rem // f, k - 1, cur + [f])

    rec(n, parts, [])
    return out


def radix_chain(n, top):
    """Radices of n with power-of-two radices at most `top` (largest first),
    then odd primes; None when a factor above 13 remains."""
    rad = []
    for p in (13, 11, 7, 5, 3):
        while n % p == 0:
            rad.append(p)
            n //= p
    pw = []
    r = top
    while n > 1 and r >= 2:
        while n % r == 0:
            pw.append(r)
            n //= r
        r //= 2
    return pw + rad if n == 1 else None


def _chain_factor(f):
    """A pass length with a compile-time chain kernel (powers of two, 81, 100)."""
    return f in (81, 100) or (f & (f - 1)) == 0


def candidates(n, max_per_p=4):
    """Tuning dicts (splits, passes) for length n; {} is the default plan.
    Multi-pass lengths get the most balanced factorizations of each pass
    count (also with a 4-column first pass) and the most balanced ones whose
    column factors (all but the first, the row pass) run compile-time
    chains."""
    cands = [{}]
    if n <= _ONE_PASS_MAX:
        for C in (1, 2, 4, 8):
            cands.append({"passes": [{"C": C}]})
        for top in (16, 8, 4):
            rc = radix_chain(n, top)
            if rc and len(rc) <= 16:
                cands.append({"passes": [{"radices": rc}]})
    for parts in (2, 3, 4):
        if n < 16 ** parts:
            continue
        fl = _factor_lists(n, parts)
        fl.sort(key=lambda f: (max(f) / min(f), f))
        for f in fl[:max_per_p]:
            cands.append({"splits": f})
            cands.append({"splits": f, "passes": [{"C": 4}]})
        for f in [f for f in fl if f[0] <= 1024 and all(_chain_factor(c) for c in f[1:])][:2]:
            cands.append({"splits": f})
    seen, out = set(), []
    for c in cands:
        k = repr(c)
        if k not in seen:
            seen.add(k)
            out.append(c)
    return out


def _tuning(cand):
    return _lib.make_tuning(cand.get("splits", ()), cand.get("passes", ())) if cand else None


def make_plan(n, batch, cand, sign=-1, dtype=0):
    return FFTPlan(dtype, (n,), sign, 1.0, batch, tuning=_tuning(cand))


def schedule(n, batch, cand, sign=-1, dtype=0):
    """The launch schedule a candidate compiles to (host only, no GPU)."""
    return _lib.describe_tuned(dtype, (n,), batch, sign, _tuning(cand))


def valid_candidates(n, batch, sign=-1, dtype=0):
    """(candidate, schedule) for each candidate the planner accepts; a
    candidate whose schedule repeats an earlier one is dropped (the default
    comes first, so it is always kept)."""
    out, seen = [], set()
    for cand in candidates(n):
        try:
            d = schedule(n, batch, cand, sign, dtype)
        except (ValueError, NotImplementedError):
            continue
        if d not in seen:
            seen.add(d)
            out.append((cand, d))
    return out


# ----------------------------------------------------------------- features

def _lg(v):
    return math.log2(max(float(v), 1.0)) / 24.0


def features(desc, n, batch, dtype=0):
    """Fixed-length feature vector of a plan from its launch schedule."""
    lines = [ln.strip() for ln in desc.splitlines()[1:] if ln.strip()]
    f = np.zeros(N_FEATURES, dtype=np.float32)
    f[0] = _lg(n)
    f[1] = _lg(batch)
    f[2] = _lg(n * batch)
    f[3] = float(dtype)
    f[4] = len(lines) / 4.0
    f[5] = sum(1 for ln in lines if not ln.startswith("pass")) / 4.0
    passes = [ln for ln in lines if ln.startswith("pass")][:MAX_PASSES]
    for k, ln in enumerate(passes):
        toks = ln.split()
        kv = dict(tok.split("=", 1) for tok in toks if "=" in tok)
        kernels = {tok.split("=", 1)[1] for tok in toks if tok.startswith("kernel=")}
        base = 6 + k * PASS_FEATURES
        rad = [int(r) for r in kv.get("radices", "1").split(",") if r]
        f[base + 0] = 1.0
        f[base + 1] = 1.0 if "pow2" in kernels else 0.0
        f[base + 2] = _lg(int(kv.get("n", 1)))
        f[base + 3] = _lg(int(kv.get("C", 1))) * 4.0
        f[base + 4] = _lg(int(kv.get("threads", 64))) * 2.0
        f[base + 5] = _lg(int(kv.get("groups", 1)))
        f[base + 6] = _lg(int(kv.get("in_stride", 1)))
        f[base + 7] = _lg(int(kv.get("out_stride", 1)))
        f[base + 8] = len(rad) / 8.0
        f[base + 9] = sum(1 for r in rad if r < 8 or r % 2) / 8.0
        f[base + 10] = 1.0 if "twiddle" in kv else 0.0
        f[base + 11] = 1.0 if "chain" in kernels else 0.0
        f[base + 12] = 1.0 if "lean" in kernels else 0.0
        f[base + 13] = 1.0 if "fused" in kernels else 0.0
        st = kv.get("stream", "")
        f[base + 14] = 1.0 if st in ("in", "inout") else 0.0
        f[base + 15] = 1.0 if st in ("out", "inout") else 0.0
        f[base + 16] = 1.0 if "rows_out" in toks else 0.0
        # occupancy: waves per group, and the share of a CU's 16 wave slots
        # (4 per SIMD at the kernels' register budgets) that whole groups
        # fill, also bounded by a group's LDS tile
        waves = max(1, int(kv.get("threads", 64)) // 64)
        tile = int(kv.get("C", 1)) * int(kv.get("n", 1)) * 8 * 17 // 16
        groups = min(16 // waves if waves <= 16 else 1, max(1, (160 << 10) // max(tile, 1)))
        f[base + 17] = waves / 16.0
        f[base + 18] = min(1.0, groups * waves / 16.0)
    return f


# ------------------------------------------------------------- measurement

def time_plan(p, y, x, iters=10, warmup=2, repeats=3):
    """Milliseconds of p(y, x): the median over `repeats` of the mean of
    `iters` back-to-back calls (HIP events on the current stream)."""
    for _ in range(warmup):
        p(y, x)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(repeats)]
    for a, b in ev:
        a.record()
        for _ in range(iters):
            p(y, x)
        b.record()
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) / iters for a, b in ev]))


def batch_for(n, target=1 << 22):
    return max(1, target // n)


def training_sizes(count, lo=1 << 10, hi=1 << 24, seed=7):
    """Distinct 13-smooth lengths drawn log-uniformly from [lo, hi] (splitmix64)."""
    u = (rng.uniform(seed, 64 * count) + 1.0) / 2.0
    out, seen = [], set()
    for v in u:
        m = int(round(math.exp(math.log(lo) + float(v) * (math.log(hi) - math.log(lo)))))
        while not smooth13(m):
            m += 1
        if m <= hi and m not in seen:
            seen.add(m)
            out.append(m)
        if len(out) == count:
            break
    return out


FOOTPRINTS = (1 << 22, 1 << 25)


def collect(sizes, footprints=FOOTPRINTS, log=None):
    """Measure every candidate plan of every size at each footprint (complex
    elements per plan: 32 MiB, inside the Infinity Cache, and 256 MiB, out of
    it), the exhaustive planner's work; returns a dict of arrays, one row per
    (size, footprint, candidate).  size_id numbers the lengths, case_id the
    (length, batch) pairs."""
    feats, logt, size_id, case_id, cand_id, is_def, ns, bs = [], [], [], [], [], [], [], []
    case = 0
    for si, n in enumerate(sizes):
        for fp in footprints:
            batch = batch_for(n, fp)
            x = torch.randn(n * batch, dtype=torch.complex64, device="cuda")
            y = torch.empty_like(x)
            for ci, (cand, desc) in enumerate(valid_candidates(n, batch)):
                p = make_plan(n, batch, cand)
                ms = time_plan(p, y, x, repeats=5)
                feats.append(features(desc, n, batch))
                logt.append(math.log(ms * 1e3))  # log microseconds
                size_id.append(si)
                case_id.append(case)
                cand_id.append(ci)
                is_def.append(1 if not cand else 0)
                ns.append(n)
                bs.append(batch)
                del p
            case += 1
            del x, y
        if log and (si % 10 == 0 or si == len(sizes) - 1):
            log(f"size {si + 1}/{len(sizes)} n={n}: {len(logt)} timings so far")
    return {"features": np.array(feats, dtype=np.float32), "log_us": np.array(logt, dtype=np.float32),
            "size_id": np.array(size_id, dtype=np.int32), "case_id": np.array(case_id, dtype=np.int32),
            "cand_id": np.array(cand_id, dtype=np.int32), "is_default": np.array(is_def, dtype=np.int8),
            "n": np.array(ns, dtype=np.int64), "batch": np.array(bs, dtype=np.int64)}


# --------------------------------------------------------------- the model

def init_params(in_features=N_FEATURES, hidden=HIDDEN, seed=0):
    """He-uniform weights, zero biases (seeded numpy generator)."""
    r = np.random.default_rng(seed)

    def w(o, i):
        lim = math.sqrt(6.0 / i)
        return r.uniform(-lim, lim, size=(o, i)).astype(np.float32)

    parts = [w(hidden, in_features), np.zeros(hidden, np.float32), w(hidden, hidden), np.zeros(hidden, np.float32),
             w(1, hidden), np.zeros(1, np.float32)]
    return np.concatenate([p.ravel() for p in parts])


class CostModel:
    """The MLP on the GPU (libfftjh's HIP kernels): in -> 64 -> 64 -> 1,
    predicting log microseconds.  Input features are standardized with the
    training set's mean/std, kept with the weights."""

    def __init__(self, in_features=N_FEATURES, hidden=HIDDEN, seed=0):
        self.in_features, self.hidden = in_features, hidden
        h = ctypes.c_void_p()
        check(lib().fftjh_mlp_create(ctypes.byref(h), in_features, hidden))
        self._h = h
        self.nparams = int(lib().fftjh_mlp_param_count(h))
        self.mu = np.zeros(in_features, np.float32)
        self.sd = np.ones(in_features, np.float32)
        self.log_range = (-np.inf, np.inf)  # training targets' range (log us)
        self.margin = 0.0  # predicted gain a non-default pick must show (calibrated)
        self.elems_range = (0, 0)  # n*batch range of the training cases
        self.set_params(init_params(in_features, hidden, seed))

    def __del__(self):
        # (at interpreter exit the module globals may already be gone)
        L = getattr(_lib, "_LIB", None) if _lib is not None else None
        if getattr(self, "_h", None) and L is not None:
            L.fftjh_mlp_destroy(self._h)
            self._h = None

    @staticmethod
    def _ok(rc):
        if rc:
            raise RuntimeError("fftjh_mlp: " + lib().fftjh_mlp_last_error().decode())

    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {p.size}")
        self._ok(lib().fftjh_mlp_set_params(self._h, p.ctypes.data))

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        self._ok(lib().fftjh_mlp_get_params(self._h, p.ctypes.data))
        return p

    def normalize(self, X):
        return ((np.asarray(X, np.float32) - self.mu) / self.sd).astype(np.float32)

    def _dev(self, X, raw):
        if isinstance(X, torch.Tensor) and X.is_cuda:
            return X.contiguous()
        X = np.asarray(X, np.float32)
        if not raw:
            X = self.normalize(X)
        return torch.from_numpy(np.ascontiguousarray(X)).to("cuda")

    def predict(self, X, raw=False):
        """Predicted log microseconds for feature rows X (HIP forward kernel)."""
        Xd = self._dev(X, raw)
        out = torch.empty(Xd.shape[0], dtype=torch.float32, device="cuda")
        self._ok(lib().fftjh_mlp_predict(self._h, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                         Xd.shape[0], self._stream()))
        torch.cuda.synchronize()
        return out.cpu().numpy()

    def gradients(self, X, t, raw=True):
        """Flat loss gradient and loss (HIP forward + backward kernels)."""
        Xd = self._dev(X, raw)
        td = torch.as_tensor(np.asarray(t, np.float32)).to("cuda")
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_float()
        self._ok(lib().fftjh_mlp_gradients(self._h, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_void_p(td.data_ptr()),
                                           Xd.shape[0], g.ctypes.data, ctypes.byref(loss), self._stream()))
        return g, loss.value

    def train_step(self, Xd, td, lr, want_loss=True):
        """One Adam step on device rows Xd (standardized) and targets td."""
        loss = ctypes.c_float()
        self._ok(lib().fftjh_mlp_train_step(self._h, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_void_p(td.data_ptr()),
                                            Xd.shape[0], ctypes.c_float(lr),
                                            ctypes.byref(loss) if want_loss else None, self._stream()))
        return loss.value if want_loss else None

    @staticmethod
    def standardization(X):
        """Feature means and spreads; the spread is floored at 0.05 (the
        features are O(1) by construction), so a feature that hardly varies
        in training cannot be amplified more than 20-fold at inference."""
        X = np.asarray(X, np.float32)
        return X.mean(0).astype(np.float32), np.maximum(X.std(0), 0.05).astype(np.float32)

    def fit(self, X, t, steps=6000, batch=256, lr=3e-3, seed=0, log=None, norm=None):
        """Adam on seeded mini-batches (HIP kernels); returns the loss curve
        sampled every 100 steps.  Sets the feature standardization first
        (from X, or `norm` = (mean, spread))."""
        X = np.asarray(X, np.float32)
        self.mu, self.sd = norm if norm is not None else self.standardization(X)
        tt = np.asarray(t, np.float64)
        self.log_range = (float(tt.min()), float(tt.max()))
        Xd = torch.from_numpy(self.normalize(X)).to("cuda")
        td = torch.as_tensor(np.asarray(t, np.float32)).to("cuda")
        g = torch.Generator(device="cpu").manual_seed(seed)
        curve = []
        bs = min(batch, Xd.shape[0])
        for s in range(steps):
            idx = torch.randint(0, Xd.shape[0], (bs,), generator=g).to("cuda")
            lr_s = lr * (0.05 ** (s / steps))
            want = s % 100 == 0 or s == steps - 1
            loss = self.train_step(Xd[idx].contiguous(), td[idx].contiguous(), lr_s, want_loss=want)
            if want:
                curve.append(loss)
                if log and s % 1000 == 0:
                    log(f"step {s} loss {loss:.4f}")
        torch.cuda.synchronize()
        return curve

    def save(self, path):
        np.savez(path, params=self.get_params(), mu=self.mu, sd=self.sd,
                 shape=np.array([self.in_features, self.hidden], np.int64),
                 log_range=np.array(self.log_range, np.float64), margin=np.array(self.margin, np.float64),
                 elems_range=np.array(self.elems_range, np.int64))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            m = cls(int(z["shape"][0]), int(z["shape"][1]))
            m.set_params(z["params"])
            m.mu, m.sd = z["mu"].astype(np.float32), z["sd"].astype(np.float32)
            if "log_range" in z.files:
                m.log_range = tuple(float(v) for v in z["log_range"])
            if "margin" in z.files:
                m.margin = float(z["margin"])
            if "elems_range" in z.files:
                m.elems_range = tuple(int(v) for v in z["elems_range"])
        return m


class Ensemble:
    """K cost models trained on bootstrap resamples (by length) of the
    training cases; predicts their mean.  Carries the margin a non-default
    pick must be predicted to gain and the n*batch range of the training
    cases."""

    def __init__(self, members):
        self.members = list(members)
        self.margin = 0.0
        self.elems_range = (0, 0)
        self.feature_range = None  # (lo, hi) per feature over the training rows
        lo = min(m.log_range[0] for m in self.members)
        hi = max(m.log_range[1] for m in self.members)
        self.log_range = (lo, hi)

    def predict(self, X):
        return np.mean([m.predict(X) for m in self.members], axis=0)

    def save(self, path):
        z = {"members": np.array(len(self.members), np.int64), "margin": np.array(self.margin, np.float64),
             "elems_range": np.array(self.elems_range, np.int64),
             "shape": np.array([self.members[0].in_features, self.members[0].hidden], np.int64)}
        if self.feature_range is not None:
            z["feature_lo"], z["feature_hi"] = self.feature_range
        for k, m in enumerate(self.members):
            z[f"m{k}_params"] = m.get_params()
            z[f"m{k}_mu"], z[f"m{k}_sd"] = m.mu, m.sd
            z[f"m{k}_log_range"] = np.array(m.log_range, np.float64)
        np.savez(path, **z)

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            inn, hid = int(z["shape"][0]), int(z["shape"][1])
            members = []
            for k in range(int(z["members"])):
                m = CostModel(inn, hid)
                m.set_params(z[f"m{k}_params"])
                m.mu, m.sd = z[f"m{k}_mu"].astype(np.float32), z[f"m{k}_sd"].astype(np.float32)
                m.log_range = tuple(float(v) for v in z[f"m{k}_log_range"])
                members.append(m)
            e = cls(members)
            e.margin = float(z["margin"])
            e.elems_range = tuple(int(v) for v in z["elems_range"])
            if "feature_lo" in z.files:
                e.feature_range = (z["feature_lo"].astype(np.float32), z["feature_hi"].astype(np.float32))
        return e


# ----------------------------------------------------------------- planners

def held_out_mask(data, every=5):
    """Every `every`-th size (by index) is held out of training."""
    sids = np.unique(data["size_id"])
    return np.isin(data["size_id"], sids[::every])


def _cases(data, mask):
    key = data["case_id"] if "case_id" in data else data["size_id"]
    for cid in np.unique(key[mask]):
        yield key == cid


def in_range(X, ranges):
    """Rows of X whose every feature lies within the training range
    (lo, hi); all rows when `ranges` is None."""
    X = np.atleast_2d(np.asarray(X, np.float32))
    if ranges is None:
        return np.ones(len(X), bool)
    lo, hi = ranges
    return np.all((X >= lo - 1e-6) & (X <= hi + 1e-6), axis=1)


def pick_index(pred, is_default, margin=0.0, eligible=None):
    """Row the learned planner takes among one case's candidates: the
    predicted fastest of the eligible rows (features inside the training
    range), unless its predicted time is not below the default's by the
    factor (1 - margin), or the default itself lies outside the training
    range -- then the default."""
    d = np.flatnonzero(is_default == 1)
    ok = np.ones(len(pred), bool) if eligible is None else np.asarray(eligible, bool)
    if d.size and not ok[d[0]]:
        return int(d[0])
    masked = np.where(ok, pred, np.inf)
    k = int(np.argmin(masked))
    if d.size and margin > 0 and not pred[k] < pred[d[0]] + math.log1p(-margin):
        return int(d[0])
    return k


def regret_report(data, pred, mask, margin=0.0, ranges=None):
    """Over the cases (length, batch) in `mask`: measured time of the
    learned pick and of the default plan, relative to the exhaustive
    planner's best."""
    rows = []
    elig = in_range(data["features"], ranges) if ranges is not None else np.ones(len(pred), bool)
    for m in _cases(data, mask):
        t = np.exp(data["log_us"][m].astype(np.float64))
        best = t.min()
        learned = t[pick_index(pred[m], data["is_default"][m], margin, elig[m])]
        d = data["is_default"][m] == 1
        default = t[d][0] if d.any() else t.max()
        rows.append((int(data["n"][m][0]), best, learned, default, int(m.sum())))
    r = np.array(rows, dtype=np.float64)
    reg_l = r[:, 2] / r[:, 1] - 1.0
    reg_d = r[:, 3] / r[:, 1] - 1.0
    return {
        "sizes": int(len(r)),
        "margin": float(margin),
        "candidates_per_size_mean": float(r[:, 4].mean()),
        "mean_regret_learned": float(reg_l.mean()),
        "mean_regret_default": float(reg_d.mean()),
        "median_regret_learned": float(np.median(reg_l)),
        "median_regret_default": float(np.median(reg_d)),
        "p90_regret_learned": float(np.quantile(reg_l, 0.9)),
        "p90_regret_default": float(np.quantile(reg_d, 0.9)),
        "geomean_speedup_learned_over_default": float(np.exp(np.mean(np.log(r[:, 3] / r[:, 2])))),
        "geomean_speedup_best_over_default": float(np.exp(np.mean(np.log(r[:, 3] / r[:, 1])))),
        "worst_learned_over_default": float(np.max(r[:, 2] / r[:, 3])),
        "cases_learned_5pct_slower": int(np.sum(r[:, 2] > 1.05 * r[:, 3])),
        "learned_picks_exhaustive_best": float(np.mean(r[:, 2] <= r[:, 1])),
        "default_is_exhaustive_best": float(np.mean(r[:, 3] <= r[:, 1])),
    }


MARGINS = (0.0, 0.02, 0.04, 0.06, 0.08, 0.1, 0.12, 0.15, 0.2, 0.25, 0.3)


def calibrate_margin(data, pred, mask, worst=1.05):
    """The margin, over the cases in `mask`, with the best geomean speedup
    over the default among those that leave no case slower than the default
    by more than `worst` (the largest margin if none does)."""
    best = None
    for mg in MARGINS:
        r = regret_report(data, pred, mask, mg)
        ok = r["worst_learned_over_default"] <= worst
        key = (ok, r["geomean_speedup_learned_over_default"] if ok else -mg)
        if best is None or key > best[0]:
            best = (key, mg)
    return best[1]


def load_dataset(name=DATASET, directory=DATA_DIR):
    with np.load(os.path.join(directory, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def model_available(name=DATASET, directory=DATA_DIR):
    """Whether a trained cost model is committed beside the dataset."""
    return os.path.exists(os.path.join(directory, name + "_model.npz"))


def load_model(name=DATASET, directory=DATA_DIR):
    """The committed cost model: an Ensemble, or a single CostModel from an
    older file."""
    path = os.path.join(directory, name + "_model.npz")
    with np.load(path, allow_pickle=False) as z:
        ensemble = "members" in z.files
    return Ensemble.load(path) if ensemble else CostModel.load(path)


class LearnedPlanner:
    """Picks, among a length's candidates, the one the cost model predicts
    fastest; lengths it has no candidates for fall back to the default."""

    def __init__(self, model=None):
        self.model = model if model is not None else load_model()
        self._picks = {}

    def rank(self, n, batch, sign=-1):
        """Candidates of length n ordered by predicted time: [(cand, us)]."""
        vc = valid_candidates(n, batch, sign)
        pred = self.model.predict(np.array([features(d, n, batch) for _, d in vc]))
        order = np.argsort(pred, kind="stable")
        return [(vc[k][0], float(np.exp(pred[k]))) for k in order]

    def pick(self, n, batch, sign=-1, margin=None):
        """(candidate, predicted us): the predicted fastest, unless it is not
        predicted faster than the default by the model's calibrated margin,
        then the default ({})."""
        vc = valid_candidates(n, batch, sign)
        X = np.array([features(d, n, batch) for _, d in vc])
        pred = self.model.predict(X)
        is_def = np.array([0 if c else 1 for c, _ in vc])
        elig = in_range(X, getattr(self.model, "feature_range", None))
        k = pick_index(pred, is_def, self.model.margin if margin is None else margin, elig)
        return vc[k][0], float(np.exp(pred[k]))

    def pick_cached(self, n, batch, sign=-1):
        """pick(n, batch, sign)'s candidate, remembered per (n, batch, sign)."""
        key = (int(n), int(batch), int(sign))
        if key not in self._picks:
            self._picks[key] = self.pick(n, batch, sign)[0]
        return self._picks[key]

    def choose(self, n, batch, sign=-1, margin=None):
        """Builds only the picked candidate: (plan, cand, us)."""
        cand, us = self.pick(n, batch, sign, margin)
        return make_plan(n, batch, cand, sign=sign), cand, us

    def confident(self, n, batch):
        """Whether (n, batch) lies where the model was trained: 13-smooth
        lengths from 1024 to 2^24 at n*batch within a factor 2 of the
        training cases' range."""
        lo, hi = self.model.elems_range
        return (smooth13(n) and 1024 <= n <= 1 << 24 and hi > 0
                and lo / 2 <= n * batch <= 2 * hi)

    # resampled lengths: the m-point transform's cost times this factor for
    # the deconvolution and interpolation edges (the model saw no resampled
    # schedules: the training lengths are 13-smooth)
    RESAMPLE_FACTOR = 1.5

    def per_transform_us(self, sizes):
        """Predicted microseconds per transform for each length (default
        plans at the training batch) -- per-item costs for the ragged
        planner's workgroup balance.  A non-smooth length is priced as the
        smooth m its plan resamples onto; predictions are clamped to the
        range of the training timings."""
        sizes = [int(s) for s in sizes]
        core, factor = [], []
        for n in sizes:
            m = n
            if not smooth13(n):
                d = schedule(n, 1, {})
                m = int(d.split(" m=")[1].split()[0]) if " m=" in d else n
            core.append(m)
            factor.append(self.RESAMPLE_FACTOR if m != n else 1.0)
        feats = [features(schedule(m, batch_for(m), {}), m, batch_for(m)) for m in core]
        pred = self.model.predict(np.array(feats)).astype(np.float64)
        lo, hi = self.model.log_range
        pred = np.clip(pred, lo, hi)
        per = np.exp(pred) / np.array([batch_for(m) for m in core], np.float64)
        return per * np.array(factor)


_LEARNED = None


def default_learned():
    """The planner over the cost model trained on this GPU's timings."""
    global _LEARNED
    if _LEARNED is None:
        _LEARNED = LearnedPlanner()
    return _LEARNED


def exhaustive_choose(n, batch, sign=-1, x=None, y=None, dtype=0):
    """Measure every candidate; return (plan, candidate, microseconds)."""
    if x is None:
        x = torch.randn(n * batch, dtype=torch.complex64 if dtype == 0 else torch.complex128, device="cuda")
        y = torch.empty_like(x)
    best = (None, None, float("inf"))
    for cand, _ in valid_candidates(n, batch, sign, dtype):
        p = make_plan(n, batch, cand, sign=sign, dtype=dtype)
        us = time_plan(p, y, x) * 1e3
        if us < best[2]:
            best = (p, cand, us)
    return best


# ------------------------------------------------------------------ script

MEMBERS = 5


def train(data, steps=6000, members=MEMBERS, log=None):
    """Fit the cost model on the training cases (every length but the
    held-out fifth): `members` networks, each on a bootstrap resample of the
    training lengths (seeded).  The margin is the median absolute error, out
    of bag, of the predicted log time ratio of a candidate to its default,
    rounded up to 0.01 -- a pick must be predicted faster than the default
    by more than the model's typical error on that ratio.  Returns
    (ensemble, report)."""
    test = held_out_mask(data)
    sids = np.unique(data["size_id"][~test])
    train_rows = np.flatnonzero(~test)
    norm = CostModel.standardization(data["features"][train_rows])
    models, inbag = [], []
    for k in range(members):
        r = np.random.default_rng(100 + k)
        boot = r.choice(sids, size=len(sids), replace=True)
        cnt = np.bincount(boot, minlength=int(data["size_id"].max()) + 1)
        rows = np.repeat(train_rows, cnt[data["size_id"][train_rows]])
        m = CostModel(seed=k)
        m.fit(data["features"][rows], data["log_us"][rows], steps=steps, seed=k, log=log, norm=norm)
        models.append(m)
        inbag.append(cnt > 0)
        if log:
            log(f"member {k + 1}/{members} trained")
    preds = np.array([m.predict(data["features"]) for m in models])
    # out-of-bag error of the log ratio to the default, per training case
    errs = []
    key = data["case_id"] if "case_id" in data else data["size_id"]
    for cid in np.unique(key[~test]):
        rows = np.flatnonzero(key == cid)
        out = [k for k in range(members) if not inbag[k][data["size_id"][rows[0]]]]
        d = rows[data["is_default"][rows] == 1]
        if not out or d.size == 0:
            continue
        po = preds[out][:, rows].mean(0)
        pd = preds[out][:, d[0]].mean()
        t = data["log_us"][rows].astype(np.float64)
        errs.extend(np.abs((po - pd) - (t - data["log_us"][d[0]])).tolist())
    ens = Ensemble(models)
    ens.margin = float(math.ceil(100.0 * float(np.median(errs))) / 100.0) if errs else 0.05
    elems = data["n"] * data["batch"]
    ens.elems_range = (int(elems.min()), int(elems.max()))
    Xt = data["features"][train_rows]
    ens.feature_range = (Xt.min(0).astype(np.float32), Xt.max(0).astype(np.float32))
    fr = ens.feature_range
    pred = preds.mean(0)
    rmse = lambda m: float(np.sqrt(np.mean((pred[m] - data["log_us"][m]) ** 2)))  # noqa: E731
    report = {"timings": int(len(pred)), "members": members, "train_rmse_log": rmse(~test),
              "test_rmse_log": rmse(test), "oob_median_abs_log_ratio_error": float(np.median(errs)),
              "margin": ens.margin,
              "held_out": regret_report(data, pred, test, ens.margin, fr),
              "held_out_no_margin": regret_report(data, pred, test, 0.0, fr),
              "train": regret_report(data, pred, ~test, ens.margin, fr)}
    return ens, report


def main(argv=None):
    ap = argparse.ArgumentParser(description="collect plan timings on the GPU and train the cost model")
    ap.add_argument("--out", required=True)
    ap.add_argument("--sizes", type=int, default=400)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--data", default=None, help="reuse a collected dataset (.npz) instead of measuring")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)

    def log(s):
        print(s, flush=True)

    t0 = time.time()
    if a.data:
        with np.load(a.data, allow_pickle=False) as z:
            data = {k: z[k] for k in z.files}
    else:
        sizes = training_sizes(a.sizes)
        data = collect(sizes, log=log)
        np.savez_compressed(os.path.join(a.out, DATASET + ".npz"), **data)
    log(f"{len(data['log_us'])} timings over {len(np.unique(data['size_id']))} sizes ({time.time() - t0:.0f} s)")
    model, report = train(data, steps=a.steps, log=log)
    model.save(os.path.join(a.out, DATASET + "_model.npz"))
    with open(os.path.join(a.out, DATASET + "_report.json"), "w") as fh:
        json.dump(report, fh, indent=1)
    log(json.dumps(report["held_out"]))
    log(f"done in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
