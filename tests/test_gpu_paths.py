"""GPU parity: prime (resampling) plans, real plans, N-d plans, against the
FP64 oracle on the same seeded inputs (runtests.jl:57-83, 233-262, 330-389),
plus the benchmark's prime, real and 3-D shapes at full size."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, cplx_np, inputs_c, real_np, rel_l2, to_dev, to_host, tol_for, torch_dtype

pytestmark = pytest.mark.gpu

# ------------------------------------------------------------- resampling (A9)

PRIMES = [17, 19, 23, 34, 38, 51, 97, 101, 289, 331, 361, 997, 1026, 1234, 2050, 8191, 10006, 10007, 65542]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", PRIMES)
def test_resampled(prec, n):
    x = inputs_c(n, n, prec)
    xd = to_dev(x)
    y = to_host(F.plan_fft(torch_dtype(prec), n)(xd))
    ref = O.fft(x.astype(np.complex128))
    assert rel_l2(y, ref) <= tol_for(prec, n)
    yi = to_host(F.plan_ifft(torch_dtype(prec), n)(to_dev(y)))
    assert rel_l2(yi, x) <= 2 * tol_for(prec, n)
    yb = to_host(F.plan_bfft(torch_dtype(prec), n)(to_dev(y)))
    assert rel_l2(yb, n * x.astype(np.complex128)) <= 2 * tol_for(prec, n)


@pytest.mark.parametrize("n", [65537, 1000003])
def test_config3_primes(n):
    batch = 4
    x = inputs_c(0, n * batch, "f32").reshape(batch, n)
    y = to_host(F.plan_fft(torch.float32, n, batch=batch)(to_dev(x))).reshape(batch, n)
    rows = [0, 3]
    ref = O.fft(x[rows].astype(np.complex128))
    assert rel_l2(y[rows], ref) <= tol_for("f32", n)
    # and against the reference's own route for these primes (HvdhPlan)
    assert rel_l2(y[rows], O.hvdh(x[rows].astype(np.complex128))) <= tol_for("f32", n)


# ------------------------------------------------------------------ real (A10)

REAL_N = [1, 2, 3, 4, 5, 6, 8, 11, 16, 22, 34, 64, 100, 255, 256, 257, 1024, 768, 1000, 4096, 2**14,
          3 * 2**10, 7 * 2**8, 2**18 + 2, 10007 * 2]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", REAL_N)
def test_rfft_irfft(prec, n):
    x = F.rng.normal(n, n).astype(real_np(prec))
    T = torch_dtype(prec)
    y = to_host(F.plan_rfft(T, n)(to_dev(x)))
    assert y.shape == (n // 2 + 1,)
    ref = O.rfft(x.astype(np.float64))
    assert rel_l2(y, ref) <= tol_for(prec, n)
    xr = to_host(F.plan_irfft(T, n)(to_dev(y)))
    assert rel_l2(xr, x) <= 2 * tol_for(prec, n)
    yb = to_host(F.plan_brfft(T, n)(to_dev(ref.astype(cplx_np(prec)))))
    assert rel_l2(yb, O.irfft(ref, n, normalized=False)) <= tol_for(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [64, 128, 512, 2048, 8192, 16384])
@pytest.mark.parametrize("batch", [1, 3, 37])
def test_real_rows_batched(prec, n, batch):
    # batched rows: power-of-two half lengths take the one-launch kernel with
    # the real edges fused (partial last groups included); the rest the
    # two-launch path
    T = torch_dtype(prec)
    x = F.rng.normal(n + batch, n * batch).astype(real_np(prec)).reshape(batch, n)
    p = F.plan_rfft(T, n, batch=batch)
    Xd = p(to_dev(x))
    X = to_host(Xd).reshape(batch, n // 2 + 1)
    ref = O.rfft(x.astype(np.float64))
    assert rel_l2(X, ref) <= tol_for(prec, n)
    for b in (0, batch - 1):
        assert rel_l2(X[b], ref[b]) <= tol_for(prec, n)
    back = to_host(F.plan_irfft(T, n, batch=batch)(Xd)).reshape(batch, n)
    assert rel_l2(back, x) <= 2 * tol_for(prec, n)
    # brfft of an arbitrary spectrum (imaginary DC / Nyquist parts enter the
    # even-length pre-pass, as in the reference)
    Z = F.rng.uniform_complex(n, (n // 2 + 1) * batch).astype(cplx_np(prec)).reshape(batch, n // 2 + 1)
    yb = to_host(F.plan_brfft(T, n, batch=batch)(to_dev(Z))).reshape(batch, n)
    assert rel_l2(yb, O.irfft(Z.astype(np.complex128), n, normalized=False)) <= tol_for(prec, n)
    if "real=" in p.describe():
        assert p.describe().count("\n  ") == 1


# Multi-pass half lengths: the real edges ride on the last (R2C) or first
# (C2R) pass of the half-length transform, a workgroup holding 16 of its
# transforms T and their mirrors M - T (T = 0 and T = M/2 mirror themselves
# and share group 0).  Every row checked, the inverse on an arbitrary
# spectrum (imaginary DC and Nyquist parts) too.
PAIRED = [("f32", 2**16, 2), ("f32", 2**17, 3), ("f32", 2**18, 1), ("f32", 3 * 2**17, 2), ("f32", 2**20, 2),
          ("f32", 5 * 2**15, 2), ("f64", 2**16, 2), ("f64", 2**19, 1), ("f64", 2**21, 2)]


@pytest.mark.parametrize("n,batch", [(128, 16), (256, 48), (512, 32), (256, 4096)])
def test_lean_real_rows(n, batch):
    """FP32 real rows of 2N = 128/256/512 reals in groups of 16 on the lean
    kernel: the R2C twist at its rows-out store, the C2R pre-pass at its
    rows-in load; every row against the oracle, an arbitrary spectrum
    (imaginary DC and Nyquist parts) through both inverses."""
    h = n // 2
    x = F.rng.normal(n + batch, n * batch).astype(np.float32).reshape(batch, n)
    p = F.plan_rfft(torch.float32, n, batch=batch)
    pi = F.plan_irfft(torch.float32, n, batch=batch)
    pb = F.plan_brfft(torch.float32, n, batch=batch)
    assert "kernel=lean rows_in rows_out" in p.describe() and "real=r2c" in p.describe()
    assert "real=c2r" in pi.describe() and "kernel=lean" in pi.describe()
    Xd = p(to_dev(x))
    X = to_host(Xd).reshape(batch, h + 1)
    ref = O.rfft(x.astype(np.float64))
    for b in range(0, batch, max(1, batch // 64)):
        assert rel_l2(X[b], ref[b]) <= tol_for("f32", n), b
    assert rel_l2(to_host(pi(Xd)).reshape(batch, n), x) <= 2 * tol_for("f32", n)
    Z = F.rng.uniform_complex(n + 11, (h + 1) * batch).astype(np.complex64).reshape(batch, h + 1)
    Zd = to_dev(Z)
    yi = to_host(pi(Zd)).reshape(batch, n)
    yb = to_host(pb(Zd)).reshape(batch, n)
    for b in range(0, batch, max(1, batch // 64)):
        zr = Z[b].astype(np.complex128)
        assert rel_l2(yi[b], O.irfft(zr, n)) <= tol_for("f32", n), b
        assert rel_l2(yb[b], O.irfft(zr, n, normalized=False)) <= tol_for("f32", n), b
    assert np.array_equal(to_host(Zd), Z)


@pytest.mark.parametrize("prec,n,batch", PAIRED)
def test_real_multipass_paired(prec, n, batch):
    T = torch_dtype(prec)
    h = n // 2
    x = F.rng.normal(n + 7, n * batch).astype(real_np(prec)).reshape(batch, n)
    p = F.plan_rfft(T, n, batch=batch)
    pi = F.plan_irfft(T, n, batch=batch)
    pb = F.plan_brfft(T, n, batch=batch)
    assert "paired" in p.describe() or "paired" in pi.describe()
    Xd = p(to_dev(x))
    X = to_host(Xd).reshape(batch, h + 1)
    ref = O.rfft(x.astype(np.float64))
    for b in range(batch):
        assert rel_l2(X[b], ref[b]) <= tol_for(prec, n), b
    assert rel_l2(to_host(pi(Xd)).reshape(batch, n), x) <= 2 * tol_for(prec, n)
    Z = F.rng.uniform_complex(n + 9, (h + 1) * batch).astype(cplx_np(prec)).reshape(batch, h + 1)
    Zd = to_dev(Z)
    yi = to_host(pi(Zd)).reshape(batch, n)
    yb = to_host(pb(Zd)).reshape(batch, n)
    for b in range(batch):
        zr = Z[b].astype(np.complex128)
        assert rel_l2(yi[b], O.irfft(zr, n)) <= tol_for(prec, n), b
        assert rel_l2(yb[b], O.irfft(zr, n, normalized=False)) <= tol_for(prec, n), b
    assert np.array_equal(to_host(Zd), Z)   # the inverse leaves its input intact


def test_config3_rfft_2pow22_x4():
    n, batch = 1 << 22, 4
    x = F.rng.normal(0, n * batch).astype(np.float32).reshape(batch, n)
    p = F.plan_rfft(torch.float32, n, batch=batch)
    yd = p(to_dev(x))
    y = to_host(yd)
    assert y.shape == (batch, n // 2 + 1)
    for b in (0, batch - 1):
        assert rel_l2(y[b], O.rfft(x[b].astype(np.float64))) <= tol_for("f32", n)
    back = to_host(F.plan_irfft(torch.float32, n, batch=batch)(yd))
    assert rel_l2(back, x) <= 2 * tol_for("f32", n)


# ------------------------------------------------------------------- N-d (A11)

SHAPES = [(64, 48), (256, 256), (33, 17), (16, 11), (11, 16), (8, 1), (1, 8), (128, 3, 5), (16, 16, 16),
          (12, 10, 14), (32, 64), (60, 50), (24, 40, 3), (1024, 512), (4, 2**12), (2**12, 4), (256, 33, 17), (3, 2**17), (2**17, 2), (19, 2**15),
          (64, 4096), (3, 8192), (16, 5, 4096), (8, 6000), (7, 12288), (4096, 4096)]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dims", SHAPES)
def test_nd(prec, dims):
    n = int(np.prod(dims))
    x = inputs_c(n, n, prec).reshape(tuple(reversed(dims)))
    xd = to_dev(x)
    p = F.plan_fft(xd)
    assert F.size(p) == tuple(reversed(dims))
    y = to_host(p * xd)
    ref = O.ndfft(x.astype(np.complex128).ravel(), list(dims)).reshape(x.shape)
    tol = 1e-13 if prec == "f64" else 200 * np.finfo(np.float32).eps   # runtests.jl:332 tier
    assert rel_l2(y, ref) <= tol
    back = to_host(F.plan_ifft(xd) * to_dev(y))
    assert rel_l2(back, x) <= 2 * tol
    yb = to_host(F.plan_bfft(xd) * to_dev(y))
    assert rel_l2(yb, n * x.astype(np.complex128)) <= 2 * tol


@pytest.mark.parametrize("dims", [(64, 16), (128, 48), (256, 256, 3), (16, 256, 4), (64, 32, 48), (256, 17, 16),
                                  (33, 64, 5), (17, 256, 3), (129, 128, 2), (20, 64, 64)])
def test_lean_rows_and_columns(dims):
    """FP32 plans whose row pass runs on the lean kernel with rows in and out
    (groups of 16 rows) and whose strided axes run its column form (16
    adjacent columns), against the oracle, both directions and in place."""
    n = int(np.prod(dims))
    x = inputs_c(n + 3, n, "f32").reshape(tuple(reversed(dims)))
    xd = to_dev(x)
    p = F.plan_fft(xd)
    assert "kernel=lean" in p.describe()
    ref = O.ndfft(x.astype(np.complex128).ravel(), list(dims)).reshape(x.shape)
    y = to_host(p * xd)
    tol = 200 * np.finfo(np.float32).eps
    assert rel_l2(y, ref) <= tol
    pi = F.plan_ifft(xd)
    back = to_host(pi * to_dev(y))
    assert rel_l2(back, x) <= 2 * tol
    z = to_dev(x)
    p(z, z)
    assert np.array_equal(to_host(z), y)


@pytest.mark.parametrize("n,batch", [(64, 16 * 37), (128, 16 * 9), (256, 16 * 5), (256, 65536)])
def test_lean_row_batches(n, batch):
    x = inputs_c(n * 7 + batch, n * batch, "f32")
    xd = to_dev(x)
    p = F.plan_fft(torch.float32, n, batch=batch)
    assert "rows_in rows_out" in p.describe()
    y = to_host(p * xd)
    rows = list(range(0, batch, max(1, batch // 64))) + [batch - 1]
    xr = x.reshape(batch, n)
    ref = O.fft(xr[rows].astype(np.complex128), -1, 1.0, "f64")
    assert rel_l2(y.reshape(batch, n)[rows], ref) <= tol_for("f32", n)
    back = to_host(F.plan_ifft(torch.float32, n, batch=batch) * to_dev(y))
    assert rel_l2(back, x) <= 2 * tol_for("f32", n)


# N-d real plans (runtests.jl:358-370 on the same shapes, plus odd, ragged,
# non-smooth and long halved / strided axes): torch shape reversed(dims),
# halved along the last axis

REAL_SHAPES = SHAPES[:13] + [(9, 4, 3), (17, 33), (1, 1), (20, 1, 1), (2**15, 4), (3, 2**15), (1000, 6), (7, 5, 11)]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dims", REAL_SHAPES)
def test_nd_real(prec, dims):
    shape = tuple(reversed(dims))
    n1 = dims[0]
    x = F.rng.normal(len(dims) * 7 + n1, int(np.prod(dims))).astype(real_np(prec)).reshape(shape)
    xd = to_dev(x)
    pr = F.plan_rfft(xd)
    assert F.size(pr) == shape
    Xd = pr * xd
    X = to_host(Xd)
    assert X.shape == shape[:-1] + (n1 // 2 + 1,)
    tol = 1e-13 if prec == "f64" else 200 * np.finfo(np.float32).eps   # runtests.jl:332 tier
    ref = O.ndrfft(x.astype(np.float64))
    assert rel_l2(X, ref) <= tol
    pbr = F.plan_brfft(Xd, n1)
    assert F.size(pbr) == X.shape
    bref = O.ndirfft(X.astype(np.complex128), n1, normalized=False)
    assert rel_l2(to_host(pbr * Xd), bref) <= tol
    back = to_host(F.plan_irfft(Xd, n1) * Xd)
    assert rel_l2(back, x) <= 2 * tol
    assert np.array_equal(to_host(Xd), X)   # the inverse leaves its input intact


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_nd_real_arbitrary_spectrum(prec):
    # a complex input that is not the spectrum of a real array: the inverse
    # follows the reference's pre-pass (imaginary parts of the DC and Nyquist
    # bins enter the even-length pre-pass), as the oracle restates it
    for shape, n1 in (((6, 5, 9), 16), ((4, 7, 6), 11)):
        X = F.rng.uniform_complex(n1, int(np.prod(shape))).astype(cplx_np(prec)).reshape(shape)
        tol = 1e-13 if prec == "f64" else 200 * np.finfo(np.float32).eps
        for normalized in (True, False):
            p = (F.plan_irfft if normalized else F.plan_brfft)(to_dev(X), n1)
            y = to_host(p * to_dev(X))
            assert y.shape == shape[:-1] + (n1,)
            ref = O.ndirfft(X.astype(np.complex128), n1, normalized=normalized)
            assert rel_l2(y, ref) <= tol


def test_nd_real_256cubed_properties():
    # full size: the half spectrum is the prefix of the complex plan's
    # spectrum of the same data, and irfft inverts it
    shape = (256, 256, 256)
    x = F.rng.normal(3, 256 ** 3).astype(np.float32).reshape(shape)
    xd = to_dev(x)
    Xd = F.rfft(xd)
    Y = F.fft(xd.to(torch.complex64))
    tol = 200 * np.finfo(np.float32).eps
    assert rel_l2(to_host(Xd), to_host(Y[..., :129])) <= tol
    assert rel_l2(to_host(F.irfft(Xd, 256)), x) <= 2 * tol
    assert rel_l2(to_host(Xd), O.ndrfft(x.astype(np.float64))) <= tol


def test_config3_nd_256cubed():
    dims = (256, 256, 256)
    n = 256 ** 3
    x = inputs_c(0, n, "f32").reshape(dims)
    xd = to_dev(x)
    y = to_host(F.fft(xd))
    ref = O.ndfft(x.astype(np.complex128).ravel(), list(dims)).reshape(dims)
    assert rel_l2(y, ref) <= 200 * np.finfo(np.float32).eps


# ------------------------------------------------- composite route (F3)

COMPOSITE_N = [2 * 10007, 6 * 1009, 64 * 10007, 17 * 4096, 48 * 1009, 32 * 17, 12 * 65537, 2 * 1000003]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", COMPOSITE_N)
def test_composite_route(prec, n):
    # n = p*s, s the 13-smooth part: one exact step over s around the
    # resampled length p (composite.jl:1-101), against the oracle, with both
    # routes agreeing
    batch = 2
    x = inputs_c(n + 1, n * batch, prec).reshape(batch, n)
    T = torch_dtype(prec)
    p = F.plan_fft(T, n, batch=batch, route="composite")
    d = p.describe()
    assert d.splitlines()[1].strip().startswith("pass x->y") and "rs_deconv" in d
    y = to_host(p(to_dev(x))).reshape(batch, n)
    ref = O.fft(x.astype(np.complex128))
    assert rel_l2(y, ref) <= tol_for(prec, n)
    assert rel_l2(y[1], ref[1]) <= tol_for(prec, n)
    yr = to_host(F.plan_fft(T, n, batch=batch, route="resample")(to_dev(x))).reshape(batch, n)
    assert rel_l2(y, yr) <= 2 * tol_for(prec, n)
    back = to_host(F.plan_ifft(T, n, batch=batch, route="composite")(to_dev(y.astype(x.dtype)))).reshape(batch, n)
    assert rel_l2(back, x) <= 2 * tol_for(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [48 * 1009, 100 * 1009, 1000 * 1009, 1024 * 1031])
def test_default_route_nonsmooth(prec, n):
    # lengths the planner sends down the composite route by default
    x = inputs_c(n + 3, n, prec)
    p = F.plan_fft(torch_dtype(prec), n)
    assert p.describe().splitlines()[1].strip().startswith("pass x->y")
    y = to_host(p(to_dev(x)))
    assert rel_l2(y, O.fft(x.astype(np.complex128))) <= tol_for(prec, n)
