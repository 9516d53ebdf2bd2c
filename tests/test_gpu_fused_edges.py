"""Sweeps over the fused edge paths against the oracle, every row of every
batch checked:

* real edges on the last (R2C) / first (C2R) pass of a multi-pass half
  length, with power-of-two and other pair counts M (16 transforms T and
  their mirrors M - T per workgroup), forward, irfft and brfft of an
  arbitrary spectrum;
* the resampling deconvolution at the load of the m-point transform's first
  pass, 1-D batched and along the contiguous axis of an N-d plan.

Each case asserts from the plan's schedule that it takes the fused path."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, cplx_np, inputs_c, real_np, rel_l2, to_dev, to_host, tol_for, torch_dtype

pytestmark = pytest.mark.gpu

# half lengths h = n/2 = M * N1 with M a multiple of 32: 3*2^15 .. 9*2^16
REAL_PAIRED = [3 * 2**16, 5 * 2**16, 9 * 2**15, 7 * 2**15, 15 * 2**14, 3 * 2**18, 15 * 2**16, 9 * 2**17, 2**19]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", REAL_PAIRED)
def test_real_paired_sweep(prec, n):
    T = torch_dtype(prec)
    h = n // 2
    batch = 3 if n < 2**19 else 2
    p, pi, pb = F.plan_rfft(T, n, batch=batch), F.plan_irfft(T, n, batch=batch), F.plan_brfft(T, n, batch=batch)
    assert "paired" in p.describe() or "paired" in pi.describe()
    x = F.rng.normal(n + 11, n * batch).astype(real_np(prec)).reshape(batch, n)
    X = to_host(p(to_dev(x))).reshape(batch, h + 1)
    ref = O.rfft(x.astype(np.float64))
    for b in range(batch):
        assert rel_l2(X[b], ref[b]) <= tol_for(prec, n), b
    Z = F.rng.uniform_complex(n + 13, (h + 1) * batch).astype(cplx_np(prec)).reshape(batch, h + 1)
    Zd = to_dev(Z)
    yi = to_host(pi(Zd)).reshape(batch, n)
    yb = to_host(pb(Zd)).reshape(batch, n)
    for b in range(batch):
        zr = Z[b].astype(np.complex128)
        assert rel_l2(yi[b], O.irfft(zr, n)) <= tol_for(prec, n), b
        assert rel_l2(yb[b], O.irfft(zr, n, normalized=False)) <= tol_for(prec, n), b


# primes and non-smooth lengths whose m-point transform starts with a
# power-of-two pass (m = 2^k * 9 and friends)
RS_FUSED = [65537, 2 * 10007, 1009, 100003, 524309]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", RS_FUSED)
def test_resample_fused_deconv(prec, n):
    batch = 3
    T = torch_dtype(prec)
    p = F.plan_fft(T, n, batch=batch)
    assert " deconv n=" in p.describe() and "rs_deconv" not in p.describe()
    x = inputs_c(n + 5, n * batch, prec).reshape(batch, n)
    y = to_host(p(to_dev(x))).reshape(batch, n)
    ref = O.fft(x.astype(np.complex128))
    for b in range(batch):
        assert rel_l2(y[b], ref[b]) <= tol_for(prec, n), b
    yi = to_host(F.plan_ifft(T, n, batch=batch)(to_dev(y))).reshape(batch, n)
    assert rel_l2(yi, x) <= 2 * tol_for(prec, n)


@pytest.mark.parametrize("dims", [(1009, 8), (17, 4099), (65537, 2)])
def test_resample_fused_deconv_nd(dims):
    # the contiguous axis (dims[0]) resampled with the deconvolution fused;
    # a strided non-smooth axis keeps its own deconvolution launch
    n = int(np.prod(dims))
    x = inputs_c(n + 1, n, "f32").reshape(tuple(reversed(dims)))
    xd = to_dev(x)
    p = F.plan_fft(xd)
    assert f" deconv n={dims[0]} " in p.describe()
    y = to_host(p * xd)
    ref = O.ndfft(x.astype(np.complex128).ravel(), list(dims)).reshape(x.shape)
    assert rel_l2(y, ref) <= 200 * np.finfo(np.float32).eps   # runtests.jl:332 tier
    back = to_host(F.plan_ifft(xd) * to_dev(y))
    assert rel_l2(back, x) <= 400 * np.finfo(np.float32).eps
