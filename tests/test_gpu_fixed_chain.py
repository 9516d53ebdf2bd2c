"""The compile-time chain kernel on non-power-of-two lengths (100 = 10 x 10,
10 elements per thread, 32 transforms per workgroup; 81 = 9 x 9, 27 elements
per thread, 64 transforms per workgroup, on passes with strided loads only):
rows with partial last groups,
strided N-d axes, and the multi-pass lengths built from them, against the
oracle.  Each case asserts that the chain kernel runs."""
import numpy as np
import pytest

from oracle import oracle as O
from conftest import F, inputs_c, rel_l2, to_dev, to_host, tol_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,n,batch", [("f32", 100, 1), ("f32", 100, 7), ("f32", 100, 33), ("f32", 100, 1000),
                                          ("f32", 10**6, 1), ("f32", 10**6, 3), ("f32", 3**11, 5), ("f32", 3**12, 2),
                                          ("f64", 3**11, 3), ("f64", 3**12, 1)])
def test_chain_rows_and_multipass(prec, n, batch):
    import torch
    T = torch.float32 if prec == "f32" else torch.float64
    if n == 10**6:
        # (the library's own 10^6 plan runs the square lean kernel; a tuned
        # split of the same factors keeps the chain kernel)
        t = F._lib.make_tuning([100, 100, 100], (), 0)
        p = F.FFTPlan(0, (n,), -1, 1.0, batch, tuning=t)
    else:
        p = F.plan_fft(T, n, batch=batch, planner="heuristic")
    assert "kernel=chain" in p.describe()
    x = inputs_c(n + batch, n * batch, prec).reshape(batch, n)
    y = to_host(p(to_dev(x))).reshape(batch, n)
    ref = O.fft(x.astype(np.complex128))
    for b in range(batch):
        assert rel_l2(y[b], ref[b]) <= tol_for(prec, n), b
    back = to_host(F.plan_ifft(T, n, batch=batch)(to_dev(y))).reshape(batch, n)
    assert rel_l2(back, x) <= 2 * tol_for(prec, n)


@pytest.mark.parametrize("dims", [(100, 100), (40, 100), (100, 3, 7), (7, 9, 100), (100, 100, 100), (81, 81),
                                  (64, 81), (81, 81, 81)])
def test_chain_nd_axes(dims):
    n = int(np.prod(dims))
    x = inputs_c(n + 3, n, "f32").reshape(tuple(reversed(dims)))
    xd = to_dev(x)
    p = F.plan_fft(xd)
    assert "kernel=chain" in p.describe()
    y = to_host(p * xd)
    ref = O.ndfft(x.astype(np.complex128).ravel(), list(dims)).reshape(x.shape)
    assert rel_l2(y, ref) <= 200 * np.finfo(np.float32).eps   # runtests.jl:332 tier
    back = to_host(F.plan_ifft(xd) * to_dev(y))
    assert rel_l2(back, x) <= 400 * np.finfo(np.float32).eps
