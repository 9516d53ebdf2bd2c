"""GPU parity on the rest of the reference's threaded test matrix and on the
plans the other suites do not reach: the 1-D lengths of threading.jl:80-81
and the both-axes-non-smooth shape of threading.jl:124, in both precisions;
C2R 2^22 x 4 compared row by row with the oracle's irfft; the
2^24 transform forced onto the 4096 x 4096 two-pass split; FP64 plans large
enough for the streaming-hint kernel variants."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, cplx_np, inputs_c, rel_l2, to_dev, to_host, tol_for, torch_dtype

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [1441440, 3 * 2**21, 2**22])
def test_threading_lengths(prec, n):
    # threading.jl:80-81 (1,441,440 = 2^5 3^2 5 7 11 13)
    print(f"c2c {prec} n={n}", flush=True)
    x = inputs_c(n % 1000, n, prec)
    xd = to_dev(x)
    y = to_host(F.plan_fft(torch_dtype(prec), n)(xd))
    x64 = x.astype(np.complex128)
    assert rel_l2(y, O.fft(x64)) <= tol_for(prec, n)
    yi = to_host(F.plan_ifft(torch_dtype(prec), n)(xd))
    assert rel_l2(yi, O.fft(x64, sign=1, scale=1.0 / n)) <= tol_for(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_nd_both_axes_nonsmooth(prec):
    # threading.jl:124: (513, 255), 513 = 3^3 19 and 255 = 3 5 17 both
    # resampled on their axes (the first contiguous, the second strided)
    dims = (513, 255)
    n = dims[0] * dims[1]
    print(f"nd {prec} dims={dims}", flush=True)
    x = inputs_c(124, n, prec).reshape(tuple(reversed(dims)))
    xd = to_dev(x)
    p = F.plan_fft(xd)
    y = to_host(p * xd)
    ref = O.ndfft(x.astype(np.complex128).ravel(), list(dims)).reshape(x.shape)
    tol = 1e-13 if prec == "f64" else 200 * np.finfo(np.float32).eps   # runtests.jl:332 tier
    assert rel_l2(y, ref) <= tol
    back = to_host(F.plan_ifft(xd) * to_dev(y))
    assert rel_l2(back, x) <= 2 * tol


def test_config3_c2r_2pow22_x4_rows():
    # C2R 2^22 x 4: every row against the oracle's irfft (and
    # brfft), on a spectrum that is not the transform of real data
    n, batch = 1 << 22, 4
    h = n // 2 + 1
    print(f"c2r f32 n={n} batch={batch}", flush=True)
    Z = F.rng.uniform_complex(22, h * batch).astype(np.complex64).reshape(batch, h)
    Zd = to_dev(Z)
    yi = to_host(F.plan_irfft(torch.float32, n, batch=batch)(Zd)).reshape(batch, n)
    yb = to_host(F.plan_brfft(torch.float32, n, batch=batch)(Zd)).reshape(batch, n)
    for b in range(batch):
        zr = Z[b].astype(np.complex128)
        assert rel_l2(yi[b], O.irfft(zr, n)) <= tol_for("f32", n), b
        assert rel_l2(yb[b], O.irfft(zr, n, normalized=False)) <= tol_for("f32", n), b
    assert np.array_equal(to_host(Zd), Z)


def test_2pow24_forced_4096x4096_split():
    # the four-step split of 2^24 into 4096 x 4096,
    # forced through the tuned entry point, against the oracle
    n = 1 << 24
    print("c2c f32 n=2^24 splits=4096x4096", flush=True)
    L = F._lib
    x = inputs_c(24, n, "f32")
    p = F.FFTPlan(L.F32, (n,), -1, 1.0, 1, None, L.make_tuning((4096, 4096)))
    d = p.describe()
    assert d.count("n=4096 ") == 2 and "launches=2" in d
    y = to_host(p(to_dev(x)))
    assert rel_l2(y, O.fft(x.astype(np.complex128))) <= tol_for("f32", n)


@pytest.mark.parametrize("n,batch,tag", [(4096, 8192, "stream=inout"), (1 << 24, 1, "stream=in")])
def test_f64_streaming_hint_variants(n, batch, tag):
    # FP64 plans of 384 MiB and more take the non-temporal register-I/O
    # variants of the power-of-two kernel (loads of x, stores of y)
    print(f"c2c f64 n={n} batch={batch}", flush=True)
    x = inputs_c(n + batch, n * batch, "f64").reshape(batch, n)
    xd = to_dev(x)
    p = F.plan_fft(torch.float64, n, batch=batch)
    d = p.describe()
    assert tag in d and ("stream=out" in d or tag == "stream=inout")
    y = to_host(p(xd)).reshape(batch, n)
    rows = [0, batch - 1] if batch > 1 else [0]
    assert rel_l2(y[rows], O.fft(x[rows])) <= tol_for("f64", n)
    back = to_host(F.plan_ifft(torch.float64, n, batch=batch)(to_dev(y))).reshape(batch, n)
    assert rel_l2(back, x) <= 2 * tol_for("f64", n)


def test_fixed_chain_fallbacks():
    # the generic kernel where the 81-point chain fits no group size: a
    # strided 81 axis of 40 columns, and a tuned C = 4 on an 81 pass
    L = F._lib
    print("c2c f32 dims=40x81 and 3^12 x 2 tuned C=4", flush=True)
    x = inputs_c(81, 40 * 81, "f32").reshape(81, 40)
    xd = to_dev(x)
    y = to_host(F.plan_fft(xd) * xd)
    ref = O.ndfft(x.astype(np.complex128).ravel(), [40, 81]).reshape(x.shape)
    assert rel_l2(y, ref) <= 200 * np.finfo(np.float32).eps
    n, batch = 3**12, 2
    x = inputs_c(12, n * batch, "f32").reshape(batch, n)
    p = F.FFTPlan(L.F32, (n,), -1, 1.0, batch, None, L.make_tuning((81, 6561), [{}, {"C": 4}]))
    assert "C=4 " in p.describe()
    y = to_host(p(to_dev(x))).reshape(batch, n)
    assert rel_l2(y, O.fft(x.astype(np.complex128))) <= tol_for("f32", n)


def test_chain_column_split_out_of_cache():
    # 4*3^12 x 16 (272 MB per array): the default split runs both column
    # passes on the 81-point chain and the 324-point pass on the rows
    n, batch = 4 * 3**12, 16
    print(f"c2c f32 n={n} batch={batch} chain columns", flush=True)
    x = inputs_c(3, n * batch, "f32").reshape(batch, n)
    xd = to_dev(x)
    p = F.plan_fft(torch.float32, n, batch=batch, planner="heuristic")
    d = p.describe()
    assert d.count("n=81 radices=9,9 kernel=chain") == 2 and "n=324 " in d
    y = p(xd)
    rows = [0, 7, batch - 1]
    yh = to_host(y).reshape(batch, n)
    assert rel_l2(yh[rows], O.fft(x[rows].astype(np.complex128))) <= tol_for("f32", n)
    back = to_host(F.plan_ifft(torch.float32, n, batch=batch)(y)).reshape(batch, n)
    assert rel_l2(back, x) <= 2 * tol_for("f32", n)
