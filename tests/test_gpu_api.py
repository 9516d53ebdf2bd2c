"""The reference's API idioms on the GPU plans (runtests.jl:264-328,
423-427): array-based planning, mul_, in-place plans, inv(p).p, Scaled,
brfft, one-shots, and the argument / dimension errors."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, inputs_c, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("T,n", [(torch.float32, 96), (torch.float32, 11), (torch.float64, 96), (torch.float64, 11)])
def test_fftw_compat_api(T, n):
    prec = "f32" if T == torch.float32 else "f64"
    tol = 100 * (np.finfo(np.float32).eps if prec == "f32" else np.finfo(np.float64).eps)
    x = to_dev(inputs_c(1, n, prec))
    xh = to_host(x).astype(np.complex128)
    ref = O.fft(xh)
    p = F.plan_fft(x, flags="ignored")
    assert rel_l2(to_host(p * x), ref) <= tol
    y = torch.empty_like(x)
    assert F.mul_(y, p, x) is y
    assert rel_l2(to_host(y), ref) <= tol
    pb = F.plan_bfft(x)
    assert rel_l2(to_host(pb * x), O.fft(xh, sign=1)) <= tol
    # in place: DSP's complex conv idiom p! * buf; inv(p!).p
    pi_ = F.plan_fft_(x)
    z = x.clone()
    assert (pi_ * z) is z
    assert rel_l2(to_host(z), ref) <= tol
    ip = F.inv(pi_).p
    assert (ip * z) is z
    assert rel_l2(to_host(z), n * xh) <= tol
    full = F.inv(pi_)
    z2 = x.clone()
    pi_ * z2
    full * z2
    assert rel_l2(to_host(z2), xh) <= tol
    # rfft / brfft
    xr = to_dev(F.rng.normal(5, n).astype(np.float32 if prec == "f32" else np.float64))
    pr = F.plan_rfft(xr)
    X = pr * xr
    assert rel_l2(to_host(X), O.rfft(to_host(xr).astype(np.float64))) <= tol
    pbr = F.plan_brfft(X, n)
    assert rel_l2(to_host(pbr * X), O.irfft(to_host(X).astype(np.complex128), n, normalized=False)) <= tol
    yr = torch.empty_like(xr)
    assert F.mul_(yr, pbr, X) is yr
    # hilbert idiom: output into a view of a longer buffer
    if n % 2 == 0:
        Xbig = torch.zeros(n, dtype=X.dtype, device=X.device)
        F.mul_(Xbig[: n // 2 + 1], pr, xr)
        assert rel_l2(to_host(Xbig[: n // 2 + 1]), to_host(X)) == 0
    # one-shots
    assert rel_l2(to_host(F.fft(x)), ref) <= tol
    assert rel_l2(to_host(F.ifft(x)), O.fft(xh, sign=1, scale=1.0 / n)) <= tol
    assert rel_l2(to_host(F.bfft(x)), O.fft(xh, sign=1)) <= tol
    assert rel_l2(to_host(F.irfft(X, n)), to_host(xr)) <= tol
    assert rel_l2(to_host(F.fft_(x.clone())), ref) <= tol
    assert rel_l2(to_host(F.ifft_(x.clone())), O.fft(xh, sign=1, scale=1.0 / n)) <= tol
    # region validation
    with pytest.raises(F.ArgumentError):
        F.plan_fft(x, region=(1,))


def test_inv_algebra():
    n = 64
    p = F.plan_fft(torch.float64, n)
    s = F.inv(p)
    assert isinstance(s, F.Scaled) and s.p.sign == 1 and s.p.scale == 1 and s.scale == 1 / n
    q = F.inv(F.plan_ifft(torch.float64, n))
    assert isinstance(q, F.FFTPlan) and q.sign == -1
    r = F.inv(F.plan_bfft(torch.float64, n))
    assert isinstance(r, F.Scaled) and r.p.sign == -1 and r.scale == 1 / n
    rr = F.inv(s)
    assert isinstance(rr, F.Scaled) and rr.p.sign == -1 and abs(rr.scale - 1.0) < 1e-15
    x = to_dev(inputs_c(2, n, "f64"))
    assert rel_l2(to_host(rr * x), to_host(p * x)) < 1e-15
    assert F.size(p) == (n,) and F.size(F.plan_rfft(torch.float32, 10)) == (10,)
    assert F.size(F.plan_irfft(torch.float32, 10)) == (6,)


def test_dimension_and_argument_errors():
    p = F.plan_fft(torch.float64, 8)
    with pytest.raises(F.DimensionMismatch):
        p(torch.zeros(4, dtype=torch.complex128, device="cuda:0"))
    with pytest.raises(F.DimensionMismatch):
        p(torch.zeros(4, dtype=torch.complex128, device="cuda:0"), torch.zeros(8, dtype=torch.complex128,
                                                                                 device="cuda:0"))
    with pytest.raises(F.ArgumentError):
        F.plan_fft(torch.float64, 0)
    with pytest.raises(F.DimensionMismatch):
        F.plan_irfft(torch.zeros(5, dtype=torch.complex64, device="cuda:0"), 16)
    with pytest.raises(TypeError):
        p(torch.zeros(8, dtype=torch.complex64, device="cuda:0"))


def test_nd_real_api():
    """N-d real plans from example arrays (api.jl:131-154): size(p), mul!,
    one-shots (runtests.jl:385-387) and the dimension errors."""
    dev = "cuda:0"
    Ar = to_dev(F.rng.normal(11, 240).reshape(12, 20).astype(np.float32))
    X = F.rfft(Ar)
    assert X.shape == (12, 11)
    tol = 200 * np.finfo(np.float32).eps
    assert rel_l2(to_host(X), O.ndrfft(to_host(Ar).astype(np.float64))) <= tol
    assert rel_l2(to_host(F.irfft(X, 20)), to_host(Ar)) <= 2 * tol
    assert rel_l2(to_host(F.brfft(X, 20)), 240 * to_host(Ar)) <= 2 * tol
    pr = F.plan_rfft(Ar)
    assert F.size(pr) == (12, 20)
    Y = torch.empty(12, 11, dtype=torch.complex64, device=dev)
    assert F.mul_(Y, pr, Ar) is Y and torch.equal(Y, X)
    pi = F.plan_irfft(X, 20)
    assert F.size(pi) == (12, 11)
    with pytest.raises(F.DimensionMismatch):
        pr(torch.zeros(20, 12, dtype=torch.float32, device=dev))
    with pytest.raises(F.DimensionMismatch):
        pr(torch.empty(12, 10, dtype=torch.complex64, device=dev), Ar)
    with pytest.raises(F.DimensionMismatch):   # leading length 21 needs 11 rows
        F.plan_irfft(X, 21 + 2)
    with pytest.raises(F.DimensionMismatch):
        pi(torch.empty(11, 11, dtype=torch.complex64, device=dev))
    with pytest.raises(F.DimensionMismatch):
        pi(torch.empty(12, 21, dtype=torch.float32, device=dev), X)
    with pytest.raises(F.ArgumentError):
        F.plan_rfft(Ar, region=(1,))
    # odd leading length: 21 -> 11 rows as well
    x21 = to_dev(F.rng.normal(12, 4 * 21).reshape(4, 21).astype(np.float64))
    X21 = F.rfft(x21)
    assert X21.shape == (4, 11)
    assert rel_l2(to_host(F.irfft(X21, 21)), to_host(x21)) <= 1e-13


def test_graph_capture_replays():
    """Execution allocates nothing and is capture safe: a captured plan
    replays to the eager result on every replay."""
    n, batch = 4096, 8
    x = to_dev(inputs_c(4, n * batch, "f32"))
    p = F.plan_fft(torch.float32, n, batch=batch)
    ref = to_host(p(x))
    y = torch.empty_like(x)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        p(y, x)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        p(y, x)
    for _ in range(3):
        y.zero_()
        torch.cuda.synchronize()
        g.replay()
        assert np.array_equal(to_host(y), ref)
