"""GPU parity of the HvdH route (the reference's HvdhPlan, hvdh.jl:154-241:
Gaussian spread onto a smooth t coprime to n, the t-point transform, chunked
banded solves) for odd non-smooth FP32 lengths: against the FP64 DFT and the
oracle's restatement of HvdhPlan on the same seeded inputs, both directions,
in place, batched -- the primes 65,537 x 4 and 1,000,003 x 4 at full size."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, inputs_c, rel_l2, to_dev, to_host, tol_for

pytestmark = pytest.mark.gpu


def _plans(n, batch):
    pf = F.plan_fft(torch.float32, n, batch=batch, route="hvdh")
    pi = F.plan_ifft(torch.float32, n, batch=batch, route="hvdh")
    assert "hvdh_solve" in pf.describe() and "hvdh_spread" in pi.describe()
    return pf, pi


@pytest.mark.parametrize("n,batch", [(65537, 4), (1000003, 4)])
def test_config3_primes_hvdh(n, batch):
    print(f"hvdh c2c f32 n={n} batch={batch}", flush=True)
    x = inputs_c(0, n * batch, "f32").reshape(batch, n)
    pf, pi = _plans(n, batch)
    y = to_host(pf(to_dev(x))).reshape(batch, n)
    rows = [0, batch - 1]
    x64 = x[rows].astype(np.complex128)
    ref = O.fft(x64)
    err = rel_l2(y[rows], ref)
    # the reference's HvdH tier is 50 eps(F32) = 5.96e-6 (runtests.jl:121, 183)
    assert err <= 5.96e-6, err
    # the oracle's HvdhPlan restatement (FP64) on the same rows
    ref_h = np.stack([O.hvdh(r, prec="f64") for r in x64])
    assert rel_l2(y[rows], ref_h) <= 5.96e-6
    # inverse (scale 1/n) and the round trip
    xi = to_host(pi(to_dev(x))).reshape(batch, n)
    assert rel_l2(xi[rows], O.fft(x64, sign=1, scale=1.0 / n)) <= 5.96e-6
    back = to_host(pi(to_dev(y.astype(np.complex64)))).reshape(batch, n)
    assert rel_l2(back, x) <= 2 * 5.96e-6


@pytest.mark.parametrize("n", [16387, 30031, 100003, 3 * 5 * 7 * 11 * 13 * 17 * 19 + 2])
def test_hvdh_lengths(n):
    if n % 2 == 0:
        n += 1
    print(f"hvdh c2c f32 n={n} batch=3", flush=True)
    x = inputs_c(n, 3 * n, "f32").reshape(3, n)
    try:
        pf, _ = _plans(n, 3)
    except F.UnsupportedError:
        pytest.skip(f"no HvdH grid within the solve kernel's halo for n={n}")
    y = to_host(pf(to_dev(x))).reshape(3, n)
    assert rel_l2(y, O.fft(x.astype(np.complex128))) <= 5.96e-6


def test_hvdh_in_place_and_bfft():
    n, batch = 65537, 2
    print(f"hvdh in place n={n}", flush=True)
    x = inputs_c(5, n * batch, "f32")
    pf = F.plan_fft(torch.float32, n, batch=batch, route="hvdh")
    y1 = to_host(pf(to_dev(x)))
    z = to_dev(x)
    assert pf(z, z) is z
    assert np.array_equal(to_host(z), y1)
    pb = F.plan_bfft(torch.float32, n, batch=batch, route="hvdh")
    yb = to_host(pb(to_dev(x))).reshape(batch, n)
    ref = O.fft(x.reshape(batch, n).astype(np.complex128), sign=1)
    assert rel_l2(yb, ref) <= 5.96e-6


# the reference's SplitPlan (composite.jl:103-200): an even length n = s*p
# has no HvdH grid coprime to it, so the route takes one length-s step and
# runs the s*batch length-p transforms by HvdH, the output permutation
# writing at stride s; s = 16 is CompositePlan's batched HvdH pass
# (composite.jl:42-43, 94-96)
@pytest.mark.parametrize("n,batch", [(2 * 65537, 3), (6 * 65537, 2), (10 * 100003, 1), (16 * 65537, 1),
                                     (4 * 1000003, 1)])
def test_split_route(n, batch):
    print(f"hvdh split c2c f32 n={n} batch={batch}", flush=True)
    x = inputs_c(n + 7, n * batch, "f32").reshape(batch, n)
    pf, pi = _plans(n, batch)
    d = pf.describe()
    s = n // next(p for p in (65537, 100003, 1000003) if n % p == 0)
    assert f"pass x->y n={s} " in d and f"out_stride={s}" in d, d
    y = to_host(pf(to_dev(x))).reshape(batch, n)
    x64 = x.astype(np.complex128)
    assert rel_l2(y, O.fft(x64)) <= 5.96e-6
    # the oracle's SplitPlan restatement (FP64) on the first row
    assert rel_l2(y[0], O.split(x64[0])) <= 5.96e-6
    xi = to_host(pi(to_dev(x))).reshape(batch, n)
    assert rel_l2(xi, O.fft(x64, sign=1, scale=1.0 / n)) <= 5.96e-6
    back = to_host(pi(to_dev(y.astype(np.complex64)))).reshape(batch, n)
    assert rel_l2(back, x) <= 2 * 5.96e-6
    z = to_dev(x)
    assert pf(z, z) is z
    assert np.array_equal(to_host(z).reshape(batch, n), y)
