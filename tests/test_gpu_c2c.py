"""GPU parity: 1-D complex plans (direct chain, four-step, batched) against
the FP64 oracle on the same seeded inputs (the reference's test matrix,
runtests.jl:6-56, 171-231), plus the benchmark's 1-D shapes at full size."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, inputs_c, rel_l2, to_dev, to_host, tol_for, torch_dtype

pytestmark = pytest.mark.gpu


def _check_c2c(prec, n, batch=1, seed=1, inverse_too=True):
    x = inputs_c(seed, n * batch, prec).reshape(batch, n) if batch > 1 else inputs_c(seed, n, prec)
    xd = to_dev(x)
    p = F.plan_fft(torch_dtype(prec), n, batch=batch)
    y = to_host(p(xd))
    ref = O.fft(x.astype(np.complex128))
    err = rel_l2(y, ref)
    assert err <= tol_for(prec, n), (prec, n, batch, err)
    if inverse_too:
        pi = F.plan_ifft(torch_dtype(prec), n, batch=batch)
        xi = to_host(pi(xd))
        refi = O.fft(x.astype(np.complex128), sign=1, scale=1.0 / n)
        assert rel_l2(xi, refi) <= tol_for(prec, n)
        back = to_host(pi(to_dev(y)))
        assert rel_l2(back, x) <= 2 * tol_for(prec, n)
    return y


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("logn", range(0, 19))
def test_pow2(prec, logn):
    _check_c2c(prec, 1 << logn, seed=logn)


MIXED = [3, 5, 6, 7, 9, 10, 12, 14, 15, 20, 21, 24, 25, 27, 35, 45, 48, 49, 60, 63, 105, 120, 315, 27 * 32,
         3 * 2**8, 9 * 2**8, 5 * 2**8, 7 * 2**8, 15 * 2**7, 105 * 2**4, 5**3 * 2**6, 7**2 * 2**6,
         3**4 * 5 * 7 * 2**5, 11, 13, 22, 26, 429, 352, 416, 1001, 11**2 * 2**4, 13**2 * 2**4,
         2**4 * 3**2 * 11 * 13, 3 * 5 * 7 * 11 * 13 * 2**4, 100]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", MIXED)
def test_mixed_radix(prec, n):
    _check_c2c(prec, n, seed=n)


FOURSTEP = [2**17, 2**18, 2**20, 3 * 2**16, 5 * 2**16, 7 * 2**16, 9 * 2**15, 15 * 2**15,
            3**8 * 4, 5**6 * 2, 7**5 * 3, 3**4 * 11 * 13 * 8, 13**4, 2 * 3**9, 3**10 * 2, 10**5]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", FOURSTEP)
def test_fourstep_sizes(prec, n):
    _check_c2c(prec, n, seed=n)


@pytest.mark.parametrize("prec,n,batch", [("f32", 4096, 64), ("f64", 1024, 33), ("f32", 60, 1000),
                                          ("f32", 1000, 8), ("f32", 2**17, 3), ("f64", 3 * 2**16, 2),
                                          ("f32", 100, 37), ("f32", 10**4, 5)])
def test_batched(prec, n, batch):
    _check_c2c(prec, n, batch=batch, seed=batch)


@pytest.mark.parametrize("prec,n,batch", [("f32", 4096, 16), ("f32", 2**20, 2), ("f64", 1000, 5)])
def test_inplace_and_determinism(prec, n, batch):
    x = inputs_c(9, n * batch, prec)
    p = F.plan_fft(torch_dtype(prec), n, batch=batch)
    y1 = to_host(p(to_dev(x)))
    y2 = to_host(p(to_dev(x)))
    assert np.array_equal(y1, y2)                      # run-to-run bitwise
    z = to_dev(x)
    assert p(z, z) is z                                 # in place, alias safe
    assert np.array_equal(to_host(z), y1)
    # a transform's result does not depend on its slot in the batch
    x2 = np.roll(x.reshape(batch, n), 1, axis=0)
    y3 = to_host(p(to_dev(x2.ravel()))).reshape(batch, n)
    assert np.array_equal(np.roll(y3, -1, axis=0), y1.reshape(batch, n))


def test_bfft_is_n_times_ifft():
    n = 3 * 1024
    x = to_dev(inputs_c(3, n, "f64"))
    yb = to_host(F.plan_bfft(torch.float64, n)(x))
    yi = to_host(F.plan_ifft(torch.float64, n)(x))
    assert rel_l2(yb, n * yi) < 1e-14


# ---- the benchmark's 1-D shapes at full size ------------------------------------

def test_config1_f64_1024(golden):
    x = golden["cfg1_x"]
    y = to_host(F.plan_fft(torch.float64, 1024)(to_dev(x)))
    assert rel_l2(y, O.fft(x)) <= 1e-13
    assert rel_l2(y, golden["cfg1_y"]) <= 1e-13


def test_config2a_4096x16384():
    n, batch = 4096, 16384
    x = inputs_c(0, n * batch, "f32")
    xd = to_dev(x)
    p = F.plan_fft(torch.float32, n, batch=batch)
    pi = F.plan_ifft(torch.float32, n, batch=batch)
    yd = p(xd)
    y = to_host(yd).reshape(batch, n)
    rows = [0, 1, 4095, 8191, 16383]
    ref = O.fft(x.reshape(batch, n)[rows].astype(np.complex128))
    assert rel_l2(y[rows], ref) <= tol_for("f32", n)
    back = to_host(pi(yd))
    assert rel_l2(back, x) <= 2 * tol_for("f32", n)
    # Parseval over the whole batch
    y64, x64 = y.astype(np.complex128), x.astype(np.complex128)
    assert abs(np.vdot(y64, y64).real / (n * np.vdot(x64, x64).real) - 1) < 1e-5


def test_config2b_2pow24():
    n = 1 << 24
    x = inputs_c(0, n, "f32")
    xd = to_dev(x)
    y = to_host(F.plan_fft(torch.float32, n)(xd))
    ref = O.fft(x.astype(np.complex128))
    assert rel_l2(y, ref) <= tol_for("f32", n)
    back = to_host(F.plan_ifft(torch.float32, n)(to_dev(y)))
    assert rel_l2(back, x) <= 2 * tol_for("f32", n)


def test_config2c_1e6x8():
    n, batch = 10**6, 8
    x = inputs_c(0, n * batch, "f32").reshape(batch, n)
    y = to_host(F.plan_fft(torch.float32, n, batch=batch)(to_dev(x)))
    ref = O.fft(x[[0, 7]].astype(np.complex128))
    assert rel_l2(y.reshape(batch, n)[[0, 7]], ref) <= tol_for("f32", n)


@pytest.mark.parametrize("batch", [1, 3])
def test_lean_square_1e6(batch):
    """10^6 = 100^3 on the square lean kernel (100 = 10 x 10, 50 columns per
    group, twiddle periods 10^4 and 10^6): forward vs the oracle, the 1/n
    inverse back, and the in-place call (x == y through the scratch)."""
    n = 10**6
    x = inputs_c(9 + batch, n * batch, "f32").reshape(batch, n)
    p = F.plan_fft(torch.float32, n, batch=batch, planner="heuristic")
    assert p.describe().count("radices=10,10 kernel=lean") == 3
    xd = to_dev(x)
    y = to_host(p * xd).reshape(batch, n)
    rows = [0, batch - 1] if batch > 1 else [0]
    assert rel_l2(y[rows], O.fft(x[rows].astype(np.complex128))) <= tol_for("f32", n)
    pi = F.plan_ifft(torch.float32, n, batch=batch, planner="heuristic")
    back = to_host(pi * to_dev(y))
    assert rel_l2(back.reshape(batch, n), x) <= 2 * tol_for("f32", n)
    z = to_dev(x)
    p(z, z)
    assert np.array_equal(to_host(z).reshape(batch, n), y)


# composite-radix butterflies (prime-factor 6, 10, 12, 14, 15 and 3 x 3 = 9):
# each forced as a stage of an explicit chain, first, middle and last
# position, both directions (the fused generic kernel and the plain one)
COMPOSITE_CHAINS = [[6, 16], [16, 9], [10, 16], [16, 12], [14, 16], [16, 15], [6, 6, 6], [9, 9, 9],
                    [10, 10, 10], [12, 5, 12], [14, 14], [15, 15], [2, 15, 3], [9, 7, 10]]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("chain", COMPOSITE_CHAINS, ids=lambda c: "x".join(map(str, c)))
def test_composite_radix_chains(prec, chain):
    L = F._lib
    n = int(np.prod(chain))
    batch = 3
    x = inputs_c(n, n * batch, prec).reshape(batch, n)
    code = L.F32 if prec == "f32" else L.F64
    for sign in (-1, 1):
        p = F.FFTPlan(code, (n,), sign, 1.0, batch, None, L.make_tuning((), [{"radices": chain}]))
        assert f"radices={','.join(map(str, chain))} " in p.describe()
        y = to_host(p(to_dev(x)))
        ref = O.fft(x.astype(np.complex128), sign=sign)
        assert rel_l2(y, ref) <= tol_for(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [90, 196, 225, 1000, 4000, 6000, 3 ** 10, 10 ** 6])
def test_default_composite_plans(prec, n):
    # lengths whose default chains take composite radices
    _check_c2c(prec, n, batch=2)


@pytest.mark.parametrize("prec,n,batch,launches", [("f32", 147456, 4, 2), ("f32", 3 * 2**17, 2, 2),
                                                   ("f64", 2**19, 1, 2), ("f32", 2**18, 2, 3)])
def test_small_footprint_two_passes(prec, n, batch, launches):
    # arrays of 8 MiB and less above the two-pass limit: two passes with a
    # power-of-two column pass (FP32 powers of two keep the three lean passes)
    T = torch_dtype(prec)
    x = inputs_c(n % 97 + batch, n * batch, prec).reshape(batch, n)
    for sign in (-1, 1):
        p = F.plan_fft(T, n, batch=batch) if sign < 0 else F.plan_ifft(T, n, batch=batch)
        d = p.describe()
        assert f"launches={launches}" in d, d
        y = to_host(p(to_dev(x))).reshape(batch, n)
        ref = O.fft(x.astype(np.complex128), sign=sign, scale=1.0 if sign < 0 else 1.0 / n)
        for b in range(batch):
            assert rel_l2(y[b], ref[b]) <= tol_for(prec, n), (sign, b)
