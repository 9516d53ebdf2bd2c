"""Deterministic randomized sweep over every plan kind against the oracle:
1-D complex (13-smooth, non-smooth and prime lengths, batches), 1-D real
(even and odd, forward and both inverses on an arbitrary spectrum), N-d
complex and real shapes (prime and ragged dims), and ragged batches.  The
cases are drawn once from a fixed seed, so a failure names a reproducible
case; every row of every batch is checked."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, cplx_np, inputs_c, real_np, rel_l2, to_dev, to_host, tol_for, torch_dtype

pytestmark = pytest.mark.gpu

_R = np.random.default_rng(20261203)


def _smooth(n):
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


def _is_prime(n):
    if n < 2:
        return False
    i = 2
    while i * i <= n:
        if n % i == 0:
            return False
        i += 1
    return True


def _draw_length(kind, lo, hi):
    while True:
        n = int(np.exp(_R.uniform(np.log(lo), np.log(hi))))
        if kind == "smooth" and _smooth(n):
            return n
        if kind == "prime" and _is_prime(n):
            return n
        if kind == "other" and not _smooth(n) and not _is_prime(n):
            return n


C2C = []
for i in range(36):
    prec = "f32" if i % 3 else "f64"
    kind = ("smooth", "smooth", "prime", "other")[i % 4]
    hi = 2**22 if prec == "f32" else 2**20
    lo = 2 if i % 2 else 2**12  # half of the cases multi-pass sized
    n = _draw_length(kind, lo, hi if kind == "smooth" else 2**19)
    batch = int(_R.choice([1, 2, 3, 5]))
    while n * batch > 2**22:
        batch = max(1, batch // 2) if batch > 1 else 1
        if batch == 1:
            break
    C2C.append((prec, n, batch, int(_R.choice([-1, 1]))))


@pytest.mark.parametrize("prec,n,batch,sign", C2C)
def test_sweep_c2c(prec, n, batch, sign):
    T = torch_dtype(prec)
    x = inputs_c(n + 17, n * batch, prec).reshape(batch, n)
    p = F.plan_fft(T, n, batch=batch) if sign < 0 else F.plan_bfft(T, n, batch=batch)
    y = to_host(p(to_dev(x))).reshape(batch, n)
    ref = O.fft(x.astype(np.complex128), sign=sign)
    for b in range(batch):
        assert rel_l2(y[b], ref[b]) <= tol_for(prec, n), b


REAL = []
for i in range(20):
    prec = "f32" if i % 2 else "f64"
    if i % 4 == 0:
        # odd lengths take the widened complex route
        n = int(np.exp(_R.uniform(np.log(2), np.log(2**18)))) | 1
    else:
        n = 2 * int(np.exp(_R.uniform(np.log(2), np.log(2**20 if prec == "f32" else 2**18))))
    REAL.append((prec, n, int(_R.choice([1, 2, 3]))))


@pytest.mark.parametrize("prec,n,batch", REAL)
def test_sweep_real(prec, n, batch):
    T = torch_dtype(prec)
    h = n // 2 + 1
    x = F.rng.normal(n + 19, n * batch).astype(real_np(prec)).reshape(batch, n)
    X = to_host(F.plan_rfft(T, n, batch=batch)(to_dev(x))).reshape(batch, h)
    ref = O.rfft(x.astype(np.float64))
    Z = F.rng.uniform_complex(n + 23, h * batch).astype(cplx_np(prec)).reshape(batch, h)
    Zd = to_dev(Z)
    yi = to_host(F.plan_irfft(T, n, batch=batch)(Zd)).reshape(batch, n)
    yb = to_host(F.plan_brfft(T, n, batch=batch)(Zd)).reshape(batch, n)
    for b in range(batch):
        assert rel_l2(X[b], ref[b]) <= tol_for(prec, n), b
        zr = Z[b].astype(np.complex128)
        assert rel_l2(yi[b], O.irfft(zr, n)) <= tol_for(prec, n), b
        assert rel_l2(yb[b], O.irfft(zr, n, normalized=False)) <= tol_for(prec, n), b


def _draw_shape():
    rank = int(_R.choice([2, 2, 3]))
    while True:
        dims = tuple(int(np.exp(_R.uniform(0, np.log(400 if rank == 2 else 80)))) for _ in range(rank))
        if 2 <= int(np.prod(dims)) <= 2**20:
            return dims


ND = [(("f32", "f64")[i % 2], _draw_shape()) for i in range(16)]


@pytest.mark.parametrize("prec,dims", ND)
def test_sweep_nd(prec, dims):
    n = int(np.prod(dims))
    tol = 1e-13 if prec == "f64" else 200 * np.finfo(np.float32).eps   # runtests.jl:332 tier
    x = inputs_c(n + 29, n, prec).reshape(tuple(reversed(dims)))
    xd = to_dev(x)
    y = to_host(F.plan_fft(xd) * xd)
    ref = O.ndfft(x.astype(np.complex128).ravel(), list(dims)).reshape(x.shape)
    assert rel_l2(y, ref) <= tol
    # N-d real on the same dims (dims[0], the torch last axis, halved)
    xr = F.rng.normal(n + 31, n).astype(real_np(prec)).reshape(tuple(reversed(dims)))
    Xr = to_host(F.plan_rfft(to_dev(xr)) * to_dev(xr))
    assert rel_l2(Xr, O.ndrfft(xr.astype(np.float64))) <= tol
    back = to_host(F.plan_irfft(to_dev(Xr.astype(cplx_np(prec))), dims[0]) * to_dev(Xr.astype(cplx_np(prec))))
    assert rel_l2(back, xr) <= 2 * tol


RAGGED = []
for i in range(4):
    sizes = []
    for _ in range(40):
        kind = ("smooth", "smooth", "smooth", "prime", "other")[int(_R.integers(5))]
        sizes.append(_draw_length(kind, 2, 2**16 if kind == "smooth" else 9000))
    RAGGED.append((("f32", "f64")[i % 2], tuple(sizes)))


@pytest.mark.parametrize("prec,sizes", RAGGED)
def test_sweep_ragged(prec, sizes):
    sizes = np.asarray(sizes, dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    x = inputs_c(int(sizes[0]), int(sizes.sum()), prec)
    y = to_host(F.plan_ragged(torch_dtype(prec), sizes)(to_dev(x)))
    for n, o in zip(sizes.tolist(), offs.tolist()):
        assert rel_l2(y[o:o + n], O.fft(x[o:o + n].astype(np.complex128))) <= tol_for(prec, n), n
