"""The C-ABI called directly (raw ctypes, device pointers from torch), as a
non-Python host would call it: plan constructors, fftjh_plan_info,
fftjh_execute_checked's length checks before any work is queued, and the
N-d real entry points with batch > 1, against the oracle."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

L = F._lib
P64 = ctypes.POINTER(ctypes.c_int64)


def _lib():
    return L.lib()


def _dims(d):
    return (ctypes.c_int64 * len(d))(*d)


def _info(h):
    kind, dtype, sign, rank = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    dims = (ctypes.c_int64 * 8)()
    batch, ie, oe = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    scale = ctypes.c_double()
    assert _lib().fftjh_plan_info(h, ctypes.byref(kind), ctypes.byref(dtype), ctypes.byref(sign), ctypes.byref(rank),
                                  dims, ctypes.byref(batch), ctypes.byref(scale), ctypes.byref(ie),
                                  ctypes.byref(oe)) == L.OK
    return dict(kind=kind.value, dtype=dtype.value, sign=sign.value, rank=rank.value,
                dims=tuple(dims[i] for i in range(rank.value)), batch=batch.value, scale=scale.value,
                in_elems=ie.value, out_elems=oe.value)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_c2c_checked_execution():
    lib = _lib()
    n, batch = 1000, 3
    h = ctypes.c_void_p()
    assert lib.fftjh_plan_c2c(ctypes.byref(h), L.F64, 1, _dims([n]), batch, -1, ctypes.c_double(1.0), 0) == L.OK
    try:
        info = _info(h)
        assert info == dict(kind=L.KIND_C2C, dtype=L.F64, sign=-1, rank=1, dims=(n,), batch=batch, scale=1.0,
                            in_elems=n * batch, out_elems=n * batch)
        x = F.rng.uniform_complex(7, n * batch).reshape(batch, n)
        xd = to_dev(x)
        yd = torch.full_like(xd, 7.0)
        # a length mismatch is refused before any work is queued: y untouched
        rc = lib.fftjh_execute_checked(h, _ptr(yd), n * batch, _ptr(xd), n * batch - 1, _stream())
        assert rc == L.ERR_DIMENSION and "input length" in lib.fftjh_last_error().decode()
        rc = lib.fftjh_execute_checked(h, _ptr(yd), n * batch + 1, _ptr(xd), n * batch, _stream())
        assert rc == L.ERR_DIMENSION and "output length" in lib.fftjh_last_error().decode()
        assert torch.all(yd == 7.0)
        assert lib.fftjh_execute_checked(h, _ptr(yd), n * batch, _ptr(xd), n * batch, _stream()) == L.OK
        assert rel_l2(to_host(yd), O.fft(x)) <= 1e-13
        # in place: y == x
        assert lib.fftjh_execute(h, _ptr(xd), _ptr(xd), _stream()) == L.OK
        assert rel_l2(to_host(xd), O.fft(x)) <= 1e-13
        assert lib.fftjh_execute(h, None, _ptr(xd), _stream()) == L.ERR_ARGUMENT
    finally:
        assert lib.fftjh_plan_destroy(h) == L.OK


@pytest.mark.parametrize("dims", [(64, 48), (9, 4, 3), (256, 5)])
def test_nd_real_batched_entry_points(dims):
    # column-major dims (dims[0] halved), two arrays back to back
    lib = _lib()
    batch = 2
    n1, rest = dims[0], int(np.prod(dims[1:]))
    h1 = n1 // 2 + 1
    shape = tuple(reversed(dims))
    x = F.rng.normal(n1 + rest, batch * n1 * rest).reshape((batch,) + shape)
    hr, hc = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.fftjh_plan_r2c_nd(ctypes.byref(hr), L.F64, len(dims), _dims(dims), batch, 0) == L.OK
    assert lib.fftjh_plan_c2r_nd(ctypes.byref(hc), L.F64, len(dims), _dims(dims), batch, 1, 0) == L.OK
    try:
        ir, ic = _info(hr), _info(hc)
        assert ir["kind"] == L.KIND_R2C and ir["dims"] == dims and ir["in_elems"] == batch * n1 * rest
        assert ir["out_elems"] == batch * h1 * rest and ic["in_elems"] == ir["out_elems"]
        assert ic["kind"] == L.KIND_C2R and abs(ic["scale"] - 1.0 / (n1 * rest)) < 1e-18
        xd = to_dev(x)
        Xd = torch.empty(batch * h1 * rest, dtype=torch.complex128, device="cuda:0")
        assert lib.fftjh_execute_checked(hr, _ptr(Xd), Xd.numel(), _ptr(xd), xd.numel(), _stream()) == L.OK
        X = to_host(Xd).reshape((batch,) + shape[:-1] + (h1,))
        for b in range(batch):
            assert rel_l2(X[b], O.ndrfft(x[b])) <= 1e-13
        yd = torch.empty_like(xd)
        Xin = Xd.clone()
        assert lib.fftjh_execute_checked(hc, _ptr(yd), yd.numel(), _ptr(Xd), Xd.numel(), _stream()) == L.OK
        assert rel_l2(to_host(yd), x) <= 2e-13
        assert torch.equal(Xd, Xin)   # the inverse leaves its input intact
    finally:
        lib.fftjh_plan_destroy(hr)
        lib.fftjh_plan_destroy(hc)


def test_constructor_errors_leave_no_plan():
    lib = _lib()
    h = ctypes.c_void_p(1234)
    assert lib.fftjh_plan_r2c_nd(ctypes.byref(h), L.F32, 0, _dims([8]), 1, 0) == L.ERR_ARGUMENT
    assert h.value is None
    h = ctypes.c_void_p(1234)
    assert lib.fftjh_plan_c2r_nd(ctypes.byref(h), L.F32, 2, _dims([8, 0]), 1, 1, 0) == L.ERR_ARGUMENT
    assert h.value is None and "positive" in lib.fftjh_last_error().decode()
    h = ctypes.c_void_p(1234)
    assert lib.fftjh_plan_r2c_nd(ctypes.byref(h), L.F32, 2, _dims([8, 4]), 1, 0x10) == L.ERR_ARGUMENT
    assert h.value is None
