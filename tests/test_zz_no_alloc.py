"""fftjh_execute_static (the allocation-free variant of fftjh_execute) through
the raw C-ABI: its output must equal fftjh_execute's on the same plan and
inputs (bit for bit: the same kernels run, only the scratch lives elsewhere)
and meet the oracle's bar, on consecutive calls whose later inputs are
smaller (the fixed array keeps the earlier call's contents), for multi-pass,
resampled and ragged plans; plans needing more scratch than the array holds
are refused before any work."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import F, inputs_c, rel_l2, to_dev, to_host, tol_for

pytestmark = pytest.mark.gpu

L = F._lib


def _lib():
    return L.lib()


def _dims(d):
    return (ctypes.c_int64 * len(d))(*d)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _c2c(dtype, n, batch, sign=-1):
    h = ctypes.c_void_p()
    assert _lib().fftjh_plan_c2c(ctypes.byref(h), dtype, 1, _dims([n]), batch, sign, ctypes.c_double(1.0), 0) == L.OK
    return h


def _ragged(sizes):
    h = ctypes.c_void_p()
    arr = (ctypes.c_int64 * len(sizes))(*sizes)
    assert _lib().fftjh_plan_ragged(ctypes.byref(h), L.F32, len(sizes), arr, None, -1, 0, 0) == L.OK
    return h


def _both(h, x):
    """(fftjh_execute output, fftjh_execute_static output) of plan h on x."""
    lib = _lib()
    y_plain = torch.full_like(x, 3.0)
    y_static = torch.full_like(x, -5.0)
    assert lib.fftjh_execute(h, _ptr(y_plain), _ptr(x), _stream()) == L.OK
    rc = lib.fftjh_execute_static(h, _ptr(y_static), _ptr(x), _stream())
    assert rc == L.OK, lib.fftjh_last_error().decode()
    return to_host(y_plain), to_host(y_static)


def test_static_scratch_matches_execute_on_shrinking_calls():
    lib = _lib()
    # call 1: 2^20 x 4 FP32 (three passes, 32 MiB intermediate); call 2: a
    # smaller two-pass length; call 3: a resampled prime (two grids of
    # scratch); each result against fftjh_execute and the FP64 oracle
    cases = [("f32", 1 << 20, 4, 11), ("f32", 100000, 2, 12), ("f32", 65537, 2, 13), ("f64", 1 << 18, 2, 14),
             ("f64", 3 ** 7 * 5, 3, 15)]
    for prec, n, batch, seed in cases:
        print(f"execute_static c2c {prec} n={n} batch={batch}", flush=True)
        h = _c2c(L.F32 if prec == "f32" else L.F64, n, batch)
        try:
            x = inputs_c(seed, n * batch, prec).reshape(batch, n)
            y_plain, y_static = _both(h, to_dev(x))
            assert np.array_equal(y_plain, y_static), f"{prec} n={n}: static scratch result differs"
            rows = [0, batch - 1]
            ref = O.fft(x[rows].astype(np.complex128), -1, 1.0, "f64")
            assert rel_l2(y_static[rows], ref) <= tol_for(prec, n)
        finally:
            assert lib.fftjh_plan_destroy(h) == L.OK


def test_static_scratch_ragged_workspace_twice_smaller():
    lib = _lib()
    # two-level items (lengths above one tile use the per-workgroup
    # workspace) and a resampled prime; the second batch is smaller
    for sizes, seed in (([65536, 20000, 4096, 1009, 48000, 9216], 21), ([30000, 1031, 512, 16384], 22)):
        print(f"execute_static ragged sizes={sizes}", flush=True)
        h = _ragged(sizes)
        try:
            x = inputs_c(seed, sum(sizes), "f32")
            y_plain, y_static = _both(h, to_dev(x))
            assert np.array_equal(y_plain, y_static), f"ragged {sizes}: static scratch result differs"
            off = 0
            for n in sizes:
                ref = O.fft(x[off:off + n].astype(np.complex128)[None, :], -1, 1.0, "f64")[0]
                assert rel_l2(y_static[off:off + n], ref) <= tol_for("f32", n), n
                off += n
        finally:
            assert lib.fftjh_plan_destroy(h) == L.OK


def test_static_scratch_real_paired_passes():
    # R2C / C2R on the paired lean passes (2^20 x 2: three passes through
    # scratch); static scratch against fftjh_execute, bit for bit, and the
    # forward against the oracle
    lib = _lib()
    n, batch, h1 = 1 << 20, 2, (1 << 19) + 1
    print(f"execute_static r2c/c2r f32 n={n} batch={batch}", flush=True)
    x = F.rng.normal(31, n * batch).astype(np.float32)
    xd = to_dev(x)
    hr, hc = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.fftjh_plan_r2c(ctypes.byref(hr), L.F32, n, batch, 0) == L.OK
    assert lib.fftjh_plan_c2r(ctypes.byref(hc), L.F32, n, batch, 1, 0) == L.OK
    try:
        X1 = torch.full((batch * h1,), 3.0, dtype=torch.complex64, device="cuda")
        X2 = torch.full_like(X1, -5.0)
        assert lib.fftjh_execute(hr, _ptr(X1), _ptr(xd), _stream()) == L.OK
        assert lib.fftjh_execute_static(hr, _ptr(X2), _ptr(xd), _stream()) == L.OK, lib.fftjh_last_error().decode()
        a, b = to_host(X1), to_host(X2)
        assert np.array_equal(a, b)
        ref = O.rfft(x.reshape(batch, n).astype(np.float64))
        assert rel_l2(b.reshape(batch, h1)[batch - 1], ref[batch - 1]) <= tol_for("f32", n)
        z1 = torch.full((batch * n,), 3.0, dtype=torch.float32, device="cuda")
        z2 = torch.full_like(z1, -5.0)
        assert lib.fftjh_execute(hc, _ptr(z1), _ptr(X1), _stream()) == L.OK
        assert lib.fftjh_execute_static(hc, _ptr(z2), _ptr(X1), _stream()) == L.OK, lib.fftjh_last_error().decode()
        assert np.array_equal(to_host(z1), to_host(z2))
        assert rel_l2(to_host(z2), x) <= 2 * tol_for("f32", n)
    finally:
        assert lib.fftjh_plan_destroy(hr) == L.OK
        assert lib.fftjh_plan_destroy(hc) == L.OK


def test_static_scratch_refuses_plans_beyond_its_size():
    lib = _lib()
    print("execute_static refusal: c2c f32 n=2^24 (128 MiB intermediate)", flush=True)
    h = _c2c(L.F32, 1 << 24, 1)
    try:
        x = torch.zeros(1 << 24, dtype=torch.complex64, device="cuda:0")
        y = torch.full_like(x, 2.0)
        rc = lib.fftjh_execute_static(h, _ptr(y), _ptr(x), _stream())
        assert rc == L.ERR_UNSUPPORTED
        msg = lib.fftjh_last_error().decode()
        assert "scratch" in msg and str(L.STATIC_SCRATCH_BYTES) in msg, msg
        torch.cuda.synchronize()
        assert bool((y == 2.0).all()), "a refused call must queue no work"
    finally:
        assert lib.fftjh_plan_destroy(h) == L.OK
    assert lib.fftjh_execute_static(None, None, None, None) == L.ERR_ARGUMENT


def test_static_scratch_plan_without_scratch():
    lib = _lib()
    print("execute_static single pass: c2c f32 n=4096 batch=8", flush=True)
    h = _c2c(L.F32, 4096, 8)
    try:
        x = inputs_c(31, 4096 * 8, "f32").reshape(8, 4096)
        y_plain, y_static = _both(h, to_dev(x))
        assert np.array_equal(y_plain, y_static)
    finally:
        assert lib.fftjh_plan_destroy(h) == L.OK
