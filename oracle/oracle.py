"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, always as the checker or the CPU baseline, never as a product path.
All arrays are numpy; complex data is interleaved (complex64 / complex128 /
clongdouble).  Precision suffixes: "f32", "f64", "ld".
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_CT = {"f32": ctypes.c_float, "f64": ctypes.c_double, "ld": ctypes.c_longdouble}
_NP_REAL = {"f32": np.float32, "f64": np.float64, "ld": np.longdouble}
_NP_CPLX = {"f32": np.complex64, "f64": np.complex128, "ld": np.clongdouble}


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i64, vp = ctypes.c_int64, ctypes.c_void_p
        for p in ("f32", "f64", "ld"):
            T = _CT[p]
            getattr(L, f"oracle_fft_{p}").argtypes = [i64, ctypes.c_int, T, vp, vp]
            getattr(L, f"oracle_fft_batch_{p}").argtypes = [i64, i64, ctypes.c_int, T, vp, vp]
            getattr(L, f"oracle_fourstep_{p}").argtypes = [i64, i64, i64, ctypes.c_int, T, vp, vp, ctypes.c_int]
            getattr(L, f"oracle_resample_{p}").argtypes = [i64, ctypes.c_int, T, vp, vp]
            getattr(L, f"oracle_hvdh_{p}").argtypes = [i64, ctypes.c_int, T, vp, vp]
            getattr(L, f"oracle_hvdh_halo_{p}").argtypes = [i64, ctypes.c_int, T, vp, vp, i64]
            getattr(L, f"oracle_split_{p}").argtypes = [i64, ctypes.c_int, T, vp, vp]
            getattr(L, f"oracle_rfft_{p}").argtypes = [i64, vp, vp]
            getattr(L, f"oracle_irfft_{p}").argtypes = [i64, vp, vp, ctypes.c_int]
            getattr(L, f"oracle_ndfft_{p}").argtypes = [ctypes.c_int, vp, ctypes.c_int, T, vp, vp]
            getattr(L, f"dft_naive_{p}").argtypes = [i64, ctypes.c_int, vp, vp]
            getattr(L, f"baseline_fft_batch_{p}").argtypes = [i64, i64, ctypes.c_int, T, vp, vp, ctypes.c_int]
            getattr(L, f"baseline_ndfft_{p}").argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, T, vp,
                                                          ctypes.c_int]
            getattr(L, f"baseline_real_batch_{p}").argtypes = [i64, i64, ctypes.c_int, ctypes.c_int, vp, vp,
                                                               ctypes.c_int]
        L.oracle_chain_radices.argtypes = [i64, ctypes.c_int, vp]
        L.oracle_nextsmooth.argtypes = [i64]
        L.oracle_nextsmooth.restype = i64
        L.oracle_hvdh_t.argtypes = [i64]
        L.oracle_hvdh_t.restype = i64
        L.oracle_gauss_legendre.argtypes = [ctypes.c_int, vp, vp]
        L.oracle_chunk_range.argtypes = [i64, ctypes.c_int, ctypes.c_int, i64,
                                         ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.oracle_expi.argtypes = [i64, i64, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]
        L.oracle_splitmix64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, i64, vp]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _cplx(x, p):
    return np.ascontiguousarray(x, dtype=_NP_CPLX[p])


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} failed (rc={rc})")


def fft(x, sign=-1, scale=1.0, prec="f64"):
    """1-D transform of the last axis's rows (any leading batch shape)."""
    x = _cplx(x, prec)
    n = x.shape[-1]
    batch = x.size // n if n else 0
    y = np.empty_like(x)
    _check(getattr(lib(), f"oracle_fft_batch_{prec}")(n, batch, sign, scale, _ptr(x), _ptr(y)), "fft")
    return y


def fourstep(x, N1, N2, sign=-1, scale=1.0, prec="f64", nthreads=1):
    x = _cplx(x, prec)
    y = np.empty_like(x)
    _check(getattr(lib(), f"oracle_fourstep_{prec}")(x.size, N1, N2, sign, scale, _ptr(x), _ptr(y), nthreads),
           "fourstep")
    return y


def resample(x, sign=-1, scale=1.0, prec="f64"):
    x = _cplx(x, prec)
    y = np.empty_like(x)
    _check(getattr(lib(), f"oracle_resample_{prec}")(x.size, sign, scale, _ptr(x), _ptr(y)), "resample")
    return y


def hvdh(x, sign=-1, scale=1.0, prec="f64", halo=None):
    """1-D transform of each row by the HvdH Gaussian-resampling route (the
    reference's route for odd non-smooth lengths, hvdh.jl).  halo overrides
    the probed halo rows (tests plant a truncated one)."""
    x = _cplx(x, prec)
    n = x.shape[-1]
    rows = x.reshape(-1, n)
    y = np.empty_like(rows)
    for i in range(rows.shape[0]):
        r = np.ascontiguousarray(rows[i])
        o = np.empty_like(r)
        if halo is None:
            rc = getattr(lib(), f"oracle_hvdh_{prec}")(n, sign, scale, _ptr(r), _ptr(o))
        else:
            rc = getattr(lib(), f"oracle_hvdh_halo_{prec}")(n, sign, scale, _ptr(r), _ptr(o), halo)
        _check(rc, "hvdh")
        y[i] = o
    return y.reshape(x.shape)


def split(x, sign=-1, scale=1.0, prec="f64"):
    """1-D transform of each row by the reference's SplitPlan (composite.jl:
    103-200): length-p HvdH sub-transforms of the decimated inputs, twiddled,
    then the length-s transforms, s the 7-smooth part and p odd."""
    x = _cplx(x, prec)
    n = x.shape[-1]
    rows = x.reshape(-1, n)
    y = np.empty_like(rows)
    for i in range(rows.shape[0]):
        r = np.ascontiguousarray(rows[i])
        o = np.empty_like(r)
        _check(getattr(lib(), f"oracle_split_{prec}")(n, sign, scale, _ptr(r), _ptr(o)), "split")
        y[i] = o
    return y.reshape(x.shape)


def hvdh_t(n):
    """HvdH's smooth grid length t for n (-1: none)."""
    return int(lib().oracle_hvdh_t(n))


def rfft(x, prec="f64"):
    x = np.ascontiguousarray(x, dtype=_NP_REAL[prec])
    n = x.shape[-1]
    rows = x.reshape(-1, n)
    out = np.empty((rows.shape[0], n // 2 + 1), dtype=_NP_CPLX[prec])
    for i in range(rows.shape[0]):
        r = np.ascontiguousarray(rows[i])
        o = np.empty(n // 2 + 1, dtype=_NP_CPLX[prec])
        _check(getattr(lib(), f"oracle_rfft_{prec}")(n, _ptr(r), _ptr(o)), "rfft")
        out[i] = o
    return out.reshape(x.shape[:-1] + (n // 2 + 1,))


def irfft(X, n, normalized=True, prec="f64"):
    X = _cplx(X, prec)
    rows = X.reshape(-1, X.shape[-1])
    out = np.empty((rows.shape[0], n), dtype=_NP_REAL[prec])
    for i in range(rows.shape[0]):
        r = np.ascontiguousarray(rows[i])
        o = np.empty(n, dtype=_NP_REAL[prec])
        _check(getattr(lib(), f"oracle_irfft_{prec}")(n, _ptr(r), _ptr(o), int(normalized)), "irfft")
        out[i] = o
    return out.reshape(X.shape[:-1] + (n,))


def ndfft(x, dims, sign=-1, scale=1.0, prec="f64"):
    """N-d transform; dims column-major (dims[0] fastest), x flat of prod(dims)."""
    x = _cplx(x, prec).ravel()
    d = np.asarray(dims, dtype=np.int64)
    y = np.empty_like(x)
    _check(getattr(lib(), f"oracle_ndfft_{prec}")(len(dims), _ptr(d), sign, scale, _ptr(x), _ptr(y)), "ndfft")
    return y


def _axis_fft(y, ax, sign, prec):
    # complex transform of every line along axis ax (rows of the oracle's fft)
    return np.moveaxis(fft(np.ascontiguousarray(np.moveaxis(y, ax, -1)), sign, 1.0, prec), -1, ax)


def ndrfft(x, prec="f64"):
    """N-d real forward transform of a row-major real array (shape s0..s_{N-1}
    is the column-major array of the reversed dims), halved along the last,
    contiguous axis: real rows along that axis first, then the complex
    transform of every other axis (ndplan.jl:370-418: _real_pencils! into
    the intermediate, then _exec_nd! of dims 2..N)."""
    y = rfft(x, prec)
    for ax in range(np.ndim(x) - 1):
        y = _axis_fft(y, ax, -1, prec)
    return np.ascontiguousarray(y)


def ndirfft(X, n, normalized=True, prec="f64"):
    """N-d complex-to-real inverse (ndplan.jl:478-521): the complex inverse
    of every axis but the last, scaled by 1/prod of those dims when
    normalized (restscale, ndplan.jl:456), then the real rows along the
    last axis (1/n when normalized, brfft otherwise)."""
    y = _cplx(X, prec)
    for ax in range(y.ndim - 1):
        y = _axis_fft(y, ax, +1, prec)
    if normalized and y.ndim > 1:
        y = (y / np.prod(y.shape[:-1])).astype(_NP_CPLX[prec])
    return irfft(y, n, normalized, prec)


def dft_naive(x, sign=-1, prec="ld"):
    x = _cplx(x, prec)
    y = np.empty_like(x)
    _check(getattr(lib(), f"dft_naive_{prec}")(x.size, sign, _ptr(x), _ptr(y)), "dft")
    return y


def baseline_fft_batch(x, sign=-1, scale=1.0, prec="f32", nthreads=1):
    """The reference's loop of 1-D plans over a batch, on nthreads threads."""
    x = _cplx(x, prec)
    n = x.shape[-1]
    y = np.empty_like(x)
    _check(getattr(lib(), f"baseline_fft_batch_{prec}")(n, x.size // n, sign, scale, _ptr(x), _ptr(y), nthreads),
           "baseline")
    return y


def baseline_ndfft(x, dims, sign=-1, scale=1.0, prec="f32", nthreads=1, first=0):
    """The reference's N-d plan on nthreads threads (column-major dims, axes
    from `first`); returns a transformed copy of x."""
    y = _cplx(x, prec).copy()
    d = np.asarray(dims, dtype=np.int64)
    _check(getattr(lib(), f"baseline_ndfft_{prec}")(len(dims), _ptr(d), first, sign, scale, _ptr(y), nthreads),
           "baseline_ndfft")
    return y


def baseline_rfft_batch(x, prec="f32", nthreads=1):
    """rfft of the rows of x (even length) on nthreads threads."""
    x = np.ascontiguousarray(x, dtype=_NP_REAL[prec])
    n = x.shape[-1]
    y = np.empty(x.shape[:-1] + (n // 2 + 1,), dtype=_NP_CPLX[prec])
    _check(getattr(lib(), f"baseline_real_batch_{prec}")(n, x.size // n, 0, 0, _ptr(x), _ptr(y), nthreads),
           "baseline_real_batch")
    return y


def baseline_irfft_batch(X, n, normalized=True, prec="f32", nthreads=1):
    """irfft (brfft when not normalized) of the rows of X on nthreads threads."""
    X = _cplx(X, prec)
    y = np.empty(X.shape[:-1] + (n,), dtype=_NP_REAL[prec])
    _check(getattr(lib(), f"baseline_real_batch_{prec}")(n, X.size // X.shape[-1], 1, int(normalized), _ptr(X),
                                                         _ptr(y), nthreads), "baseline_real_batch")
    return y


def chain_radices(n, W):
    out = np.zeros(64, dtype=np.int32)
    k = lib().oracle_chain_radices(n, W, _ptr(out))
    return None if k < 0 else [int(v) for v in out[:k]]


def nextsmooth(n):
    return int(lib().oracle_nextsmooth(n))


def gauss_legendre(nq):
    z = np.zeros(nq)
    w = np.zeros(nq)
    lib().oracle_gauss_legendre(nq, _ptr(z), _ptr(w))
    return z, w


def chunk_range(length, nt, c, align=1):
    lo, hi = ctypes.c_int64(), ctypes.c_int64()
    lib().oracle_chunk_range(length, nt, c, align, ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def expi(e, L, sign):
    c, s = ctypes.c_double(), ctypes.c_double()
    lib().oracle_expi(e, L, sign, ctypes.byref(c), ctypes.byref(s))
    return complex(c.value, s.value)


def splitmix64(seed, count, first=0):
    out = np.empty(count, dtype=np.uint64)
    lib().oracle_splitmix64(seed, first, count, _ptr(out))
    return out
