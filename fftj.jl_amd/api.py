"""Plan/apply API of the GPU engine, mirroring FFTJ.jl's (api.jl, plan.jl,
real.jl, ndplan.jl) so a GPU plan stands in for a CPU one.

    Julia (reference)                 here
    plan_fft(T, n)                    plan_fft(torch.float32, n)       plan.jl:207-213
    plan_fft(x::Array)                plan_fft(x)                      api.jl:107-112
    p(y, x) / p(x) / p * x            p(y, x) / p(x) / p * x           plan.jl:359-398, api.jl:34-35
    mul!(y, p, x)                     mul_(y, p, x)                    api.jl:34-55
    plan_fft!(T, n) / InplacePlan     plan_fft_(T, n) / InplacePlan    api.jl:15-24, 42-49
    inv(p) -> Scaled(p, s)            inv(p) -> Scaled(p, s)           api.jl:72-94
    size(p)                           size(p)                          api.jl:57-60, 166-168
    fft, ifft, bfft, fft!, rfft, ...  fft, ifft, bfft, fft_, rfft, ...  api.jl:97-99, 181-203

Data are torch tensors on the GPU (complex64/complex128 for complex sides,
float32/float64 for real sides), contiguous.  torch is only the memory and
stream plumbing: every transform runs in the HIP kernels of libfftjh.so on
torch's current stream.  N-d arrays: a row-major tensor of shape
(s0, ..., s_{N-1}) is the column-major array of dims (s_{N-1}, ..., s0), so
"transform all dims" is the same operation; size(p) reports the tensor shape.
Python regions are 0-based.  An extension over the reference: 1-D plans take
batch=B for B transforms stored back to back.  nthreads and FFTW planner
flags are accepted and ignored (the reference ignores flags, api.jl:107-108;
the GPU has no thread count).
"""
import ctypes
import math

import torch

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, check, dims_array, lib

F32, F64 = _lib.F32, _lib.F64
_CPLX = {F32: torch.complex64, F64: torch.complex128}
_REAL = {F32: torch.float32, F64: torch.float64}


def _code(T):
    """Precision code of a type argument (the reference's T)."""
    if T in (torch.float32, torch.complex64, "f32", "float32", "Float32"):
        return F32
    if T in (torch.float64, torch.complex128, "f64", "float64", "Float64"):
        return F64
    raise ArgumentError(f"element type must be float32 or float64 (got {T})")


def _code_of_tensor(x):
    if x.dtype in (torch.complex64, torch.float32):
        return F32
    if x.dtype in (torch.complex128, torch.float64):
        return F64
    raise ArgumentError(f"unsupported element type {x.dtype}")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require(x, dtype, what):
    if not isinstance(x, torch.Tensor):
        raise ArgumentError(f"{what} must be a torch tensor")
    if x.dtype != dtype:
        raise TypeError(f"{what} has element type {x.dtype}, plan needs {dtype}")
    if not x.is_cuda:
        raise ArgumentError(f"{what} must be on the GPU")
    if not x.is_contiguous():
        raise ArgumentError(f"{what} must be contiguous")


class _Handle:
    # keeps its own reference to the destroy entry point: at interpreter exit
    # module globals may already be cleared when the last plans are collected
    __slots__ = ("ptr", "_destroy")

    def __init__(self, ptr):
        self.ptr = ptr
        self._destroy = lib().fftjh_plan_destroy

    def __del__(self):
        if self.ptr:
            self._destroy(self.ptr)
            self.ptr = None


class _PlanBase:
    def _exec(self, y, x):
        check(lib().fftjh_execute(self._h.ptr, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                  _stream()))
        return y

    def describe(self):
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib().fftjh_plan_describe(self._h.ptr, buf, len(buf)))
        return buf.value.decode()

    def profile_launches(self, y, x, iters=20):
        """Mean milliseconds of each kernel launch of this plan over `iters` runs."""
        cap = ctypes.c_int(64)
        ms = (ctypes.c_double * 64)()
        check(lib().fftjh_profile_launches(self._h.ptr, ctypes.c_void_p(y.data_ptr()),
                                           ctypes.c_void_p(x.data_ptr()), _stream(), iters, ms, ctypes.byref(cap)))
        return [ms[i] for i in range(cap.value)]


class FFTPlan(_PlanBase):
    """Complex plan: 1-D (the reference's FFTPlan/FourStepPlan/ResamplePlan,
    chosen by the GPU planner) or N-d (NdPlan).  Fields n, sign, scale."""

    def __init__(self, dtype, dims, sign, scale, batch=1, shape=None, tuning=None, flags=0):
        self.dtype_code = dtype
        self.tuning = tuning
        self.flags = int(flags)
        self.dims = tuple(int(d) for d in dims)  # column-major, dims[0] fastest
        self.sign = int(sign)
        self.scale = float(scale)
        self.batch = int(batch)
        self.n = math.prod(self.dims)
        self.shape = tuple(shape) if shape is not None else (
            (self.n,) if len(self.dims) == 1 else tuple(reversed(self.dims)))
        h = ctypes.c_void_p()
        if tuning is None:
            check(lib().fftjh_plan_c2c(ctypes.byref(h), dtype, len(self.dims), dims_array(self.dims), self.batch,
                                       self.sign, self.scale, self.flags))
        else:
            if self.flags:
                tuning = _lib.Tuning.from_buffer_copy(tuning)
                tuning.flags |= self.flags
            check(lib().fftjh_plan_c2c_tuned(ctypes.byref(h), dtype, len(self.dims), dims_array(self.dims),
                                             self.batch, self.sign, self.scale, ctypes.byref(tuning)))
        self._h = _Handle(h.value)
        self._scaled = {}

    @property
    def dtype(self):
        return _CPLX[self.dtype_code]

    def _len_check(self, x, what):
        want = self.n * self.batch
        if x.numel() != want:
            raise DimensionMismatch(f"plan length {want}, {what} length {x.numel()}")
        if len(self.dims) > 1 and self.batch == 1 and x.dim() == len(self.dims) and tuple(x.shape) != self.shape:
            raise DimensionMismatch(f"plan dims {self.shape}, {what} dims {tuple(x.shape)}")

    def __call__(self, y, x=None):
        if x is None:
            x, y = y, None
        _require(x, self.dtype, "input")
        self._len_check(x, "input")
        if y is None:
            y = torch.empty_like(x)
        _require(y, self.dtype, "output")
        self._len_check(y, "output")
        return self._exec(y, x)

    def __mul__(self, x):
        return self(x)

    def _with_scale(self, s):
        """The same transform with its scale multiplied by s (Scaled applies this)."""
        key = float(s)
        if key not in self._scaled:
            self._scaled[key] = FFTPlan(self.dtype_code, self.dims, self.sign, self.scale * key, self.batch,
                                        self.shape, self.tuning, self.flags)
        return self._scaled[key]

    def _sibling(self, sign, scale):
        return FFTPlan(self.dtype_code, self.dims, sign, scale, self.batch, self.shape, self.tuning, self.flags)

    def __repr__(self):
        return f"FFTPlan(dims={self.dims}, sign={self.sign}, scale={self.scale}, batch={self.batch})"


class RfftPlan(_PlanBase):
    """Real forward plan: n reals -> n//2+1 complex (real.jl:6-13), or, with
    shape=(s0, ..., s_{N-1}), the N-d plan of that real tensor, halved along
    the last (contiguous) axis (ndplan.jl:355-418)."""

    def __init__(self, dtype, n, batch=1, shape=None):
        self.dtype_code = dtype
        self.batch = int(batch)
        self.shape = None if shape is None or len(shape) <= 1 else tuple(int(d) for d in shape)
        self.dims = (int(n),) if self.shape is None else tuple(reversed(self.shape))
        self.n = self.dims[0]
        h = ctypes.c_void_p()
        if self.shape is None:
            check(lib().fftjh_plan_r2c(ctypes.byref(h), dtype, self.n, self.batch, 0))
        else:
            check(lib().fftjh_plan_r2c_nd(ctypes.byref(h), dtype, len(self.dims), dims_array(self.dims),
                                          self.batch, 0))
        self._h = _Handle(h.value)

    @property
    def cshape(self):
        return self.shape[:-1] + (self.n // 2 + 1,)

    def __call__(self, y, x=None):
        if x is None:
            x, y = y, None
        _require(x, _REAL[self.dtype_code], "input")
        if self.shape is not None:
            if tuple(x.shape) != self.shape:
                raise DimensionMismatch(f"plan dims {self.shape}, input {tuple(x.shape)}")
            if y is None:
                y = torch.empty(self.cshape, dtype=_CPLX[self.dtype_code], device=x.device)
            _require(y, _CPLX[self.dtype_code], "output")
            if tuple(y.shape) != self.cshape:
                raise DimensionMismatch(f"rfft output dims must be {self.cshape}, got {tuple(y.shape)}")
            return self._exec(y, x)
        if x.numel() != self.n * self.batch:
            raise DimensionMismatch(f"plan length {self.n * self.batch}, input length {x.numel()}")
        nout = (self.n // 2 + 1) * self.batch
        if y is None:
            shape = (self.n // 2 + 1,) if self.batch == 1 else (self.batch, self.n // 2 + 1)
            y = torch.empty(shape, dtype=_CPLX[self.dtype_code], device=x.device)
        _require(y, _CPLX[self.dtype_code], "output")
        if y.numel() != nout:
            raise DimensionMismatch(f"rfft output must have length {nout}, got {y.numel()}")
        return self._exec(y, x)

    def __mul__(self, x):
        return self(x)


class IrfftPlan(_PlanBase):
    """Real inverse plan: n//2+1 complex -> n reals (real.jl:15-23), or, with
    shape=(s0, ..., s_{N-1}) the real output's shape, the N-d plan from the
    (s0, ..., s_{N-1}//2+1) complex tensor (ndplan.jl:435-521).
    normalized=True is irfft, False is brfft."""

    def __init__(self, dtype, n, normalized=True, batch=1, shape=None):
        self.dtype_code = dtype
        self.batch = int(batch)
        self.normalized = bool(normalized)
        self.shape = None if shape is None or len(shape) <= 1 else tuple(int(d) for d in shape)
        self.dims = (int(n),) if self.shape is None else tuple(reversed(self.shape))
        self.n = self.dims[0]
        h = ctypes.c_void_p()
        if self.shape is None:
            check(lib().fftjh_plan_c2r(ctypes.byref(h), dtype, self.n, self.batch, int(self.normalized), 0))
        else:
            check(lib().fftjh_plan_c2r_nd(ctypes.byref(h), dtype, len(self.dims), dims_array(self.dims),
                                          self.batch, int(self.normalized), 0))
        self._h = _Handle(h.value)

    @property
    def cshape(self):
        return self.shape[:-1] + (self.n // 2 + 1,)

    def __call__(self, y, X=None):
        if X is None:
            X, y = y, None
        _require(X, _CPLX[self.dtype_code], "input")
        if self.shape is not None:
            what = "irfft" if self.normalized else "brfft"
            if tuple(X.shape) != self.cshape:
                raise DimensionMismatch(f"{what} input dims must be {self.cshape}, got {tuple(X.shape)}")
            if y is None:
                y = torch.empty(self.shape, dtype=_REAL[self.dtype_code], device=X.device)
            _require(y, _REAL[self.dtype_code], "output")
            if tuple(y.shape) != self.shape:
                raise DimensionMismatch(f"plan dims {self.shape}, output {tuple(y.shape)}")
            return self._exec(y, X)
        nin = (self.n // 2 + 1) * self.batch
        if X.numel() != nin:
            raise DimensionMismatch(f"irfft input must have length {nin}, got {X.numel()}")
        if y is None:
            shape = (self.n,) if self.batch == 1 else (self.batch, self.n)
            y = torch.empty(shape, dtype=_REAL[self.dtype_code], device=X.device)
        _require(y, _REAL[self.dtype_code], "output")
        if y.numel() != self.n * self.batch:
            raise DimensionMismatch(f"plan length {self.n * self.batch}, output length {y.numel()}")
        return self._exec(y, X)

    def __mul__(self, X):
        return self(X)


class InplacePlan:
    """In-place wrapper (api.jl:15-24, 42-49): ip * x transforms x in place."""

    def __init__(self, p):
        self.p = p

    def __mul__(self, x):
        self.p(x, x)
        return x

    def __call__(self, y, x=None):
        if x is None:
            return self * y
        self.p(y, x)
        return y


class Scaled:
    """A plan times a constant, as inv returns (api.jl:26-31); .p and .scale
    are public and .p stays the unnormalized plan.  Application runs one
    fused plan whose kernels fold the combined scale into their stores."""

    def __init__(self, p, scale):
        self.p = p
        self.scale = scale

    def _fused(self):
        inner = self.p.p if isinstance(self.p, InplacePlan) else self.p
        return inner._with_scale(self.scale)

    def __mul__(self, x):
        f = self._fused()
        if isinstance(self.p, InplacePlan):
            f(x, x)
            return x
        return f(x)

    def __call__(self, y, x=None):
        if x is None:
            return self * y
        self._fused()(y, x)
        return y


def mul_(y, p, x):
    """mul!(y, p, x): apply p to x into y and return y (api.jl:34-55)."""
    if isinstance(p, (InplacePlan, Scaled)):
        p(y, x)
        return y
    return p(y, x)


def inv(p):
    """inv(p) (api.jl:72-94, 170-180): forward -> Scaled(bfft, 1/n) (so
    inv(p).p is the unnormalized backward plan); ifft -> plain fft;
    bfft -> Scaled(fft, 1/n); inv(c*P) = inv(P)/c."""
    if isinstance(p, InplacePlan):
        s = inv(p.p)
        return Scaled(InplacePlan(s.p), s.scale) if isinstance(s, Scaled) else InplacePlan(s)
    if isinstance(p, Scaled):
        q = inv(p.p)
        return Scaled(q.p, q.scale / p.scale) if isinstance(q, Scaled) else Scaled(q, 1.0 / p.scale)
    if isinstance(p, FFTPlan):
        if p.sign == -1:
            return Scaled(p._sibling(+1, 1.0), 1.0 / p.n)
        if p.scale == 1:
            return Scaled(p._sibling(-1, 1.0), 1.0 / p.n)
        return p._sibling(-1, 1.0)
    raise ArgumentError(f"inv is not defined for {type(p).__name__}")


def size(p):
    """size(p) (api.jl:57-60, 166-168)."""
    if isinstance(p, (InplacePlan,)):
        return size(p.p)
    if isinstance(p, Scaled):
        return size(p.p)
    if isinstance(p, FFTPlan):
        return p.shape
    if isinstance(p, RfftPlan):
        return (p.n,) if p.shape is None else p.shape
    if isinstance(p, IrfftPlan):
        return (p.n // 2 + 1,) if p.shape is None else p.cshape
    raise ArgumentError(f"size is not defined for {type(p).__name__}")


# ------------------------------------------------------------ constructors

def _check_region(region, N):
    # _checkregion_nd (api.jl:103-105): N-d plans transform all dims
    if region is None:
        return
    r = tuple(region) if not isinstance(region, int) else (region,)
    if r != tuple(range(N)):
        raise ArgumentError(f"FFTJ N-d plans transform all dims (expected region {tuple(range(N))}, got {r})")


def _positive_n(n):
    n = int(n)
    if n < 1:
        raise ArgumentError(f"transform length must be positive (got {n})")
    return n


def _precision_flags(precision):
    """precision=None (plain FP32/FP64 stages) or "f16" / "bf16" / "bf16x3":
    radix-9/16/32 stages of an FP32 plan as split-operand MFMA products."""
    if precision is None:
        return 0
    if precision not in _lib.PRECISION_FLAGS:
        raise ArgumentError(f"precision must be None, 'f16', 'bf16' or 'bf16x3' (got {precision!r})")
    return _lib.PRECISION_FLAGS[precision]


def _route_flags(route):
    """route=None (the planner's choice) or, for a length that is not
    13-smooth, "composite" (one exact step over its smooth part around the
    resampled rest, composite.jl), "resample" (the whole length resampled,
    resample.jl) or "hvdh" (FP32: Gaussian resampling onto a coprime smooth
    grid and chunked banded solves, hvdh.jl; an even length takes one exact
    step over its smooth part around HvdH transforms of the odd rest, the
    reference's SplitPlan, composite.jl:103-200)."""
    if route is None:
        return 0
    if route not in _lib.ROUTE_FLAGS:
        raise ArgumentError(f"route must be None, 'composite', 'resample' or 'hvdh' (got {route!r})")
    return _lib.ROUTE_FLAGS[route]


def _plan_c2c(T_or_x, n, sign, scale_kind, region, batch, planner=None, precision=None, route=None):
    flags = _precision_flags(precision) | _route_flags(route)
    if isinstance(T_or_x, torch.Tensor):
        x = T_or_x
        if not x.is_complex():
            raise ArgumentError("complex plans need a complex array")
        code = _code_of_tensor(x)
        N = x.dim()
        _check_region(region, N)
        shape = tuple(x.shape)
        dims = (x.numel(),) if N <= 1 else tuple(reversed(shape))
        if any(d < 1 for d in dims):
            raise ArgumentError(f"dimensions must be positive (got {shape})")
        length = math.prod(dims)
        scale = 1.0 / length if scale_kind == "inv" else 1.0
        return FFTPlan(code, dims, sign, scale, 1, shape if N > 1 else (length,), flags=flags)
    code = _code(T_or_x)
    n = _positive_n(n)
    scale = 1.0 / n if scale_kind == "inv" else 1.0
    tuning = None if flags else _planner_tuning(planner, code, n, batch, sign)
    return FFTPlan(code, (n,), sign, scale, batch, tuning=tuning, flags=flags)


def _planner_tuning(planner, code, n, batch, sign):
    """Search-space choice for a 1-D complex plan.  None: the learned
    planner's pick where it is confident (FP32, 13-smooth lengths 1024..2^24
    at n*batch near its training cases, a committed model) and the pick is
    predicted faster than the library's own schedule by the model's
    calibrated margin, else the library's own schedule; "heuristic": always
    the library's own schedule; "learned": the learned pick for any FP32
    length (FP64 keeps the heuristic); "exhaustive": times every candidate
    on the current stream and keeps the fastest."""
    if planner == "heuristic":
        return None
    if planner not in (None, "learned", "exhaustive"):
        raise ArgumentError(f"planner must be None, 'heuristic', 'learned' or 'exhaustive' (got {planner!r})")
    from . import planner as _pl
    if planner == "exhaustive":
        return _pl._tuning(_pl.exhaustive_choose(n, batch, sign, dtype=code)[1])
    if code != _lib.F32:
        return None
    if planner is None:
        if not _pl.model_available():
            return None
        lp = _pl.default_learned()
        if not lp.confident(n, batch):
            return None
        return _pl._tuning(lp.pick_cached(n, batch, sign))
    return _pl._tuning(_pl.default_learned().pick(n, batch, sign)[0])


def plan_fft(T_or_x, n=None, region=None, *, batch=1, nthreads=None, planner=None, precision=None,
              route=None, **flags):
    """plan_fft(T, n) (plan.jl:207-208) or plan_fft(x[, region]) (api.jl:107-112).
    planner= picks the schedule search for 1-D plans of a given length;
    precision= selects split-precision MFMA stages for FP32 plans."""
    return _plan_c2c(T_or_x, n, -1, "one", region, batch, planner, precision, route)


def plan_ifft(T_or_x, n=None, region=None, *, batch=1, nthreads=None, planner=None, precision=None,
               route=None, **flags):
    """plan_ifft (plan.jl:209-210, api.jl:119-124): inverse scaled by 1/n."""
    return _plan_c2c(T_or_x, n, +1, "inv", region, batch, planner, precision, route)


def plan_bfft(T_or_x, n=None, region=None, *, batch=1, nthreads=None, planner=None, precision=None,
               route=None, **flags):
    """plan_bfft (plan.jl:212-213, api.jl:113-118): unnormalized inverse."""
    return _plan_c2c(T_or_x, n, +1, "one", region, batch, planner, precision, route)


def plan_fft_(T_or_x, n=None, region=None, **kw):
    """plan_fft! (api.jl:19-20, 125-126)."""
    return InplacePlan(plan_fft(T_or_x, n, region, **kw))


def plan_ifft_(T_or_x, n=None, region=None, **kw):
    return InplacePlan(plan_ifft(T_or_x, n, region, **kw))


def plan_bfft_(T_or_x, n=None, region=None, **kw):
    return InplacePlan(plan_bfft(T_or_x, n, region, **kw))


def plan_rfft(T_or_x, n=None, region=None, *, batch=1, nthreads=None, **flags):
    """plan_rfft(T, n) (real.jl:41-52) or plan_rfft(x) (api.jl:131-136)."""
    if isinstance(T_or_x, torch.Tensor):
        x = T_or_x
        if x.is_complex():
            raise ArgumentError("rfft plans need a real array")
        _check_region(region, x.dim())
        if x.dim() > 1:
            if any(d < 1 for d in x.shape):
                raise ArgumentError(f"dimensions must be positive (got {tuple(reversed(x.shape))})")
            return RfftPlan(_code_of_tensor(x), x.shape[-1], 1, tuple(x.shape))
        return RfftPlan(_code_of_tensor(x), _positive_n(x.numel()), 1)
    return RfftPlan(_code(T_or_x), _positive_n(n), batch)


def _plan_c2r(T_or_X, n, normalized, region, batch):
    if isinstance(T_or_X, torch.Tensor):
        X = T_or_X
        _check_region(region, X.dim())
        n = int(n)
        lead = X.shape[-1] if X.dim() else 1
        if lead != n // 2 + 1:
            what = "irfft" if normalized else "brfft"
            raise DimensionMismatch(f"{what} of leading length {n} needs {n // 2 + 1} rows, got {lead}")
        if X.dim() > 1:
            shape = tuple(X.shape[:-1]) + (n,)
            if any(d < 1 for d in shape):
                raise ArgumentError(f"dimensions must be positive (got {tuple(reversed(shape))})")
            return IrfftPlan(_code_of_tensor(X), n, normalized, 1, shape)
        return IrfftPlan(_code_of_tensor(X), _positive_n(n), normalized, 1)
    return IrfftPlan(_code(T_or_X), _positive_n(n), normalized, batch)


def plan_irfft(T_or_X, n, region=None, *, batch=1, nthreads=None, **flags):
    """plan_irfft(T, n) (real.jl:57-70) or plan_irfft(X, n1) (api.jl:146-154)."""
    return _plan_c2r(T_or_X, n, True, region, batch)


def plan_brfft(T_or_X, n, region=None, *, batch=1, nthreads=None, **flags):
    """plan_brfft(T, n) (real.jl:72-73) or plan_brfft(X, n1) (api.jl:137-145)."""
    return _plan_c2r(T_or_X, n, False, region, batch)


# ---------------------------------------------------------------- one-shots

def _as_complex(x):
    # promoting fallbacks (api.jl:192-203): integers and reals -> float complex
    if not x.is_complex():
        if not x.is_floating_point():
            x = x.to(torch.float64)
        x = x.to(torch.complex64 if x.dtype == torch.float32 else torch.complex128)
    return x.contiguous()


def _as_real(x):
    if not x.is_floating_point():
        x = x.to(torch.float64)
    return x.contiguous()


def fft(x):
    x = _as_complex(x)
    return plan_fft(x) * x


def ifft(x):
    x = _as_complex(x)
    return plan_ifft(x) * x


def bfft(x):
    x = _as_complex(x)
    return plan_bfft(x) * x


def fft_(x):
    """fft!(x) (api.jl:97): in place."""
    return plan_fft_(x) * x


def ifft_(x):
    return plan_ifft_(x) * x


def bfft_(x):
    return plan_bfft_(x) * x


def rfft(x):
    x = _as_real(x)
    return plan_rfft(x) * x


def irfft(X, n):
    X = _as_complex(X)
    return plan_irfft(X, n) * X


def brfft(X, n):
    X = _as_complex(X)
    return plan_brfft(X, n) * X


# backend-neutral utilities the reference re-exports (FFTJ.jl:8)
fftshift = torch.fft.fftshift
ifftshift = torch.fft.ifftshift
fftfreq = torch.fft.fftfreq
rfftfreq = torch.fft.rfftfreq
