"""ctypes binding of libfftjh.so (include/fftjh.h).

The library is built in-tree by fftj.jl_amd/build.py.  There is no fallback:
if the library is missing, loading fails with an error that says how to
build it.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FFTJH_LIB names another build of the same library (kernel variants)
LIB_PATH = os.environ.get("FFTJH_LIB") or os.path.join(HERE, "lib", "libfftjh.so")

OK, ERR_ARGUMENT, ERR_DIMENSION, ERR_HIP, ERR_UNSUPPORTED = 0, 1, 2, 3, 4
F32, F64 = 0, 1
KIND_C2C, KIND_R2C, KIND_C2R, KIND_RAGGED = 0, 1, 2, 3
STATIC_SCRATCH_BYTES = 64 << 20  # FFTJH_STATIC_SCRATCH_BYTES of include/fftjh.h

# every entry point include/fftjh.h declares
EXPORTS = (
    "fftjh_plan_c2c", "fftjh_plan_r2c", "fftjh_plan_c2r", "fftjh_plan_r2c_nd", "fftjh_plan_c2r_nd",
    "fftjh_execute", "fftjh_execute_checked", "fftjh_execute_static",
    "fftjh_plan_info", "fftjh_plan_describe", "fftjh_describe_request", "fftjh_describe_tuned", "fftjh_plan_destroy",
    "fftjh_last_error", "fftjh_version", "fftjh_profile_launches", "fftjh_plan_c2c_tuned", "fftjh_plan_ragged",
    "fftjh_plan_ragged_tuned", "fftjh_mlp_create", "fftjh_mlp_destroy", "fftjh_mlp_param_count",
    "fftjh_mlp_set_params", "fftjh_mlp_get_params", "fftjh_mlp_predict", "fftjh_mlp_gradients",
    "fftjh_mlp_train_step", "fftjh_mlp_last_error",
)


class RaggedTuning(ctypes.Structure):
    _fields_ = [("groups", ctypes.c_int32), ("radix_pref", ctypes.c_int32), ("tile_max", ctypes.c_int32),
                ("pass_min", ctypes.c_int32), ("costs", ctypes.c_void_p)]


class PassTuning(ctypes.Structure):
    _fields_ = [("C", ctypes.c_int32), ("nradix", ctypes.c_int32), ("radices", ctypes.c_int32 * 16),
                ("xcd_remap", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("splits", ctypes.c_int64 * 4), ("pass_", PassTuning * 4), ("flags", ctypes.c_uint32),
                ("pad_", ctypes.c_uint32)]


FLAG_MFMA_F16, FLAG_MFMA_BF16, FLAG_MFMA_BF16X3 = 0x1, 0x2, 0x4
PRECISION_FLAGS = {"f16": FLAG_MFMA_F16, "bf16": FLAG_MFMA_BF16, "bf16x3": FLAG_MFMA_BF16X3}
FLAG_ROUTE_COMPOSITE, FLAG_ROUTE_RESAMPLE, FLAG_ROUTE_HVDH = 0x10, 0x20, 0x40
ROUTE_FLAGS = {"composite": FLAG_ROUTE_COMPOSITE, "resample": FLAG_ROUTE_RESAMPLE, "hvdh": FLAG_ROUTE_HVDH}


def make_tuning(splits=(), passes=(), flags=0):
    """passes: up to 4 dicts with optional keys C, radices, xcd_remap (0 planner, 1 off, 2 on);
    flags: FLAG_MFMA_* (split-precision MFMA stages)."""
    t = Tuning()
    t.flags = int(flags)
    for i, f in enumerate(splits or ()):
        t.splits[i] = int(f)
    for i, p in enumerate(passes):
        t.pass_[i].C = int(p.get("C", 0))
        rad = p.get("radices") or []
        t.pass_[i].nradix = len(rad)
        for j, r in enumerate(rad):
            t.pass_[i].radices[j] = int(r)
        t.pass_[i].xcd_remap = int(p.get("xcd_remap", 0))
    return t


class ArgumentError(ValueError):
    """Invalid plan parameters (the reference's ArgumentError)."""


class DimensionMismatch(ValueError):
    """Array lengths or shapes that do not match the plan (DimensionMismatch)."""


class HipError(RuntimeError):
    """A HIP runtime call inside the library failed."""


class UnsupportedError(NotImplementedError):
    """A valid request this library does not plan."""


_ERRS = {ERR_ARGUMENT: ArgumentError, ERR_DIMENSION: DimensionMismatch, ERR_HIP: HipError,
         ERR_UNSUPPORTED: UnsupportedError}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python fftj.jl_amd/build.py` "
                              "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        i64, vp, ci = ctypes.c_int64, ctypes.c_void_p, ctypes.c_int
        P64 = ctypes.POINTER(i64)
        L.fftjh_plan_c2c.argtypes = [ctypes.POINTER(vp), ci, ci, P64, i64, ci, ctypes.c_double, ctypes.c_uint]
        L.fftjh_plan_c2c_tuned.argtypes = [ctypes.POINTER(vp), ci, ci, P64, i64, ci, ctypes.c_double,
                                           ctypes.POINTER(Tuning)]
        L.fftjh_plan_ragged.argtypes = [ctypes.POINTER(vp), ci, i64, P64, P64, ci, ci, ctypes.c_uint]
        L.fftjh_plan_ragged_tuned.argtypes = [ctypes.POINTER(vp), ci, i64, P64, P64, ci, ci,
                                              ctypes.POINTER(RaggedTuning)]
        L.fftjh_mlp_create.argtypes = [ctypes.POINTER(vp), ci, ci]
        L.fftjh_mlp_destroy.argtypes = [vp]
        L.fftjh_mlp_param_count.argtypes = [vp]
        L.fftjh_mlp_param_count.restype = i64
        L.fftjh_mlp_set_params.argtypes = [vp, vp]
        L.fftjh_mlp_get_params.argtypes = [vp, vp]
        L.fftjh_mlp_predict.argtypes = [vp, vp, vp, ci, vp]
        L.fftjh_mlp_gradients.argtypes = [vp, vp, vp, ci, vp, ctypes.POINTER(ctypes.c_float), vp]
        L.fftjh_mlp_train_step.argtypes = [vp, vp, vp, ci, ctypes.c_float, ctypes.POINTER(ctypes.c_float), vp]
        L.fftjh_mlp_last_error.restype = ctypes.c_char_p
        L.fftjh_plan_r2c.argtypes = [ctypes.POINTER(vp), ci, i64, i64, ctypes.c_uint]
        L.fftjh_plan_c2r.argtypes = [ctypes.POINTER(vp), ci, i64, i64, ci, ctypes.c_uint]
        L.fftjh_plan_r2c_nd.argtypes = [ctypes.POINTER(vp), ci, ci, ctypes.POINTER(i64), i64, ctypes.c_uint]
        L.fftjh_plan_c2r_nd.argtypes = [ctypes.POINTER(vp), ci, ci, ctypes.POINTER(i64), i64, ci, ctypes.c_uint]
        L.fftjh_execute.argtypes = [vp, vp, vp, vp]
        L.fftjh_execute_checked.argtypes = [vp, vp, i64, vp, i64, vp]
        L.fftjh_execute_static.argtypes = [vp, vp, vp, vp]
        L.fftjh_plan_info.argtypes = [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci),
                                      ctypes.POINTER(ci), P64, P64, ctypes.POINTER(ctypes.c_double), P64, P64]
        L.fftjh_plan_describe.argtypes = [vp, ctypes.c_char_p, i64]
        L.fftjh_describe_request.argtypes = [ci, ci, ci, P64, i64, ci, ctypes.c_char_p, i64]
        L.fftjh_describe_tuned.argtypes = [ci, ci, P64, i64, ci, ctypes.POINTER(Tuning), ctypes.c_char_p, i64]
        L.fftjh_plan_destroy.argtypes = [vp]
        L.fftjh_last_error.restype = ctypes.c_char_p
        L.fftjh_profile_launches.argtypes = [vp, vp, vp, vp, ci, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ci)]
        _LIB = L
    return _LIB


def check(rc):
    if rc != OK:
        msg = lib().fftjh_last_error().decode()
        raise _ERRS.get(rc, RuntimeError)(msg)


def dims_array(dims):
    return (ctypes.c_int64 * max(1, len(dims)))(*dims)


def describe_request(kind, dtype, dims, batch=1, sign=-1):
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().fftjh_describe_request(kind, dtype, len(dims), dims_array(dims), batch, sign, buf, len(buf)))
    return buf.value.decode()


def describe_tuned(dtype, dims, batch=1, sign=-1, tuning=None):
    """Host-only schedule of a complex plan with search-space choices."""
    buf = ctypes.create_string_buffer(1 << 16)
    t = tuning if tuning is not None else Tuning()
    check(lib().fftjh_describe_tuned(dtype, len(dims), dims_array(dims), batch, sign, ctypes.byref(t), buf, len(buf)))
    return buf.value.decode()
