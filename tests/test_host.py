"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/fftjh.h declares, argument errors come back as
the reference's exception types before any GPU work, and the planner's
schedules for the benchmark configurations."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, F

_lib = F._lib


def _declared():
    src = open(os.path.join(ROOT, "include", "fftjh.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(fftjh_\w+)\s*\(", src, re.M)))


def test_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 12
    L = _lib.lib()
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.EXPORTS)
    assert L.fftjh_version() >= 100


def test_argument_errors_before_gpu_work():
    # n = 0 -> ArgumentError (runtests.jl:424), raised on the host
    with pytest.raises(F.ArgumentError, match="positive"):
        F.plan_fft("f64", 0)
    with pytest.raises(F.ArgumentError):
        F.plan_rfft("f32", -3)
    with pytest.raises(F.ArgumentError):
        F.plan_fft("f16", 8)
    with pytest.raises(F.ArgumentError, match="positive"):
        _lib.describe_request(_lib.KIND_C2C, _lib.F32, [0])
    with pytest.raises(F.ArgumentError):
        _lib.describe_request(_lib.KIND_C2C, _lib.F32, [8], batch=0)
    with pytest.raises(F.ArgumentError):
        _lib.describe_request(_lib.KIND_C2C, 7, [8])


def test_execute_errors_without_a_plan():
    L = _lib.lib()
    assert L.fftjh_execute(None, None, None, None) == _lib.ERR_ARGUMENT
    assert "plan is NULL" in L.fftjh_last_error().decode()
    assert L.fftjh_plan_destroy(None) == _lib.OK


def _passes(desc):
    return [ln.strip() for ln in desc.splitlines()[1:] if ln.strip()]


def test_schedule_config1_single_launch():
    d = _lib.describe_request(_lib.KIND_C2C, _lib.F64, [1024])
    steps = _passes(d)
    assert len(steps) == 1 and "n=1024" in steps[0] and "radices=16,16,4" in steps[0] and "kernel=pow2" in steps[0]


def test_schedule_config2():
    # (a) 4096 x 16384 contiguous: one launch, one transform per workgroup row
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [4096], batch=16384))
    assert len(s) == 1 and "batch=16384" in s[0]
    # (b) 2^24: three 256-point passes (nested four-step) on the lean kernel
    # inside y: columns of x to rows of y with the w_65536 twiddle, then two
    # in-place column passes (the second with w_2^24); an in-place call
    # (x == y) runs the same launches through s0
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [1 << 24]))
    assert len(s) == 3 and all("n=256" in p and "kernel=lean" in p for p in s)
    assert s[0].startswith("pass x->y") and s[1].startswith("pass y->y") and s[2].startswith("pass y->y")
    assert "in_stride=65536" in s[0] and "rows_out" in s[0] and "twiddle=w_65536" in s[0]
    assert "in_stride=256" in s[1] and "twiddle=w_16777216" in s[1]
    assert "in_stride=65536" in s[2] and "out_stride=65536" in s[2] and "twiddle" not in s[2]
    assert s[0].endswith("aliased=x->s0") and s[1].endswith("aliased=s0->s0") and s[2].endswith("aliased=s0->y")
    # FP64 keeps the generic three-pass layout
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F64, [1 << 24]))
    assert len(s) == 3 and all("kernel=pow2" in p for p in s) and s[1].startswith("pass s0->s0")
    # (c) 10^6 = 100 x 100 x 100
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [10**6], batch=8))
    assert len(s) == 3 and all("n=100 " in p for p in s)


def test_schedule_config3():
    # the deconvolution rides on the first (power-of-two) pass of the m-point
    # transform; the interpolation is its own launch.  The 147,456-point grid
    # x 4 (4.5 MiB) is a small footprint: two passes, 64 x 2304
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [65537], batch=4))
    assert len(s) == 3 and s[0].startswith("pass x->") and "deconv n=65537 m=147456" in s[0]
    assert " n=64 " in s[0] and " n=2304 " in s[1]
    assert s[-1].startswith("rs_interp")
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [1000003], batch=4))
    assert "deconv n=1000003 m=2097152" in s[0]
    # a grid of three lean factors (2^21 = 128^3) runs the lean layout from x
    # into the grid, the first pass loading through the deconvolution
    assert len(s) == 4 and all("kernel=lean" in p for p in s[:3]) and "rows_out" in s[0]
    assert s[0].startswith("pass x->s0") and s[2].startswith("pass s0->s0") and s[3].startswith("rs_interp s0->y")
    # a generic first pass keeps the deconvolution as a launch of its own
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F64, [10007], batch=3))
    assert s[0].startswith("rs_deconv") and "deconv" not in s[1]
    # the real edges ride on the half-length transform's last / first pass
    # (paired columns), three launches each
    s = _passes(_lib.describe_request(_lib.KIND_R2C, _lib.F32, [1 << 22], batch=4))
    assert len(s) == 3 and "real=r2c-paired" in s[-1] and "C=32 " in s[-1] and "->y " in s[-1]
    s = _passes(_lib.describe_request(_lib.KIND_C2R, _lib.F32, [1 << 22], batch=4))
    assert len(s) == 3 and "real=c2r-paired" in s[0] and "x->" in s[0] and "scale=0.5" in s[0]
    # a generic last pass keeps the twist as its own launch
    s = _passes(_lib.describe_request(_lib.KIND_R2C, _lib.F32, [3 * 2**17], batch=2))
    assert s[-1].startswith("r2c_twist") and not any("paired" in p for p in s)
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [256, 256, 256]))
    assert len(s) == 3 and "in_stride=1 " in s[0] and "in_stride=256 " in s[1] and "in_stride=65536 " in s[2]


def test_lean_paired_real_schedule():
    # the paired real edges of 64- and 128-point multi-pass passes run on the
    # lean pair kernel (two 16-column sub-groups, 2n threads); 256-point ones
    # and C2R passes whose twiddle period is not a power of two keep pow2
    for n, batch, kind, edge in ((1 << 22, 4, _lib.KIND_R2C, -1), (1 << 22, 4, _lib.KIND_C2R, 0),
                                 (1 << 20, 16, _lib.KIND_R2C, -1), (1 << 20, 16, _lib.KIND_C2R, 0),
                                 (786432, 3, _lib.KIND_R2C, -1)):
        s = _passes(_lib.describe_request(kind, _lib.F32, [n], batch=batch))
        p = s[edge]
        assert "kernel=lean" in p and "paired" in p and "C=32 " in p, p
        m = int(p.split(" n=")[1].split()[0])
        assert f"threads={2 * m} " in p and m in (64, 128)
    s = _passes(_lib.describe_request(_lib.KIND_R2C, _lib.F32, [1 << 18], batch=64))
    assert "kernel=pow2 real=r2c-paired" in s[-1]
    s = _passes(_lib.describe_request(_lib.KIND_C2R, _lib.F32, [196608], batch=3))
    assert "kernel=pow2 real=c2r-paired" in s[0] and "twiddle=w_3072" in s[0]


def test_lean_rows_and_columns_schedule():
    # FP32 rows of 64/128/256 points in groups of 16 take the lean kernel with
    # both edges transposed through LDS; N-d axes whose column blocks hold a
    # multiple of 16 columns take its column form; the rest keep pow2
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [256, 256, 256]))
    assert "kernel=lean rows_in rows_out" in s[0] and "in_stride=1 out_stride=1" in s[0]
    assert all("kernel=lean" in p and "rows_" not in p for p in s[1:])
    assert "in_stride=256 " in s[1] and "in_stride=65536 " in s[2]
    for n, batch in ((64, 48), (128, 4096), (256, 16)):
        s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [n], batch=batch))
        assert len(s) == 1 and "kernel=lean rows_in rows_out" in s[0] and f"groups={batch // 16} " in s[0]
    for n, batch in ((128, 40), (512, 64), (256, 1)):
        s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [n], batch=batch))
        assert "kernel=lean" not in s[0]
    # FP64 keeps the power-of-two kernel; a block of 129 columns ends in a
    # partial group of one column (9 groups per block)
    assert "kernel=lean" not in _lib.describe_request(_lib.KIND_C2C, _lib.F64, [256], batch=64)
    s = _passes(_lib.describe_request(_lib.KIND_R2C, _lib.F32, [256, 256, 256]))
    assert "kernel=lean" in s[1] and "in_stride=129 " in s[1] and "groups=2304 " in s[1] and "kernel=lean" in s[2]
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [33, 64, 5]))
    assert "n=64 " in s[1] and "kernel=lean" in s[1] and "groups=15 " in s[1]


def test_lean_square_schedule():
    # 10^6 = 100^3 (FP32): the lean layout on the square kernel, 100 = 10 x 10,
    # 50 columns per group of 500 threads; twiddle periods 10^4 and 10^6
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [10**6], batch=8))
    assert len(s) == 3 and all("n=100 radices=10,10 kernel=lean" in p and "C=50 threads=500" in p for p in s)
    assert "rows_out" in s[0] and "twiddle=w_10000 " in s[0] and "twiddle=w_1000000 " in s[1]
    assert "twiddle" not in s[2] and s[2].endswith("aliased=s0->y")
    # FP64 keeps the fused chain kernel
    assert "kernel=lean" not in _lib.describe_request(_lib.KIND_C2C, _lib.F64, [10**6], batch=8)


def test_stream_hints_rule():
    # footprints of 384 MiB and up: the pass reading x streams its loads, the
    # pass writing y last streams its stores; smaller plans keep the caches
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F64, [1 << 24]))
    assert s[0].endswith("stream=in") and "stream=" not in s[1] and s[2].endswith("stream=out")
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [4096], batch=16384))
    assert len(s) == 1 and s[0].endswith("stream=inout")
    for kind, dims, batch in ((_lib.KIND_C2C, [256, 256, 256], 1), (_lib.KIND_C2C, [10**6], 8),
                              (_lib.KIND_R2C, [1 << 22], 4), (_lib.KIND_C2C, [1000003], 4)):
        assert "stream=" not in _lib.describe_request(kind, _lib.F32, dims, batch)


def test_fixed_chain_schedule():
    # passes of 100 = 10 x 10 outside the lean layout (a tuned 10^6 split,
    # the axes of a 100^3 array) run on the compile-time chain kernel (10
    # elements per thread, 32 transforms per workgroup); FP64's 320 threads
    # would exceed its limit, so it keeps the runtime chain
    t = _lib.make_tuning([100, 100, 100], (), 0)
    s = _passes(_lib.describe_tuned(_lib.F32, [10**6], batch=8, tuning=t))
    assert len(s) == 3 and all("n=100 radices=10,10 kernel=chain" in p and "C=32 threads=320" in p for p in s)
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [100, 100, 100]))
    assert len(s) == 3 and all("n=100 radices=10,10 kernel=chain" in p for p in s)
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F64, [10**6], batch=2))
    assert len(s) == 3 and not any("kernel=chain" in p for p in s)


def test_chain_column_split():
    # out of the Infinity Cache (128 MiB per array and more), 3-smooth-heavy
    # lengths put their strided passes on the 81-point chain; in cache, or
    # where no such split exists, the balanced split stays
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [4 * 3**12], batch=64))
    assert [p.split(" n=")[1].split()[0] for p in s] == ["81", "81", "324"]
    assert "kernel=chain" in s[0] and "kernel=chain" in s[1] and "in_stride=1 " in s[2]
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [4 * 3**12], batch=2))
    assert [p.split(" n=")[1].split()[0] for p in s] == ["162", "162", "81"]
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F64, [4 * 3**12], batch=64))
    assert "n=324 " not in " ".join(s)
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [1 << 24], batch=8))
    assert all("n=81 " not in p for p in s)


def test_fixed_chain_gives_way_to_generic():
    # the 81 = 9 x 9 chain runs three threads per transform, so only 64
    # columns fill its waves; 64 columns take it, 40 fall back to the generic
    # kernel with compact radices 9,9 (and, a small pass whose groups would
    # not cover the CUs, one column per group)
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [64, 81]))
    assert "n=81 radices=9,9 kernel=chain" in s[1] and "C=64 " in s[1]
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [40, 81]))
    assert "n=81 radices=9,9" in s[1] and "kernel=chain" not in s[1] and "C=1 " in s[1] and "groups=40 " in s[1]
    # a tuned group size the chain cannot take runs the generic kernel
    # instead of refusing the plan; one that fits no kernel is still refused
    t = _lib.make_tuning((81, 6561), [{}, {"C": 4}])
    s = _passes(_lib.describe_tuned(_lib.F32, [3**12], batch=64, tuning=t))
    assert "n=81 radices=9,9" in s[1] and "kernel=chain" not in s[1] and "C=4 " in s[1]
    with pytest.raises(_lib.ArgumentError):
        _lib.describe_tuned(_lib.F32, [3**12], batch=64, tuning=_lib.make_tuning((81, 6561), [{"C": 4}, {}]))


def test_nd_long_and_nonsmooth_axes():
    # a long strided axis runs the multi-pass schedule on its column block
    # directly (offsets scaled by the column count), no gather / scatter
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [2, 1 << 20]))
    assert len(s) == 4 and not any(p.startswith("strided_copy") for p in s)
    assert s[1].startswith("pass y->y") and "in_stride=16384 " in s[1] and "C=16 " in s[1]
    assert s[2].startswith("pass y->s0") and s[3].startswith("pass s0->y") and "out_stride=32768" in s[3]
    # a strided axis of 4096 or more splits into 16-column passes
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [4096, 4096]))
    assert len(s) == 3 and "n=64 " in s[1] and "C=16 " in s[1] and "in_stride=262144 " in s[1]
    assert "in_stride=4096 " in s[2] and "out_stride=262144" in s[2]
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F64, [1024, 2048]))
    assert len(s) == 2
    # a non-smooth axis resamples with strided deconvolution / interpolation
    s = _passes(_lib.describe_request(_lib.KIND_C2C, _lib.F32, [33, 17]))
    assert s[1].startswith("rs_deconv") and s[-1].startswith("rs_interp")


def test_nd_real_schedules():
    """N-d real plans: real rows along dims[0] then in-place complex passes
    of the other dims (R2C); complex passes into scratch then the real rows
    (C2R), irfft scale split 1/prod(dims[1:]) and 1/dims[0]."""
    # power-of-two half length: one launch for the rows, real edges fused
    s = _passes(_lib.describe_request(_lib.KIND_R2C, _lib.F32, [64, 48]))
    assert s[0].startswith("pass x->y") and "n=32 " in s[0] and "real=r2c" in s[0]
    assert s[1].startswith("pass y->y") and "n=48 " in s[1] and "in_stride=33 " in s[1] and len(s) == 2
    s = _passes(_lib.describe_request(_lib.KIND_C2R, _lib.F32, [64, 48]))
    assert s[0].startswith("pass x->s0") and "n=48 " in s[0] and "scale=0.0208333" in s[0]
    assert s[1].startswith("pass s0->y") and "real=c2r" in s[1] and "scale=0.015625" in s[1] and len(s) == 2
    # other even lengths: the half-length transform, then the twist
    s = _passes(_lib.describe_request(_lib.KIND_R2C, _lib.F32, [96, 48]))
    assert s[0].startswith("pass x->s0") and "n=48 " in s[0] and s[1].startswith("r2c_twist s0->y")
    s = _passes(_lib.describe_request(_lib.KIND_C2R, _lib.F32, [96, 48]))
    assert s[1].startswith("c2r_pre s0->s1") and s[2].startswith("pass s1->y") and "scale=0.0208333" in s[2]
    # odd leading length, three dims
    s = _passes(_lib.describe_request(_lib.KIND_R2C, _lib.F64, [9, 4, 3]))
    assert [p.split()[0] for p in s] == ["real_to_complex", "pass", "copy_prefix", "pass", "pass"]
    s = _passes(_lib.describe_request(_lib.KIND_C2R, _lib.F64, [9, 4, 3]))
    assert [p.split()[0] for p in s] == ["pass", "pass", "hermitian_fill", "pass", "take_real"]
    # every dim past the first is 1: the rows alone, straight from x
    s = _passes(_lib.describe_request(_lib.KIND_C2R, _lib.F32, [8, 1, 1]))
    assert s[0].startswith("c2r_pre x->s0") and len(s) == 2
    s = _passes(_lib.describe_request(_lib.KIND_C2R, _lib.F32, [64, 1, 1]))
    assert s[0].startswith("pass x->y") and "real=c2r" in s[0] and len(s) == 1
    with pytest.raises(F.ArgumentError, match="positive"):
        _lib.describe_request(_lib.KIND_R2C, _lib.F32, [8, 0])


def test_route_flags():
    """Non-smooth lengths: the composite route is one pass over the smooth
    part (twiddle w_n at its store, in place into y) and the resampled rest;
    the resample route resamples the whole length; at most one route bit."""
    L = _lib
    n = 64 * 10007
    s = _passes(L.describe_tuned(L.F32, (n,), 1, -1, L.make_tuning((), (), L.FLAG_ROUTE_COMPOSITE)))
    assert s[0].startswith("pass x->y n=64 ") and "in_stride=10007 " in s[0] and f"twiddle=w_{n}" in s[0]
    assert s[1].startswith("rs_deconv y->") and "n=10007 " in s[1] and s[-1].startswith("rs_interp") and "->y" in s[-1]
    def resampled_from_x(step, n):
        # the deconvolution as its own launch, or at the load of the first pass
        return (step.startswith("rs_deconv x->") and f"n={n} " in step) or \
            (step.startswith("pass x->") and f" deconv n={n} " in step)

    s = _passes(L.describe_tuned(L.F32, (n,), 1, -1, L.make_tuning((), (), L.FLAG_ROUTE_RESAMPLE)))
    assert resampled_from_x(s[0], n)
    # the planner's own choice: composite where the resampled length-p
    # transforms fit one pass and the whole length's would not
    s = _passes(L.describe_request(L.KIND_C2C, L.F32, [100 * 1009]))
    assert s[0].startswith("pass x->y n=100 ") and s[1].startswith("rs_deconv")
    s = _passes(L.describe_request(L.KIND_C2C, L.F32, [2 * 10007]))
    assert resampled_from_x(s[0], 2 * 10007)
    # a prime has no smooth part: both routes resample it
    s = _passes(L.describe_tuned(L.F32, (65537,), 1, -1, L.make_tuning((), (), L.FLAG_ROUTE_COMPOSITE)))
    assert resampled_from_x(s[0], 65537)
    with pytest.raises(L.ArgumentError):
        L.describe_tuned(L.F32, (n,), 1, -1, L.make_tuning((), (), L.FLAG_ROUTE_COMPOSITE | L.FLAG_ROUTE_RESAMPLE))
    with pytest.raises(F.ArgumentError):
        F.plan_fft("f32", n, route="bluestein")


def test_composite_radix_schedules():
    """Fewer stages with composite radices where the group size holds."""
    L = _lib
    def radices(n, batch):
        d = L.describe_request(L.KIND_C2C, L.F32, [n], batch=batch)
        return [ln.split("radices=")[1].split()[0] for ln in d.splitlines()[1:] if "radices=" in ln]
    assert radices(1000, 131072) == ["10,10,10"]
    assert radices(10 ** 6, 8) == ["10,10"] * 3
    assert radices(196, 100) == ["14,14"] and radices(225, 100) == ["15,15"] and radices(90, 1000) == ["6,15"]
    assert radices(12, 1 << 22) == ["12"]
    # powers of two keep the fused power-of-two chain
    assert radices(4096, 4) == ["16,16,16"]
    # tuned chains may name composite radices; radix 17 has no kernel
    d = L.describe_tuned(L.F32, (160,), 2, -1, L.make_tuning((), [{"radices": [10, 16]}]))
    assert "radices=10,16 " in d
    with pytest.raises(L.ArgumentError):
        L.describe_tuned(L.F32, (17 * 16,), 1, -1, L.make_tuning((), [{"radices": [17, 16]}]))


def test_region_check():
    import torch
    x = torch.zeros(4, 4, dtype=torch.complex64)
    with pytest.raises(F.ArgumentError, match="all dims"):
        F.plan_fft(x, region=(1,))


def test_mfma_flags_and_schedules():
    """Split-precision flags: one FP32 mode at a time; the chain takes radix
    16 and 9 stages on MFMA (radix 32 by tuning), the rest on VALU."""
    L = _lib
    for bits, name in ((L.FLAG_MFMA_F16, "mfma_f16"), (L.FLAG_MFMA_BF16, "mfma_bf16"),
                       (L.FLAG_MFMA_BF16X3, "mfma_bf16x3")):
        d = L.describe_tuned(L.F32, (4096,), 4, -1, L.make_tuning((), (), bits))
        assert f"kernel={name} " in d and "radices=16,16,16" in d
    d = L.describe_tuned(L.F32, (3 ** 6 * 2,), 2, -1, L.make_tuning((), (), L.FLAG_MFMA_F16))
    assert "radices=2,9,9,9" in d
    d = L.describe_tuned(L.F32, (1024,), 2, -1, L.make_tuning((), [{"radices": [32, 32]}], L.FLAG_MFMA_BF16))
    assert "radices=32,32 kernel=mfma_bf16 " in d
    with pytest.raises(L.ArgumentError):
        L.describe_tuned(L.F64, (4096,), 1, -1, L.make_tuning((), (), L.FLAG_MFMA_F16))
    with pytest.raises(L.ArgumentError):
        L.describe_tuned(L.F32, (4096,), 1, -1, L.make_tuning((), (), L.FLAG_MFMA_F16 | L.FLAG_MFMA_BF16))
    with pytest.raises(L.ArgumentError):
        L.describe_tuned(L.F32, (4096,), 1, -1, L.make_tuning((), (), 0x100))
    with pytest.raises(L.ArgumentError):  # radix 32 has no VALU kernel
        L.describe_tuned(L.F32, (1024,), 1, -1, L.make_tuning((), [{"radices": [32, 32]}]))
    # three-pass power-of-two plans take the lean kernel's MFMA variant; a
    # tuned split or FFTJH_LEAN_MAT=0 keeps the generic MFMA chain
    for bits, name in ((L.FLAG_MFMA_F16, "mfma_f16"), (L.FLAG_MFMA_BF16X3, "mfma_bf16x3")):
        s = _passes(L.describe_tuned(L.F32, (1 << 20,), 64, -1, L.make_tuning((), (), bits)))
        assert len(s) == 3 and all(f"kernel=lean kernel={name} " in p for p in s) and "rows_out" in s[0]
    s = _passes(L.describe_tuned(L.F32, (1 << 20,), 64, -1, L.make_tuning((1024, 1024), (), L.FLAG_MFMA_F16)))
    assert len(s) == 2 and not any("kernel=lean" in p for p in s)
    # every MFMA pass stays within the 512-thread bound of its kernel
    for n in (1 << 20, 3 ** 12 * 4, 1 << 16, 1458, 4374, 9 * 16 * 16 * 16):
        d = L.describe_tuned(L.F32, (n,), 64, -1, L.make_tuning((), (), L.FLAG_MFMA_BF16X3))
        for ln in d.splitlines():
            if "threads=" in ln:
                assert int(ln.split("threads=")[1].split()[0]) <= 512


def test_hvdh_route_schedule():
    # the reference's HvdhPlan parameters at the configuration primes
    # (hvdh.jl:154-190): t = 73,728 / 1,048,576, 6 taps, band 2, the probed
    # halo 9 / 13 rows; permute, spread, the t-point passes, solve, permute
    L = _lib.lib()
    import ctypes
    for n, t, halo in ((65537, 73728, 9), (1000003, 1048576, 13)):
        buf = ctypes.create_string_buffer(1 << 16)
        tune = _lib.make_tuning((), (), _lib.FLAG_ROUTE_HVDH)
        assert L.fftjh_describe_tuned(_lib.F32, 1, _lib.dims_array([n]), 4, -1, ctypes.byref(tune), buf,
                                      len(buf)) == _lib.OK
        s = _passes(buf.value.decode())
        assert s[0].startswith("hvdh_permute x->") and s[1].startswith("hvdh_spread")
        assert s[-2].startswith("hvdh_solve") and s[-1].startswith("hvdh_permute") and "->y " in s[-1]
        assert all(f"t={t} taps=6 band=2 halo={halo}" in p for p in (s[0], s[1], s[-2], s[-1]))
        assert all(p.startswith("pass") for p in s[2:-2])
    # by default the planner keeps resampling for these lengths
    assert "rs_interp" in _lib.describe_request(_lib.KIND_C2C, _lib.F32, [1000003], batch=4)
    # the route needs FP32 and a halo the solve kernel holds
    tune = _lib.make_tuning((), (), _lib.FLAG_ROUTE_HVDH)
    buf = ctypes.create_string_buffer(1 << 12)
    assert L.fftjh_describe_tuned(_lib.F64, 1, _lib.dims_array([65537]), 1, -1, ctypes.byref(tune), buf,
                                  len(buf)) == _lib.ERR_UNSUPPORTED


def test_hvdh_split_schedule():
    # an even length on the HvdH route (the reference's SplitPlan,
    # composite.jl:103-200): the twiddled length-s pass, then HvdH over the
    # s*batch length-p rows, the last permutation writing at stride s
    L = _lib.lib()
    import ctypes
    tune = _lib.make_tuning((), (), _lib.FLAG_ROUTE_HVDH)
    for n, batch, s in ((2 * 65537, 3, 2), (16 * 65537, 1, 16), (4 * 1000003, 2, 4)):
        buf = ctypes.create_string_buffer(1 << 16)
        assert L.fftjh_describe_tuned(_lib.F32, 1, _lib.dims_array([n]), batch, -1, ctypes.byref(tune), buf,
                                      len(buf)) == _lib.OK
        st = _passes(buf.value.decode())
        assert st[0].startswith(f"pass x->y n={s} ") and f"twiddle=w_{n}" in st[0]
        assert st[1].startswith("hvdh_permute y->") and f"batch={s * batch}" in st[1]
        assert st[-1].startswith("hvdh_permute") and "->y " in st[-1] and st[-1].endswith(f"out_stride={s}")
    # an odd length with a grid keeps the whole-length route
    buf = ctypes.create_string_buffer(1 << 16)
    assert L.fftjh_describe_tuned(_lib.F32, 1, _lib.dims_array([3 * 65537]), 1, -1, ctypes.byref(tune), buf,
                                  len(buf)) == _lib.OK
    assert _passes(buf.value.decode())[0].startswith("hvdh_permute x->")
    # no grid for the cofactor (10,007) or FP64: refused; by default resampling
    for dt, n in ((_lib.F32, 2 * 10007), (_lib.F64, 2 * 65537)):
        assert L.fftjh_describe_tuned(dt, 1, _lib.dims_array([n]), 1, -1, ctypes.byref(tune), buf,
                                      len(buf)) == _lib.ERR_UNSUPPORTED
    assert "hvdh" not in _lib.describe_request(_lib.KIND_C2C, _lib.F32, [2 * 65537])
