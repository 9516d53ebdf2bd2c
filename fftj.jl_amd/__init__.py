"""fftj_amd -- MI355X (gfx950) FFT engine behind FFTJ.jl's plan/apply API.

The kernels live in libfftjh.so (HIP, built in-tree by build.py); this
package is the host-side mirror of the reference's Julia API over its C-ABI
(include/fftjh.h).  The directory is named fftj.jl_amd; load it as the
module `fftj_amd` (see load_package in __graft_entry__.py).
"""
from . import _lib, planner, ragged, rng, tune
from ._lib import ArgumentError, DimensionMismatch, HipError, UnsupportedError
from .api import (FFTPlan, InplacePlan, IrfftPlan, RfftPlan, Scaled, bfft, bfft_, brfft, fft, fft_, fftfreq,
                  fftshift, ifft, ifft_, ifftshift, inv, irfft, mul_, plan_bfft, plan_bfft_, plan_brfft, plan_fft,
                  plan_fft_, plan_ifft, plan_ifft_, plan_irfft, plan_rfft, rfft, rfftfreq, size)

__all__ = [
    "ArgumentError", "DimensionMismatch", "HipError", "UnsupportedError", "FFTPlan", "InplacePlan", "IrfftPlan",
    "RfftPlan", "Scaled", "bfft", "bfft_", "brfft", "fft", "fft_", "fftfreq", "fftshift", "ifft", "ifft_",
    "ifftshift", "inv", "irfft", "mul_", "plan_bfft", "plan_bfft_", "plan_brfft", "plan_fft", "plan_fft_",
    "plan_ifft", "plan_ifft_", "plan_irfft", "plan_rfft", "rfft", "rfftfreq", "size", "rng", "ragged",
    "plan_ragged", "RaggedPlan", "planner",
]
from .ragged import RaggedPlan, plan_ragged  # noqa: E402
