"""Write tests/golden/golden.npz: known-answer vectors for the oracle.

Expected outputs come from scipy.fft (pocketfft) evaluated in 80-bit long
double (np.clongdouble, eps 1.1e-19) -- an implementation independent of
the oracle -- on the project's seeded inputs (fftj.jl_amd/rng.py).  Stored
as complex128 (inputs are exactly representable; outputs rounded once).

    python tests/golden/make_golden.py
"""
import importlib.util
import os

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
spec = importlib.util.spec_from_file_location("fftj_rng", os.path.join(ROOT, "fftj.jl_amd", "rng.py"))
rng = importlib.util.module_from_spec(spec)
spec.loader.exec_module(rng)

C2C = [1, 2, 3, 8, 16, 60, 97, 105, 128, 240, 257, 1000, 1001, 1024, 1009, 2310, 4096, 4913, 8191, 10007]
REAL = [1, 2, 5, 64, 100, 255, 256, 999, 1000, 4096]
ND = [(16, 11), (11, 16), (12, 10, 14), (33, 17), (8, 1), (24, 40, 3)]


def main():
    out = {}
    # config 1: ComplexF64 N=1024, uniform [-1, 1) from seed 0
    x = rng.uniform_complex(0, 1024)
    out["cfg1_x"] = x
    out["cfg1_y"] = sf.fft(x.astype(np.clongdouble)).astype(np.complex128)
    for i, n in enumerate(C2C):
        x = rng.uniform_complex(100 + i, n)
        out[f"c2c_{n}_x"] = x
        out[f"c2c_{n}_fwd"] = sf.fft(x.astype(np.clongdouble)).astype(np.complex128)
        out[f"c2c_{n}_inv"] = sf.ifft(x.astype(np.clongdouble)).astype(np.complex128)
    for i, n in enumerate(REAL):
        x = rng.normal(200 + i, n)
        out[f"r2c_{n}_x"] = x
        out[f"r2c_{n}_y"] = sf.rfft(x.astype(np.longdouble)).astype(np.complex128)
    for i, shape in enumerate(ND):
        # shape is the column-major dims; the row-major array is reversed
        x = rng.uniform_complex(300 + i, int(np.prod(shape))).reshape(tuple(reversed(shape)))
        tag = "x".join(map(str, shape))
        out[f"nd_{tag}_x"] = x
        out[f"nd_{tag}_y"] = sf.fftn(x.astype(np.clongdouble)).astype(np.complex128)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
