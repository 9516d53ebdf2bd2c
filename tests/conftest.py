import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

F = __graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def golden():
    with np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def fj():
    return F


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def tol_for(prec, n):
    """Parity bar: FP64 plans <= 1e-13, FP32 plans <= 3e-6*log2(N) relative L2."""
    if prec == "f64":
        return 1e-13
    return 3e-6 * max(1.0, float(np.log2(max(n, 2))))


def cplx_np(prec):
    return np.complex64 if prec == "f32" else np.complex128


def real_np(prec):
    return np.float32 if prec == "f32" else np.float64


def torch_dtype(prec):
    import torch
    return torch.float32 if prec == "f32" else torch.float64


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def inputs_c(seed, count, prec):
    """Seeded complex inputs, rounded to the plan precision."""
    return F.rng.uniform_complex(seed, count).astype(cplx_np(prec))
