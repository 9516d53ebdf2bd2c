"""Seeded synthetic inputs shared by the tests, the oracle and the benchmark.

The generator is the splitmix64 stream: draw i (0-based) of seed s is
mix64(s + (i + 1) * 0x9e3779b97f4a7c15), mix64 being splitmix64's output
function -- the same numbers splitmix64 seeded with s returns in sequence,
computed here counter-wise so numpy produces millions at once.
  uniform:  u = (z >> 11) * 2^-53,  value = 2u - 1  in [-1, 1)   (FP64)
  complex:  re is drawn before im, element by element
  normal:   Box-Muller on consecutive draw pairs (u1 from (0, 1], u2 from [0, 1))
FP32 inputs are the FP64 draws rounded to FP32.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, count, first=0):
    with np.errstate(over="ignore"):
        i = np.arange(first + 1, first + 1 + count, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _unit(z):
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform(seed, count, first=0):
    """count FP64 values uniform in [-1, 1)."""
    return 2.0 * _unit(splitmix64(seed, count, first)) - 1.0


def uniform_complex(seed, n, dtype=np.complex128):
    v = uniform(seed, 2 * n)
    out = (v[0::2] + 1j * v[1::2])
    return out.astype(dtype)


def normal(seed, count):
    """count FP64 standard normals (Box-Muller on draw pairs)."""
    m = (count + 1) // 2
    z = splitmix64(seed, 2 * m)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = _unit(z[1::2])
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2 * np.pi * u2)
    out[1::2] = r * np.sin(2 * np.pi * u2)
    return out[:count]
