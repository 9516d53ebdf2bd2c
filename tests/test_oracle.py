"""Pins the CPU oracle before anything is checked against it.

Each oracle function is checked against something independent of it:
long-double scipy golden vectors, a brute-force long-double DFT, closed
forms (impulse, constant, tone), invariants (Parseval, linearity, round
trip), and the reference's own known answers (routing chains, nextsmooth,
chunk cover).  The last test plants mistakes in oracle outputs and shows
these checks catch them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from conftest import rel_l2

EPS64 = np.finfo(np.float64).eps


# ------------------------------------------------------------ chain (A1-A5)

def test_config1_golden(golden):
    y = O.fft(golden["cfg1_x"])
    assert rel_l2(y, golden["cfg1_y"]) < 1e-15


@pytest.mark.parametrize("n", [1, 2, 3, 8, 16, 60, 97, 105, 128, 240, 257, 1000, 1001, 1024, 1009, 2310, 4096,
                               4913, 8191, 10007])
def test_fft_golden(golden, n):
    x = golden[f"c2c_{n}_x"]
    smooth = O.chain_radices(n, 8) is not None
    tol = 1e-15 if smooth else 100 * EPS64  # resampling: the reference's 100 eps tier
    assert rel_l2(O.fft(x), golden[f"c2c_{n}_fwd"]) < tol
    assert rel_l2(O.fft(x, sign=1, scale=1.0 / n), golden[f"c2c_{n}_inv"]) < tol


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 9, 11, 12, 13, 15, 16, 24, 32, 35, 64, 77, 143, 128, 256])
def test_fft_vs_bruteforce_longdouble(n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    ref = O.dft_naive(x, -1, "ld")
    yl = O.fft(x, prec="ld")
    err_ld = np.sqrt(np.sum(np.abs(yl - ref) ** 2) / np.sum(np.abs(ref) ** 2))  # in long double
    assert err_ld < 1e-17
    assert rel_l2(O.fft(x), ref) < 1e-15
    refi = O.dft_naive(x, 1, "ld")
    assert rel_l2(O.fft(x, sign=1), refi) < 1e-15


@pytest.mark.parametrize("n", [16, 1000, 4096, 3 * 5 * 7 * 11 * 13])
def test_closed_forms(n):
    k = np.arange(n)
    for j0 in (0, 1, n // 3, n - 1):
        x = np.zeros(n, complex)
        x[j0] = 1
        ref = np.exp(-2j * np.pi * ((j0 * k) % n) / n)
        assert np.abs(O.fft(x) - ref).max() < 1e-14
    assert np.abs(O.fft(np.ones(n, complex)) - n * (k == 0)).max() < 1e-12 * n
    f = n // 5 + 1
    tone = np.exp(2j * np.pi * ((f * k) % n) / n)
    assert np.abs(O.fft(tone) - n * (k == f)).max() < 1e-11 * n


@pytest.mark.parametrize("n", [1024, 1000, 97, 65537])
def test_invariants(n):
    r = np.random.default_rng(7)
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    z = r.standard_normal(n) + 1j * r.standard_normal(n)
    X = O.fft(x)
    assert abs(np.vdot(X, X).real - n * np.vdot(x, x).real) < 1e-12 * n * np.vdot(x, x).real  # Parseval
    assert rel_l2(O.fft(2 * x - 3j * z), 2 * X - 3j * O.fft(z)) < 1e-13                      # linearity
    assert rel_l2(O.fft(X, sign=1, scale=1.0 / n), x) < 1e-13                                # round trip
    assert rel_l2(O.fft(X, sign=1), n * O.fft(X, sign=1, scale=1.0 / n)) < 1e-15             # bfft = n ifft


# ------------------------------------------------------- reference KATs (A1)

@pytest.mark.parametrize("n,W,chain", [
    (1024, 8, [4, 4, 8, 8]),                 # F64 1024 (the plain radix-2/4 case)
    (4096, 16, [4, 4, 4, 8, 8]),             # F32 4096
    (10**6, 16, [4, 4, 5, 5, 5, 5, 5, 5, 4]),  # F32 10^6 serial chain
    (256, 16, [4, 4, 4, 4]),
    (8, 8, [8]), (1, 8, []), (6, 8, [3, 2]), (429, 16, [3, 11, 13]),
])
def test_chain_kats(n, W, chain):
    assert O.chain_radices(n, W) == chain


def test_nextsmooth_kats():
    # resample.jl:12-21; the two config primes' grids (65537 -> 147456, 1000003 -> 2^21)
    assert O.nextsmooth(2 * 65537) == 147456
    assert O.nextsmooth(2 * 1000003) == 2097152
    assert O.nextsmooth(1) == 64
    for n in range(1, 3000, 37):
        m = O.nextsmooth(n)
        assert m >= n and m % 64 == 0


def test_gauss_legendre_exact_on_polynomials():
    z, w = O.gauss_legendre(20)
    for deg in range(0, 39, 2):
        assert abs(np.dot(w, z ** deg) - 2.0 / (deg + 1)) < 1e-14
    assert abs(np.dot(w, z ** 5)) < 1e-15


@pytest.mark.parametrize("length,nt,align", [(100, 4, 1), (1 << 15, 4, 16), (37, 8, 8), (5, 8, 1), (16, 3, 16)])
def test_chunk_cover(length, nt, align):
    # threads.jl:9-14 contract, as threading.jl:38-73 tests it
    spans = [O.chunk_range(length, nt, c, align) for c in range(nt)]
    spans = [s for s in spans if s[0] < s[1]]
    assert spans[0][0] == 0 and spans[-1][1] == length
    for (a, b), (c, d) in zip(spans, spans[1:]):
        assert b == c
    for lo, hi in spans:
        assert lo % align == 0
        assert hi - lo >= min(align, length)


def test_expi_exact():
    for L in (7, 1000, 2**24, 1000003):
        for e in (0, 1, L // 8, L // 4, L // 2, L - 1, 3 * L // 8 + 1):
            v = O.expi(e, L, -1)
            ref = np.exp(-2j * np.pi * (e / L))
            assert abs(v - ref) < 4e-16


# ------------------------------------------------------------ four-step (A7)

@pytest.mark.parametrize("n,N1,N2", [(4096, 64, 64), (4096, 256, 16), (4096, 16, 256), (3000, 60, 50),
                                     (2 * 3**7, 54, 81), (65536, 2048, 32)])
def test_fourstep_equals_dft(n, N1, N2):
    r = np.random.default_rng(n)
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    ref = O.fft(x)
    for sign in (-1, 1):
        y = O.fourstep(x, N1, N2, sign=sign)
        assert rel_l2(y, ref if sign == -1 else O.fft(x, sign=1)) < 1e-15
    assert np.array_equal(O.fourstep(x, N1, N2, nthreads=4), O.fourstep(x, N1, N2, nthreads=1))


# ------------------------------------------------------------- resample (A9)

@pytest.mark.parametrize("n", [17, 97, 257, 1009, 8191, 65537])
def test_resample_accuracy(n):
    # the reference's contract: 100 eps(T) relative (runtests.jl:72), design
    # target 10 eps(T) (resample.jl:203)
    r = np.random.default_rng(n)
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    ref = O.dft_naive(x, -1, "ld") if n <= 1009 else O.fft(x)  # smooth-free check below for large
    if n > 1009:
        import scipy.fft as sf
        ref = sf.fft(x.astype(np.clongdouble)).astype(np.complex128)
    assert rel_l2(O.resample(x), ref) < 100 * EPS64
    x32 = x.astype(np.complex64)
    y32 = O.resample(x32, prec="f32").astype(np.complex128)
    import scipy.fft as sf
    ref32 = sf.fft(x32.astype(np.clongdouble)).astype(np.complex128)
    assert rel_l2(y32, ref32) < 100 * np.finfo(np.float32).eps


# ----------------------------------------------------------------- real (A10)

@pytest.mark.parametrize("n", [1, 2, 5, 64, 100, 255, 256, 999, 1000, 4096])
def test_rfft_golden(golden, n):
    x = golden[f"r2c_{n}_x"]
    y = O.rfft(x)
    assert rel_l2(y, golden[f"r2c_{n}_y"]) < 1e-14
    assert rel_l2(O.irfft(y, n), x) < 1e-14
    assert rel_l2(O.irfft(y, n, normalized=False), n * O.irfft(y, n)) < 1e-14  # brfft = n irfft


# ------------------------------------------------------- HvdH (F1, hvdh.jl)

def test_hvdh_grid_lengths():
    # the reference's t for the configuration primes (t = 73,728 for 65,537
    # and 1,048,576 for 1,000,003): smooth, coprime to n, >= n(1 + 1/128)
    assert O.hvdh_t(65537) == 73728 and O.hvdh_t(1000003) == 1048576
    for n in (17, 101, 1009, 10007, 20014, 255255):
        t = O.hvdh_t(n)
        assert t >= n * (1 + 1 / 128) and np.gcd(t, n) == 1 and t < 2 * n
        assert O.nextsmooth(t) >= t
    assert O.hvdh_t(20014) % 2 == 1   # even n: an odd t


@pytest.mark.parametrize("n", [17, 101, 331, 1009, 2 * 10007, 65537])
def test_hvdh_matches_the_dft(n):
    # the Gaussian-resampling route against the DFT: brute force in long
    # double for small n, numpy's FFT beyond; both directions and a scale
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = O.dft_naive(x, prec="ld").astype(np.complex128) if n <= 1009 else np.fft.fft(x)
    assert rel_l2(O.hvdh(x), ref) < 1e-14
    assert rel_l2(O.hvdh(x, sign=1, scale=1.0 / n), np.fft.ifft(x)) < 1e-14
    assert rel_l2(O.hvdh(x.astype(np.complex64), prec="f32"), ref) < 1e-6


def test_hvdh_planted_mistakes():
    # chunks solved without their halo rows (a mistake planted in the
    # oracle) fail the DFT check; outputs off by one index or permuted do too
    n = 65537
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    y = O.hvdh(x)
    assert rel_l2(y, ref) < 1e-14
    assert rel_l2(O.hvdh(x, halo=0), ref) > 1e-10
    assert rel_l2(O.hvdh(x, halo=2), ref) > 1e-12
    assert rel_l2(np.roll(y, 1), ref) > 0.1
    assert rel_l2(y[(np.arange(n) * 2) % n], ref) > 0.1


@pytest.mark.parametrize("n", [2 * 1009, 6 * 331, 14 * 101, 10 * 1009, 2 * 10007, 6 * 65537])
def test_split_matches_the_dft(n):
    # SplitPlan (decimation in time around length-p HvdH) against the DFT:
    # brute force in long double for small n, numpy's FFT beyond; a wrong
    # twiddle sign or exponent, gather stride or output order fails here
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = O.dft_naive(x, prec="ld").astype(np.complex128) if n <= 2100 else np.fft.fft(x)
    assert rel_l2(O.split(x), ref) < 1e-14
    assert rel_l2(O.split(x, sign=1, scale=1.0 / n), np.fft.ifft(x)) < 1e-14
    assert rel_l2(O.split(x.astype(np.complex64), prec="f32"), ref) < 1e-6


def test_split_rejects_lengths_without_an_odd_cofactor():
    for n in (1024, 1009, 2 * 4):
        with pytest.raises(ValueError):
            O.split(np.ones(n, np.complex128))


# ------------------------------------------------------------------ N-d (A11)

@pytest.mark.parametrize("shape", [(16, 11), (11, 16), (12, 10, 14), (33, 17), (8, 1), (24, 40, 3)])
def test_nd_golden(golden, shape):
    tag = "x".join(map(str, shape))
    x = golden[f"nd_{tag}_x"]
    y = O.ndfft(x.ravel(), list(shape))
    assert rel_l2(y, golden[f"nd_{tag}_y"].ravel()) < 1e-14


# N-d real: the oracle composes its own 1-D rfft/irfft and the complex
# transform of the other axes; pinned against the full complex N-d oracle
# (the forward half-spectrum of a real array is the prefix of its complex
# spectrum along the halved axis), numpy's rfftn/irfftn, and round trips

@pytest.mark.parametrize("shape", [(48, 64), (17, 33), (11, 16), (16, 11), (1, 8), (8, 1), (5, 3, 128),
                                   (14, 10, 12), (3, 40, 24), (4, 3, 9), (6, 7)])
def test_nd_real_oracle(shape):
    rng = np.random.default_rng(len(shape) * 1000 + shape[-1])
    x = rng.standard_normal(shape)
    n1 = shape[-1]
    y = O.ndrfft(x)
    assert y.shape == shape[:-1] + (n1 // 2 + 1,)
    full = O.ndfft(x.astype(np.complex128).ravel(), list(reversed(shape))).reshape(shape)
    assert rel_l2(y, full[..., : n1 // 2 + 1]) < 1e-14
    assert rel_l2(y, np.fft.rfftn(x)) < 1e-14
    assert rel_l2(O.ndirfft(y, n1), x) < 1e-14
    assert rel_l2(O.ndirfft(y, n1), np.fft.irfftn(y, s=shape, axes=tuple(range(len(shape))))) < 1e-14
    assert rel_l2(O.ndirfft(y, n1, normalized=False), x.size * x) < 1e-14  # brfft = prod(dims) irfft


def test_nd_real_planted_mistakes():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((6, 5, 10))
    ref = np.fft.rfftn(x)
    # an axis left untransformed
    y = O.rfft(x)
    y = O._axis_fft(y, 1, -1, "f64")
    assert rel_l2(y, ref) > 0.1
    # the halved axis taken as the first (slowest) one instead of the last
    wrong = np.moveaxis(O.ndrfft(np.ascontiguousarray(np.moveaxis(x, 0, -1))), -1, 0)
    assert wrong.shape != ref.shape or rel_l2(wrong, ref) > 0.1
    # the inverse of the other axes with the forward sign
    Y = O.ndrfft(x)
    z = O._axis_fft(O._axis_fft(Y, 0, -1, "f64"), 1, -1, "f64") / 30
    assert rel_l2(O.irfft(z, 10), x) > 0.1


def test_threaded_baselines_match_the_oracle():
    # the CPU baseline's threaded N-d and real-row restatements give the
    # serial oracle's numbers (same chains, pencils split over threads)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((6, 10, 16)) + 1j * rng.standard_normal((6, 10, 16))
    ref = O.ndfft(x.ravel(), [16, 10, 6]).reshape(x.shape)
    assert rel_l2(O.baseline_ndfft(x.ravel(), [16, 10, 6], prec="f64", nthreads=3).reshape(x.shape), ref) < 1e-15
    xr = rng.standard_normal((5, 7, 32))
    X = O.baseline_rfft_batch(xr, prec="f64", nthreads=4)
    assert rel_l2(X, O.rfft(xr)) < 1e-15
    assert rel_l2(O.baseline_irfft_batch(X, 32, prec="f64", nthreads=4), xr) < 1e-15
    assert rel_l2(O.baseline_irfft_batch(X, 32, normalized=False, prec="f64", nthreads=2), 32 * xr) < 1e-15
    Y = O.baseline_ndfft(X.ravel(), [17, 7, 5], prec="f64", nthreads=4, first=1).reshape(X.shape)
    assert rel_l2(Y, O.ndrfft(xr)) < 1e-15


# ---------------------------------------------------------------- RNG spec

def test_splitmix64_kat(fj):
    # splitmix64 seeded with 0: published first outputs
    ref = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert [int(v) for v in O.splitmix64(0, 3)] == ref
    assert [int(v) for v in fj.rng.splitmix64(0, 3)] == ref
    assert np.array_equal(O.splitmix64(12345, 1000, first=77), fj.rng.splitmix64(12345, 1000, first=77))
    u = fj.rng.uniform(0, 100000)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01


# ------------------------------------------- the pins catch planted mistakes

def test_planted_mistakes_are_caught(golden):
    x, ref = golden["c2c_1000_x"], golden["c2c_1000_fwd"]
    good = O.fft(x)
    assert rel_l2(good, ref) < 1e-15
    # wrong sign of the exponent
    assert rel_l2(O.fft(x, sign=1), ref) > 0.1
    # a dropped term (one input sample lost)
    xd = x.copy()
    xd[17] = 0
    assert rel_l2(O.fft(xd), ref) > 1e-3
    # transposed four-step output (k1 and k2 roles swapped)
    n, N1, N2 = 1000, 40, 25
    y = O.fourstep(x, N1, N2)
    swapped = y.reshape(N1, N2).T.ravel()
    assert rel_l2(swapped, ref) > 0.1
    # a twiddle off by one index: shift the output by one bin
    assert rel_l2(np.roll(good, 1), ref) > 0.1
