"""Generate dft_consts.h: exact-to-double butterfly constants for the HIP kernels.

The odd-prime butterflies use the conjugate-pair form of a length-P DFT
(X_k = u_k + i v_k, X_{P-k} = u_k - i v_k with u_k = x_0 + sum_j cos(2 pi jk/P) t_j
and v_k = sum_j sin(2 pi jk/P) d_j, the form of kernels.jl:146-228), so they need
cos/sin(2 pi jk/P) for j, k = 1..(P-1)/2.  The power-of-two butterflies need the
64th roots of unity, the radix-9 butterfly (3 x 3 with inner twiddles) the 9th.  Values are computed with mpmath at 40 digits and printed
with 17 significant digits, which round-trips every double exactly.

Run:  python gen_consts.py > dft_consts.h
"""
import mpmath

mpmath.mp.dps = 40
PRIMES = (3, 5, 7, 11, 13)


def lit(v):
    s = mpmath.nstr(v, 17, min_fixed=-1, max_fixed=1)
    if "." not in s and "e" not in s:
        s += ".0"
    return s


def main():
    out = []
    out.append("// Generated by gen_consts.py -- do not edit.")
    out.append("#pragma once")
    out.append("namespace fftjh {")
    for p in PRIMES:
        h = (p - 1) // 2
        cs, ss = [], []
        for k in range(1, h + 1):
            for j in range(1, h + 1):
                a = 2 * mpmath.pi * j * k / p
                cs.append(lit(mpmath.cos(a)))
                ss.append(lit(mpmath.sin(a)))
        out.append(f"// cos/sin(2 pi j k / {p}) at [(k-1)*{h} + (j-1)], j,k = 1..{h}")
        out.append(f"__device__ constexpr double kCos{p}[{h*h}] = {{{', '.join(cs)}}};")
        out.append(f"__device__ constexpr double kSin{p}[{h*h}] = {{{', '.join(ss)}}};")
    cs = [lit(mpmath.cos(2 * mpmath.pi * e / 9)) for e in range(9)]
    ss = [lit(mpmath.sin(2 * mpmath.pi * e / 9)) for e in range(9)]
    out.append("// cos/sin(2 pi e / 9), e = 0..8 (the radix-9 butterfly's inner twiddles)")
    out.append(f"__device__ constexpr double kCos9[9] = {{{', '.join(cs)}}};")
    out.append(f"__device__ constexpr double kSin9[9] = {{{', '.join(ss)}}};")
    cs = [lit(mpmath.cos(2 * mpmath.pi * e / 64)) for e in range(64)]
    ss = [lit(mpmath.sin(2 * mpmath.pi * e / 64)) for e in range(64)]
    out.append("// cos/sin(2 pi e / 64), e = 0..63")
    out.append(f"__device__ constexpr double kCos64[64] = {{{', '.join(cs)}}};")
    out.append(f"__device__ constexpr double kSin64[64] = {{{', '.join(ss)}}};")
    out.append("}  // namespace fftjh")
    print("\n".join(out))


if __name__ == "__main__":
    main()
