// Register butterflies for the HIP kernels: small DFTs of length R on R
// complex values held in registers, natural order in and out.
//
// Sign convention (the reference's D, kernels.jl:3): DIR = -1 forward,
// +1 inverse; X_k = sum_m x_m exp(DIR * 2*pi*i * m*k / R).
//
//  * powers of two (2..64): Cooley-Tukey split R = R1*R2 down to radix 2/4,
//    with the inner twiddles exp(DIR*2*pi*i*e/R) as compile-time constants
//    (+-i and -1 become swaps/negations);
//  * composites 6, 10, 12, 14, 15 (prime-factor split) and 9 (3 x 3 with
//    inner twiddles);
//  * odd primes 3..13: the conjugate-pair form also used by the reference's
//    generic odd butterfly (kernels.jl:146-228): with t_j = x_j + x_{P-j},
//    d_j = x_j - x_{P-j},  u_k = x_0 + sum_j cos(2 pi jk/P) t_j,
//    s_k = sum_j sin(2 pi jk/P) d_j,  X_k = u_k + DIR*i*s_k,
//    X_{P-k} = u_k - DIR*i*s_k  -- (P-1)^2/2 real multiply-adds per plane.
#pragma once
#include <hip/hip_runtime.h>
#include "dft_consts.h"

namespace fftjh {

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };
template <typename T> using V2 = typename Vec2<T>::type;

template <typename T>
__device__ __forceinline__ V2<T> mk(T a, T b) {
  V2<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ V2<T> cadd(V2<T> a, V2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ V2<T> csub(V2<T> a, V2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ V2<T> cscale(V2<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// a * w
template <typename T>
__device__ __forceinline__ V2<T> cmul(V2<T> a, V2<T> w) {
  return mk<T>(fma(a.x, w.x, -(a.y * w.y)), fma(a.x, w.y, a.y * w.x));
}
// a * conj(w)
template <typename T>
__device__ __forceinline__ V2<T> cmulc(V2<T> a, V2<T> w) {
  return mk<T>(fma(a.x, w.x, a.y * w.y), fma(a.y, w.x, -(a.x * w.y)));
}
// a * (DIR * i)
template <typename T, int DIR>
__device__ __forceinline__ V2<T> mul_di(V2<T> a) {
  if constexpr (DIR < 0) return mk<T>(a.y, -a.x);
  else return mk<T>(-a.y, a.x);
}

// a * w_R^e, w_R = exp(DIR*2*pi*i/R); e is a constant after loop unrolling
template <typename T, int DIR, int R>
__device__ __forceinline__ V2<T> tw_const(V2<T> a, int e) {
  e %= R;
  if (e == 0) return a;
  if (4 * e == R) return mul_di<T, DIR>(a);
  if (2 * e == R) return mk<T>(-a.x, -a.y);
  if (4 * e == 3 * R) return mul_di<T, -DIR>(a);
  const T c = static_cast<T>(kCos64[e * (64 / R)]);
  const T s = static_cast<T>(DIR * kSin64[e * (64 / R)]);
  return cmul<T>(a, mk<T>(c, s));
}

template <typename T, int DIR, int R>
__device__ __forceinline__ void dft_pow2(V2<T>* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    const V2<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
  } else if constexpr (R == 4) {
    const V2<T> a0 = cadd<T>(v[0], v[2]), a1 = csub<T>(v[0], v[2]);
    const V2<T> b0 = cadd<T>(v[1], v[3]);
    const V2<T> b1 = mul_di<T, DIR>(csub<T>(v[1], v[3]));
    v[0] = cadd<T>(a0, b0);
    v[1] = cadd<T>(a1, b1);
    v[2] = csub<T>(a0, b0);
    v[3] = csub<T>(a1, b1);
  } else {
    // R = R1*R2: input n = R2*n1 + n2, output k = k1 + R1*k2
    constexpr int R1 = (R >= 16) ? 4 : 2;
    constexpr int R2 = R / R1;
    V2<T> t[R];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      V2<T> u[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) u[n1] = v[R2 * n1 + n2];
      dft_pow2<T, DIR, R1>(u);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) t[n2 * R1 + k1] = tw_const<T, DIR, R>(u[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      V2<T> u[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) u[n2] = t[n2 * R1 + k1];
      dft_pow2<T, DIR, R2>(u);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = u[k2];
    }
  }
}

template <int P>
__device__ __forceinline__ double odd_cos(int idx) {
  if constexpr (P == 3) return kCos3[idx];
  else if constexpr (P == 5) return kCos5[idx];
  else if constexpr (P == 7) return kCos7[idx];
  else if constexpr (P == 11) return kCos11[idx];
  else return kCos13[idx];
}
template <int P>
__device__ __forceinline__ double odd_sin(int idx) {
  if constexpr (P == 3) return kSin3[idx];
  else if constexpr (P == 5) return kSin5[idx];
  else if constexpr (P == 7) return kSin7[idx];
  else if constexpr (P == 11) return kSin11[idx];
  else return kSin13[idx];
}

template <typename T, int DIR, int P>
__device__ __forceinline__ void dft_odd(V2<T>* v) {
  constexpr int h = (P - 1) / 2;
  V2<T> t[h], d[h];
#pragma unroll
  for (int j = 1; j <= h; ++j) {
    t[j - 1] = cadd<T>(v[j], v[P - j]);
    d[j - 1] = csub<T>(v[j], v[P - j]);
  }
  V2<T> out[P];
  V2<T> y0 = v[0];
#pragma unroll
  for (int j = 0; j < h; ++j) y0 = cadd<T>(y0, t[j]);
  out[0] = y0;
#pragma unroll
  for (int k = 1; k <= h; ++k) {
    V2<T> u = v[0];
    V2<T> s = mk<T>(T(0), T(0));
#pragma unroll
    for (int j = 1; j <= h; ++j) {
      const T c = static_cast<T>(odd_cos<P>((k - 1) * h + (j - 1)));
      const T sn = static_cast<T>(odd_sin<P>((k - 1) * h + (j - 1)));
      u.x = fma(c, t[j - 1].x, u.x);
      u.y = fma(c, t[j - 1].y, u.y);
      s.x = fma(sn, d[j - 1].x, s.x);
      s.y = fma(sn, d[j - 1].y, s.y);
    }
    const V2<T> is = mul_di<T, DIR>(s);
    out[k] = cadd<T>(u, is);
    out[P - k] = csub<T>(u, is);
  }
#pragma unroll
  for (int k = 0; k < P; ++k) v[k] = out[k];
}

// Composite lengths R = R1*R2 with gcd(R1, R2) = 1 (6, 10, 12, 14, 15): the
// prime-factor (Good-Thomas) split, no inner twiddles.  Input n = (R2*n1 +
// R1*n2) mod R, output k = (k1*R2*(R2^-1 mod R1) + k2*R1*(R1^-1 mod R2)) mod R:
// then w_R^(nk) = w_R1^(n1 k1) * w_R2^(n2 k2), so R2 DFTs of length R1 and R1
// of length R2 with the two index maps as compile-time permutations.
template <typename T, int DIR, int R>
__device__ __forceinline__ void dft_any(V2<T>* v);

constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;
}

template <typename T, int DIR, int R1, int R2>
__device__ __forceinline__ void dft_pfa(V2<T>* v) {
  constexpr int R = R1 * R2;
  constexpr int s1 = inv_mod(R2 % R1, R1), s2 = inv_mod(R1 % R2, R2);
  V2<T> t[R];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    V2<T> u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = v[(R2 * n1 + R1 * n2) % R];
    dft_any<T, DIR, R1>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) t[k1 * R2 + n2] = u[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    V2<T> u[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) u[n2] = t[k1 * R2 + n2];
    dft_any<T, DIR, R2>(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[(k1 * R2 * s1 + k2 * R1 * s2) % R] = u[k2];
  }
}

// R = 9 = 3 x 3 (Cooley-Tukey): input n = 3*n1 + n2, output k = k1 + 3*k2,
// inner twiddles w_9^(n2 k1) from the 9th roots of unity
template <typename T, int DIR>
__device__ __forceinline__ void dft9(V2<T>* v) {
  V2<T> t[9];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
    V2<T> u[3] = {v[n2], v[3 + n2], v[6 + n2]};
    dft_odd<T, DIR, 3>(u);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
      const int e = n2 * k1;
      t[n2 * 3 + k1] = e == 0 ? u[k1]
                              : cmul<T>(u[k1], mk<T>(static_cast<T>(kCos9[e]), static_cast<T>(DIR * kSin9[e])));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    V2<T> u[3] = {t[k1], t[3 + k1], t[6 + k1]};
    dft_odd<T, DIR, 3>(u);
#pragma unroll
    for (int k2 = 0; k2 < 3; ++k2) v[k1 + 3 * k2] = u[k2];
  }
}

template <typename T, int DIR, int R>
__device__ __forceinline__ void dft_any(V2<T>* v) {
  if constexpr ((R & (R - 1)) == 0) dft_pow2<T, DIR, R>(v);
  else if constexpr (R == 6) dft_pfa<T, DIR, 2, 3>(v);
  else if constexpr (R == 9) dft9<T, DIR>(v);
  else if constexpr (R == 10) dft_pfa<T, DIR, 2, 5>(v);
  else if constexpr (R == 12) dft_pfa<T, DIR, 4, 3>(v);
  else if constexpr (R == 14) dft_pfa<T, DIR, 2, 7>(v);
  else if constexpr (R == 15) dft_pfa<T, DIR, 3, 5>(v);
  else dft_odd<T, DIR, R>(v);
}

template <typename T, int DIR, int R>
__device__ __forceinline__ void dft(V2<T>* v) {
  dft_any<T, DIR, R>(v);
}

}  // namespace fftjh
