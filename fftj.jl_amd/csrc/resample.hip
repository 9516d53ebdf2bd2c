// Prime/non-smooth lengths by resampling onto a smooth length m >= 2n
// (resample.jl:1-7: a type-2 NUFFT evaluated at the points k/n).
//
//   corr_j = scale / (w * Phi(j'/m)),  j' = j folded to [-floor(n/2), ceil(n/2)),
//            Phi(nu) = int psi(2u/w) e^{-2 pi i nu u} du by a Gauss-Legendre cosine sum
//            (resample.jl:226-236)                                  -> k_rs_corr, FP64
//   g      = x .* corr placed at the folded positions of an m-grid,
//            zero in the middle (resample.jl:267-272, 281-301)      -> k_rs_deconv
//   G      = DFT_m(g)                                               -> the pass kernels
//   y_k    = sum_{i<w} psi((delta + h - 1 - i)/h) G[(q - h + 1 + i) mod m],
//            k*m = q*n + r, delta = r/n, psi(z) = exp(beta(sqrt(1-z^2)-1))
//            (resample.jl:238-244, 392-424)                         -> k_rs_interp
// The interpolation weights come from piecewise polynomials in delta fitted
// at plan time (the planner, fit_rs_weights), evaluated in the plan's
// precision from the exact integer remainder r = k*m mod n.
#include <algorithm>
#include <type_traits>

#include "butterflies.h"
#include "common.h"
#include "launch.h"

namespace fftjh {

#define GRID_STRIDE(i, count)                                                            \
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < (count); \
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)

static dim3 grid_1d(int64_t count) {
  int64_t g = (count + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  if (g < 1) g = 1;
  return dim3(static_cast<unsigned>(g));
}

__device__ __forceinline__ double es_kernel(double z, double beta) {
  return fabs(z) >= 1.0 ? 0.0 : exp(beta * (sqrt(1.0 - z * z) - 1.0));
}

template <typename T>
__global__ void k_rs_corr(T* corr, int64_t n, int64_t m, int w, double beta, const double* zq,
                          const double* wq, int nq, double scale) {
  const int64_t hp = (n + 1) / 2;
  GRID_STRIDE(j, n) {
    const int64_t jp = j < hp ? j : j - n;
    const double nu = static_cast<double>(jp) / static_cast<double>(m);
    double phi = 0.0;
    for (int q = 0; q < nq; ++q) {
      const double z = zq[q];
      if (z > 0.0) phi += wq[q] * es_kernel(z, beta) * cospi(nu * w * z);
    }
    corr[j] = static_cast<T>(scale / (w * phi));
  }
}

// transform b of a strided batch starts at (b % Bi)*s_in + (b / Bi)*s_out
__device__ __forceinline__ int64_t geo_base(const Geometry& g, int64_t Bi, int64_t b) {
  return (b % Bi) * g.s_in + (b / Bi) * g.s_out;
}

// x (transforms at geometry gx) -> g (contiguous rows of m)
template <typename T>
__global__ void k_rs_deconv(const V2<T>* x, V2<T>* g, const T* __restrict__ corr, int64_t n, int64_t m,
                            int64_t batch, int64_t Bi, Geometry gx) {
  const int64_t hp = (n + 1) / 2;
  const int64_t hi0 = m - n + hp;
  GRID_STRIDE(idx, batch * m) {
    const int64_t b = idx / m, j = idx - b * m;
    int64_t src = -1;
    if (j < hp) src = j;
    else if (j >= hi0) src = j - (m - n);
    V2<T> v = mk<T>(T(0), T(0));
    if (src >= 0) v = cscale<T>(x[geo_base(gx, Bi, b) + src * gx.stride], corr[src]);
    g[idx] = v;
  }
}

// G (contiguous rows of m) -> y (transforms at geometry gy).  A wave takes
// a span of 64*run consecutive outputs of one transform (run <= kInterpRun,
// chosen so small problems still fill the device), lane l the outputs
// l + 64u: the quotient and remainder of k*m by n are formed once per lane
// and then stepped by those of 64*m (stores stay 64-wide runs).
// The w weights of an output come from the piecewise Chebyshev polynomials
// the planner fitted (coef[((p*w + i)*(degree+1)) + k]), by Clenshaw's
// recurrence in T.
constexpr int kInterpRun = 8;

// weight of tap i from its Chebyshev coefficients a[0..deg] at t
template <typename T, int DEG>
__device__ __forceinline__ T clenshaw(const T* a, T t, int deg) {
  T b1 = T(0), b2 = T(0);
  if constexpr (DEG > 0) {
#pragma unroll
    for (int c = DEG; c >= 1; --c) {
      const T b0 = fma(T(2) * t, b1, a[c] - b2);
      b2 = b1;
      b1 = b0;
    }
  } else {
    for (int c = deg; c >= 1; --c) {
      const T b0 = fma(T(2) * t, b1, a[c] - b2);
      b2 = b1;
      b1 = b0;
    }
  }
  return fma(t, b1, a[0] - b2);
}

// W, DEG > 0: compile-time width and degree (the taps and the recurrence
// unrolled, a piece's coefficients read as one contiguous block); 0: runtime
// LDS_COEF: the coefficient table (pieces*W*(DEG+1) values) is staged in LDS
template <typename T, int W, int DEG, bool LDS_COEF>
__global__ void k_rs_interp(const V2<T>* G, V2<T>* y, const T* __restrict__ coef_g, int pieces, int degree,
                            int64_t n, int64_t m, int w_rt, int64_t batch, int64_t Bi, Geometry gy, int run) {
  extern __shared__ __align__(16) unsigned char smem[];
  const T* coef = coef_g;
  // degree 3 staged in LDS: each tap's Chebyshev coefficients a0..a3 become
  // the monomial ones of the same cubic, a0 - a2 + (a1 - 3a3) t + 2a2 t^2 +
  // 4a3 t^3 (coefficient growth at most 4), evaluated by Horner (3 FMAs a tap)
  constexpr bool kHorner = LDS_COEF && DEG == 3;
  // In LDS a piece's W*(DEG+1) values sit at a stride padded by 16 bytes:
  // lanes read the pieces of their own remainders, effectively at random,
  // and an unpadded power-of-two stride put every lane's 16-byte read on the
  // same few banks.
  constexpr int kPS = LDS_COEF ? W * (DEG + 1) + 16 / static_cast<int>(sizeof(T)) : W * (DEG + 1);
  if constexpr (LDS_COEF) {
    T* c = reinterpret_cast<T*>(smem);
    constexpr int kPer = W * (DEG + 1);
    const int cnt = pieces * kPer;
    if constexpr (kHorner) {
      for (int e = static_cast<int>(threadIdx.x); e < cnt / 4; e += static_cast<int>(blockDim.x)) {
        const T a0 = coef_g[4 * e], a1 = coef_g[4 * e + 1], a2 = coef_g[4 * e + 2], a3 = coef_g[4 * e + 3];
        T* d = c + (4 * e / kPer) * kPS + (4 * e) % kPer;
        d[0] = a0 - a2;
        d[1] = a1 - T(3) * a3;
        d[2] = T(2) * a2;
        d[3] = T(4) * a3;
      }
    } else {
      for (int e = static_cast<int>(threadIdx.x); e < cnt; e += static_cast<int>(blockDim.x))
        c[(e / kPer) * kPS + e % kPer] = coef_g[e];
    }
    __syncthreads();
    coef = c;
  }
  const int w = W > 0 ? W : w_rt;
  const int h = w / 2;
  const int N = (DEG > 0 ? DEG : degree) + 1;
  const double pn = static_cast<double>(pieces) / static_cast<double>(n);
  const int64_t span = 64 * static_cast<int64_t>(run);
  const int64_t runs = (n + span - 1) / span;
  const int64_t q64 = (64 * m) / n, r64 = 64 * m - q64 * n;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int64_t wave = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) >> 6;
  const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
  for (int64_t sp = wave; sp < batch * runs; sp += nwaves) {
    // (sp, hence b, is the same in every lane of the wave)
    const int64_t bl = sp / runs;
    const int64_t b = static_cast<int64_t>(
        (static_cast<uint64_t>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(bl >> 32))))
         << 32) |
        static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(bl & 0xffffffff))));
    int64_t k = (sp - b * runs) * span + lane;
    if (k >= n) continue;
    const int64_t km = k * m;
    int64_t q = km / n, r = km - q * n;
    const V2<T>* gb = G + b * m;
    V2<T>* yb = y + geo_base(gy, Bi, b);
    for (int u = 0; u < run && k < n; ++u) {
      const double sd = static_cast<double>(r) * pn;  // delta * pieces, in [0, pieces)
      int p = static_cast<int>(sd);
      if (p >= pieces) p = pieces - 1;
      const T t = static_cast<T>(2.0 * (sd - p) - 1.0);
      const T* cp = coef + static_cast<int64_t>(p) * (W > 0 && DEG > 0 ? kPS : w * N);
      int64_t l = q - h + 1;
      if (l < 0) l += m;
      V2<T> acc = mk<T>(T(0), T(0));
      if constexpr (W > 0 && DEG > 0) {
        T a[W * (DEG + 1)];
#pragma unroll
        for (int e = 0; e < W * (DEG + 1); ++e) a[e] = cp[e];
        T wt[W];
#pragma unroll
        for (int i = 0; i < W; ++i) {
          if constexpr (kHorner) {
            const T* ai = a + i * 4;
            wt[i] = fma(fma(fma(ai[3], t, ai[2]), t, ai[1]), t, ai[0]);
          } else {
            wt[i] = clenshaw<T, DEG>(a + i * (DEG + 1), t, DEG);
          }
        }
        // taps l + i mod m in 32 bits (m < 2^31) from the row's base
        V2<T> gv[W];
        const unsigned l32 = static_cast<unsigned>(l), m32 = static_cast<unsigned>(m);
#pragma unroll
        for (int i = 0; i < W; ++i) {
          unsigned li = l32 + static_cast<unsigned>(i);
          li = li >= m32 ? li - m32 : li;
          gv[i] = gb[li];
        }
#pragma unroll
        for (int i = 0; i < W; ++i) {
          acc.x = fma(wt[i], gv[i].x, acc.x);
          acc.y = fma(wt[i], gv[i].y, acc.y);
        }
      } else {
        for (int i = 0; i < w; ++i) {
          const T wt = clenshaw<T, 0>(cp + i * N, t, N - 1);
          int64_t li = l + i;
          if (li >= m) li -= m;
          const V2<T> gv = gb[li];
          acc.x = fma(wt, gv.x, acc.x);
          acc.y = fma(wt, gv.y, acc.y);
        }
      }
      yb[k * gy.stride] = acc;
      k += 64;
      r += r64;
      q += q64;
      if (r >= n) {
        r -= n;
        ++q;
      }
    }
  }
}

// dst[transform b, element i] = src[transform b, element i] between two
// strided geometries (gather a strided axis into rows and back)
template <typename T>
__global__ void k_strided_copy(const V2<T>* src, V2<T>* dst, int64_t n, int64_t batch, int64_t Bi, Geometry gs,
                               Geometry gd) {
  GRID_STRIDE(idx, batch * n) {
    const int64_t b = idx / n, i = idx - b * n;
    dst[geo_base(gd, Bi, b) + i * gd.stride] = src[geo_base(gs, Bi, b) + i * gs.stride];
  }
}

hipError_t launch_rs_corr(int dtype, void* corr, int64_t n, int64_t m, int w, double beta, const double* zq,
                          const double* wq, int nq, double scale, hipStream_t st) {
  const dim3 g = grid_1d(n);
  if (dtype == 0)
    hipLaunchKernelGGL(k_rs_corr<float>, g, dim3(256), 0, st, static_cast<float*>(corr), n, m, w, beta, zq,
                       wq, nq, scale);
  else
    hipLaunchKernelGGL(k_rs_corr<double>, g, dim3(256), 0, st, static_cast<double*>(corr), n, m, w, beta,
                       zq, wq, nq, scale);
  return hipGetLastError();
}

hipError_t launch_rs_deconv(int dtype, const void* x, void* g, const void* corr, int64_t n, int64_t m,
                            int64_t batch, int64_t Bi, Geometry gx, hipStream_t st) {
  const dim3 gr = grid_1d(batch * m);
  if (dtype == 0)
    hipLaunchKernelGGL(k_rs_deconv<float>, gr, dim3(256), 0, st, static_cast<const float2*>(x),
                       static_cast<float2*>(g), static_cast<const float*>(corr), n, m, batch, Bi, gx);
  else
    hipLaunchKernelGGL(k_rs_deconv<double>, gr, dim3(256), 0, st, static_cast<const double2*>(x),
                       static_cast<double2*>(g), static_cast<const double*>(corr), n, m, batch, Bi, gx);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_interp_t(const void* G, void* y, const void* coef, int pieces, int degree, int64_t n,
                                  int64_t m, int w, int64_t batch, int64_t Bi, Geometry gy, hipStream_t st) {
  // outputs per lane: up to kInterpRun, fewer when that would leave < 8192 waves
  const int64_t outs = batch * n;
  int run = static_cast<int>(std::min<int64_t>(kInterpRun, std::max<int64_t>(1, outs / (64 * 8192))));
  const dim3 gr = grid_1d(batch * ((n + 64 * run - 1) / (64 * run)) * 64);
  const auto* g = static_cast<const V2<T>*>(G);
  auto* yy = static_cast<V2<T>*>(y);
  const auto* c = static_cast<const T*>(coef);
  // the planner's usual fits: FP32 w = 8 degree 3 to 5 (table staged in
  // LDS, its piece stride W*(DEG+1) values plus 16 bytes), FP64 w = 16 degree 5
  auto w8 = [&](auto deg) {
    constexpr int D = decltype(deg)::value;
    const size_t lds = static_cast<size_t>(pieces) * (8 * (D + 1) * sizeof(T) + 16);
    if (lds <= 48 * 1024)
      hipLaunchKernelGGL((k_rs_interp<T, 8, D, true>), gr, dim3(256), lds, st, g, yy, c, pieces, degree, n, m, w, batch,
                         Bi, gy, run);
    else
      hipLaunchKernelGGL((k_rs_interp<T, 8, D, false>), gr, dim3(256), 0, st, g, yy, c, pieces, degree, n, m, w, batch,
                         Bi, gy, run);
  };
  if (w == 8 && degree == 3) {
    w8(std::integral_constant<int, 3>());
  } else if (w == 8 && degree == 4) {
    w8(std::integral_constant<int, 4>());
  } else if (w == 8 && degree == 5) {
    w8(std::integral_constant<int, 5>());
  } else if (w == 16 && degree == 5) {
    hipLaunchKernelGGL((k_rs_interp<T, 16, 5, false>), gr, dim3(256), 0, st, g, yy, c, pieces, degree, n, m, w, batch,
                       Bi, gy, run);
  } else {
    hipLaunchKernelGGL((k_rs_interp<T, 0, 0, false>), gr, dim3(256), 0, st, g, yy, c, pieces, degree, n, m, w, batch,
                       Bi, gy, run);
  }
  return hipGetLastError();
}

hipError_t launch_rs_interp(int dtype, const void* G, void* y, const void* coef, int pieces, int degree, int64_t n,
                            int64_t m, int w, int64_t batch, int64_t Bi, Geometry gy, hipStream_t st) {
  return dtype == 0 ? launch_interp_t<float>(G, y, coef, pieces, degree, n, m, w, batch, Bi, gy, st)
                    : launch_interp_t<double>(G, y, coef, pieces, degree, n, m, w, batch, Bi, gy, st);
}

hipError_t launch_strided_copy(int dtype, const void* src, void* dst, int64_t n, int64_t batch, int64_t Bi,
                               Geometry gs, Geometry gd, hipStream_t st) {
  const dim3 gr = grid_1d(batch * n);
  if (dtype == 0)
    hipLaunchKernelGGL(k_strided_copy<float>, gr, dim3(256), 0, st, static_cast<const float2*>(src),
                       static_cast<float2*>(dst), n, batch, Bi, gs, gd);
  else
    hipLaunchKernelGGL(k_strided_copy<double>, gr, dim3(256), 0, st, static_cast<const double2*>(src),
                       static_cast<double2*>(dst), n, batch, Bi, gs, gd);
  return hipGetLastError();
}

}  // namespace fftjh
