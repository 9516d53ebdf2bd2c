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
// The interpolation weights are evaluated per tap in FP64 from the exact
// integer remainder r = k*m mod n, instead of being read from a table.
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

// G (contiguous rows of m) -> y (transforms at geometry gy)
template <typename T>
__global__ void k_rs_interp(const V2<T>* G, V2<T>* y, int64_t n, int64_t m, int w, double beta,
                            int64_t batch, int64_t Bi, Geometry gy) {
  const int h = w / 2;
  GRID_STRIDE(idx, batch * n) {
    const int64_t b = idx / n, k = idx - b * n;
    const int64_t km = k * m;
    const int64_t q = km / n, r = km - q * n;
    const double delta = static_cast<double>(r) / static_cast<double>(n);
    int64_t l = q - h + 1;
    if (l < 0) l += m;
    const V2<T>* gb = G + b * m;
    V2<T> acc = mk<T>(T(0), T(0));
    for (int i = 0; i < w; ++i) {
      const T wt = static_cast<T>(es_kernel((delta + (h - 1 - i)) / h, beta));
      int64_t li = l + i;
      if (li >= m) li -= m;
      const V2<T> gv = gb[li];
      acc.x = fma(wt, gv.x, acc.x);
      acc.y = fma(wt, gv.y, acc.y);
    }
    y[geo_base(gy, Bi, b) + k * gy.stride] = acc;
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

hipError_t launch_rs_interp(int dtype, const void* G, void* y, int64_t n, int64_t m, int w, double beta,
                            int64_t batch, int64_t Bi, Geometry gy, hipStream_t st) {
  const dim3 gr = grid_1d(batch * n);
  if (dtype == 0)
    hipLaunchKernelGGL(k_rs_interp<float>, gr, dim3(256), 0, st, static_cast<const float2*>(G),
                       static_cast<float2*>(y), n, m, w, beta, batch, Bi, gy);
  else
    hipLaunchKernelGGL(k_rs_interp<double>, gr, dim3(256), 0, st, static_cast<const double2*>(G),
                       static_cast<double2*>(y), n, m, w, beta, batch, Bi, gy);
  return hipGetLastError();
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
