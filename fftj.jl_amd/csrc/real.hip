// Real-transform edges around the half-length complex transform.
//
// Even n (real.jl:80-134, 149-193): the n reals are read as n/2 complex
// points (no copy), the half-length complex transform runs, and an O(n)
// twist separates the even/odd spectra:
//   R2C  y_0 = Re Z_0 + Im Z_0,  y_h = Re Z_0 - Im Z_0,
//        y_k = xe + w_k*(-i*xo),  xe,xo = (Z_k +- conj(Z_{h-k}))/2,  w_k = e^{-2 pi i k/n}
//   C2R  Z_k = xe + i*(w_k*xd),   xe,xd = (X_k +- conj(X_{h-k}))*pre,   w_k = e^{+2 pi i k/n}
//        (pre = 1/2 for irfft, 1 for brfft), then the inverse half transform.
// Odd n (real.jl:136-140, 194-209): full-length complex transform of the
// real data (R2C) or of the Hermitian-completed spectrum (C2R).
#include "butterflies.h"
#include "common.h"
#include "launch.h"

namespace fftjh {

static dim3 grid_1d(int64_t count) {
  int64_t g = (count + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  if (g < 1) g = 1;
  return dim3(static_cast<unsigned>(g));
}

#define GRID_STRIDE(i, count)                                                            \
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < (count); \
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)

template <typename T>
__global__ void k_r2c_twist(const V2<T>* Z, V2<T>* y, const V2<T>* __restrict__ w, int64_t h,
                            int64_t batch, int64_t zs, int64_t ys) {
  const int64_t per = h + 1;
  GRID_STRIDE(idx, batch * per) {
    const int64_t b = idx / per, k = idx - b * per;
    const V2<T>* z = Z + b * zs;
    V2<T> out;
    if (k == 0) {
      const V2<T> z0 = z[0];
      out = mk<T>(z0.x + z0.y, T(0));
    } else if (k == h) {
      const V2<T> z0 = z[0];
      out = mk<T>(z0.x - z0.y, T(0));
    } else {
      const V2<T> a = z[k];
      const V2<T> zc = z[h - k];
      const V2<T> c = mk<T>(zc.x, -zc.y);
      const T half = T(0.5);
      const V2<T> xe = cscale<T>(cadd<T>(a, c), half);
      const V2<T> xo = cscale<T>(csub<T>(a, c), half);
      const V2<T> mixo = mk<T>(xo.y, -xo.x);  // -i * xo
      out = cadd<T>(xe, cmul<T>(mixo, w[k]));
    }
    y[b * ys + k] = out;
  }
}

template <typename T>
__global__ void k_c2r_pre(const V2<T>* X, V2<T>* Z, const V2<T>* __restrict__ w, int64_t h,
                          int64_t batch, int64_t xs, int64_t zs, T pre) {
  GRID_STRIDE(idx, batch * h) {
    const int64_t b = idx / h, k = idx - b * h;
    const V2<T>* xb = X + b * xs;
    const V2<T> a = xb[k];
    const V2<T> xc = xb[h - k];
    const V2<T> c = mk<T>(xc.x, -xc.y);
    const V2<T> xe = cscale<T>(cadd<T>(a, c), pre);
    const V2<T> xd = cscale<T>(csub<T>(a, c), pre);
    const V2<T> t = cmul<T>(xd, w[k]);
    Z[b * zs + k] = mk<T>(xe.x - t.y, xe.y + t.x);  // xe + i*t
  }
}

// odd R2C: real -> complex with zero imaginary parts
template <typename T>
__global__ void k_real_to_complex(const T* x, V2<T>* z, int64_t n, int64_t batch) {
  GRID_STRIDE(idx, batch * n) { z[idx] = mk<T>(x[idx], T(0)); }
}

// first `keep` complex entries of each length-n row (odd R2C output)
template <typename T>
__global__ void k_copy_prefix(const V2<T>* z, V2<T>* y, int64_t n, int64_t keep, int64_t batch) {
  GRID_STRIDE(idx, batch * keep) {
    const int64_t b = idx / keep, k = idx - b * keep;
    y[idx] = z[b * n + k];
  }
}

// odd C2R: Hermitian completion of the n/2+1 input bins to n bins
template <typename T>
__global__ void k_hermitian_fill(const V2<T>* X, V2<T>* z, int64_t n, int64_t batch) {
  const int64_t hk = n / 2 + 1;
  GRID_STRIDE(idx, batch * n) {
    const int64_t b = idx / n, k = idx - b * n;
    const V2<T>* xb = X + b * hk;
    if (k < hk) {
      z[idx] = xb[k];
    } else {
      const V2<T> v = xb[n - k];
      z[idx] = mk<T>(v.x, -v.y);
    }
  }
}

template <typename T>
__global__ void k_take_real(const V2<T>* z, T* y, int64_t count) {
  GRID_STRIDE(idx, count) { y[idx] = z[idx].x; }
}

hipError_t launch_r2c_twist(int dtype, const void* Z, void* y, const void* w, int64_t h, int64_t batch,
                            int64_t zs, int64_t ys, hipStream_t st) {
  const dim3 g = grid_1d(batch * (h + 1));
  if (dtype == 0)
    hipLaunchKernelGGL(k_r2c_twist<float>, g, dim3(256), 0, st, static_cast<const float2*>(Z),
                       static_cast<float2*>(y), static_cast<const float2*>(w), h, batch, zs, ys);
  else
    hipLaunchKernelGGL(k_r2c_twist<double>, g, dim3(256), 0, st, static_cast<const double2*>(Z),
                       static_cast<double2*>(y), static_cast<const double2*>(w), h, batch, zs, ys);
  return hipGetLastError();
}

hipError_t launch_c2r_pre(int dtype, const void* X, void* Z, const void* w, int64_t h, int64_t batch,
                          int64_t xs, int64_t zs, double pre, hipStream_t st) {
  const dim3 g = grid_1d(batch * h);
  if (dtype == 0)
    hipLaunchKernelGGL(k_c2r_pre<float>, g, dim3(256), 0, st, static_cast<const float2*>(X),
                       static_cast<float2*>(Z), static_cast<const float2*>(w), h, batch, xs, zs,
                       static_cast<float>(pre));
  else
    hipLaunchKernelGGL(k_c2r_pre<double>, g, dim3(256), 0, st, static_cast<const double2*>(X),
                       static_cast<double2*>(Z), static_cast<const double2*>(w), h, batch, xs, zs, pre);
  return hipGetLastError();
}

hipError_t launch_real_to_complex(int dtype, const void* x, void* z, int64_t n, int64_t batch,
                                  hipStream_t st) {
  const dim3 g = grid_1d(batch * n);
  if (dtype == 0)
    hipLaunchKernelGGL(k_real_to_complex<float>, g, dim3(256), 0, st, static_cast<const float*>(x),
                       static_cast<float2*>(z), n, batch);
  else
    hipLaunchKernelGGL(k_real_to_complex<double>, g, dim3(256), 0, st, static_cast<const double*>(x),
                       static_cast<double2*>(z), n, batch);
  return hipGetLastError();
}

hipError_t launch_copy_prefix(int dtype, const void* z, void* y, int64_t n, int64_t keep, int64_t batch,
                              hipStream_t st) {
  const dim3 g = grid_1d(batch * keep);
  if (dtype == 0)
    hipLaunchKernelGGL(k_copy_prefix<float>, g, dim3(256), 0, st, static_cast<const float2*>(z),
                       static_cast<float2*>(y), n, keep, batch);
  else
    hipLaunchKernelGGL(k_copy_prefix<double>, g, dim3(256), 0, st, static_cast<const double2*>(z),
                       static_cast<double2*>(y), n, keep, batch);
  return hipGetLastError();
}

hipError_t launch_hermitian_fill(int dtype, const void* X, void* z, int64_t n, int64_t batch,
                                 hipStream_t st) {
  const dim3 g = grid_1d(batch * n);
  if (dtype == 0)
    hipLaunchKernelGGL(k_hermitian_fill<float>, g, dim3(256), 0, st, static_cast<const float2*>(X),
                       static_cast<float2*>(z), n, batch);
  else
    hipLaunchKernelGGL(k_hermitian_fill<double>, g, dim3(256), 0, st, static_cast<const double2*>(X),
                       static_cast<double2*>(z), n, batch);
  return hipGetLastError();
}

hipError_t launch_take_real(int dtype, const void* z, void* y, int64_t count, hipStream_t st) {
  const dim3 g = grid_1d(count);
  if (dtype == 0)
    hipLaunchKernelGGL(k_take_real<float>, g, dim3(256), 0, st, static_cast<const float2*>(z),
                       static_cast<float*>(y), count);
  else
    hipLaunchKernelGGL(k_take_real<double>, g, dim3(256), 0, st, static_cast<const double2*>(z),
                       static_cast<double*>(y), count);
  return hipGetLastError();
}

}  // namespace fftjh
