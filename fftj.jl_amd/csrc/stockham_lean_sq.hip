// Lean pass kernel for square lengths N = R*R (FP32): C adjacent columns of
// N points per workgroup of C*R threads, one group per workgroup.  It is
// the power-of-two lean kernel's structure for the passes of three-pass
// plans whose factors are 100 = 10 x 10 (10^6 = 100^3), with the column
// count chosen so that whole waves fill (C = 50: 500 threads, 400-byte
// runs) and C*R a multiple of N (the rows-out store gives each thread one
// element index).
//
// A column of N points, element m = m_lo + R*m_hi:
//   stage 1 (thread jt = m_lo, registers u = m_hi): Y[k1] = DFT_R over m_hi,
//     times w_N^(jt*k1) (stage table in LDS);
//   exchange through LDS (real parts, then imaginary parts);
//   stage 2 (thread jt, k1 = jt): DFT_R over m_lo, X[jt + R*k2] in v[k2].
// Edges:
//   ROWS = 0: outputs back to the column layout (possibly in place):
//     y[out_base + c + (jt + R*u)*ostride];
//   ROWS = 1: outputs transposed through LDS: thread t stores element
//     e = t mod N of columns i = t / N + (C*R/N)*v, v < R, at
//     y[out_base + i*ostride + e].
// Multi-pass twiddle (period P, general or a power of two): output k of
// column c of a group with digits (d0, d1, d2) times w_P^e,
//   e = d1*(a0*(C*d0 + c) + b*k) mod P,
// from two FP64 tables (w_P^(i*2^sh), w_P^i), one product rounded to FP32;
// columns out: a start value per lane times the group's step powers S^u
// (S = w_P^(d1*b*R), staged in LDS); rows out (a0 = 0): w_P^(d1*b*e), one
// value per thread.
#include "butterflies.h"
#include "common.h"
#include "launch.h"
#include "passcommon.h"

namespace fftjh {

namespace {

__device__ __forceinline__ uint32_t sq_group(uint32_t b, uint32_t groups, int xcd) {
  if (!xcd) return b;
  const uint32_t q = groups >> 3, r = groups & 7u, x = b & 7u, s = b >> 3;
  return x * q + (x < r ? x : r) + s;
}

template <int DIR, int R, int C, int ROWS>
__global__ void __launch_bounds__(C * R) k_lean_sq(LeanDesc d, const float2* __restrict__ x, float2* y,
                                                    const float2* __restrict__ tws) {
  constexpr int N = R * R, T = C * R;
  static_assert(T % N == 0, "rows-out store: one element index per thread");
  constexpr int CS = N + 1;  // column stride of the LDS exchange (odd: conflict-free in c)
  __shared__ float lf[C * CS];
  __shared__ float2 ltw[N];
  __shared__ float2 lsu[R];
  const int tid = static_cast<int>(threadIdx.x);
  const int c = tid % C, jt = tid / C;
  if (static_cast<int>(blockIdx.x) >= d.groups) return;
  const uint32_t g = sq_group(blockIdx.x, static_cast<uint32_t>(d.groups), d.xcd);
  const uint32_t d0 = g % static_cast<uint32_t>(d.D0);
  const uint32_t gr = g / static_cast<uint32_t>(d.D0);
  const uint32_t d1 = gr % static_cast<uint32_t>(d.D1);
  const uint32_t d2 = gr / static_cast<uint32_t>(d.D1);
  const int64_t ib = d0 * d.in_s0 + d1 * d.in_s1 + d2 * d.in_s2;
  const int64_t ob = d0 * d.out_s0 + d1 * d.out_s1 + d2 * d.out_s2;

  // w_P^e for e = d1*f mod P (f reduced first)
  const double2* ptw = static_cast<const double2*>(d.ptw);
  const uint64_t P = d.pmod ? static_cast<uint64_t>(d.pmod) : (uint64_t(1) << d.psh);
  const uint64_t lmask = (uint64_t(1) << d.ptw_sh) - 1;
  auto tw_at = [&](uint64_t f) -> double2 {
    const uint64_t e = mulmod(static_cast<uint64_t>(d1) % P, f % P, P, d.pinv);
    return cmul64(ptw[e >> d.ptw_sh], ptw[d.ptw_lo + static_cast<int64_t>(e & lmask)]);
  };
  double2 w0d = make_double2(1.0, 0.0);
  if (d.psh) {
    const uint64_t bb = static_cast<uint64_t>(d.tw_b);
    if constexpr (ROWS) {
      w0d = tw_at(bb * static_cast<uint64_t>(tid % N));
    } else {
      const uint64_t col = static_cast<uint64_t>(C) * d0 + static_cast<uint64_t>(c);
      w0d = tw_at(static_cast<uint64_t>(d.tw_a0) * col + bb * static_cast<uint64_t>(jt));
      if (tid < R) {
        const double2 su = tw_at(bb * static_cast<uint64_t>(R) * static_cast<uint64_t>(tid));
        lsu[tid] = make_float2(static_cast<float>(su.x), static_cast<float>(su.y));
      }
    }
  }
  float2 v[R];
  {
    const float2* xb = x + ib + c + static_cast<int64_t>(jt) * d.istride;
#pragma unroll
    for (int u = 0; u < R; ++u) v[u] = xb[static_cast<int64_t>(R) * u * d.istride];
  }
  for (int i = tid; i < N; i += T) ltw[i] = tws[i];
  const float scale = static_cast<float>(d.scale);
  const bool scaled = d.scale != 1.0;
  __syncthreads();
  // stage 1: DFT_R over m_hi, then w_N^(jt*k1); slot [c][k1*R + jt]
  dft<float, DIR, R>(v);
#pragma unroll
  for (int k1 = 1; k1 < R; ++k1) v[k1] = cmul<float>(v[k1], ltw[jt * k1]);
  float2 in[R];
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) lf[c * CS + k1 * R + jt] = v[k1].x;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < R; ++m) in[m].x = lf[c * CS + jt * R + m];
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) lf[c * CS + k1 * R + jt] = v[k1].y;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < R; ++m) in[m].y = lf[c * CS + jt * R + m];
  // stage 2: k1 = jt, DFT_R over m_lo: X[jt + R*k2] in in[k2]
  dft<float, DIR, R>(in);
  const float2 w0 = make_float2(static_cast<float>(w0d.x), static_cast<float>(w0d.y));
  if constexpr (ROWS) {
    // transpose: column c, element jt + R*k2 at [c][jt + R*k2]; thread t then
    // stores element e = t % N of columns t / N + (T / N)*v
    float2 o[R];
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) lf[c * CS + jt + R * k2] = in[k2].x;
    __syncthreads();
    const int e = tid % N, i0 = tid / N;
#pragma unroll
    for (int q = 0; q < R; ++q) o[q].x = lf[(i0 + (T / N) * q) * CS + e];
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) lf[c * CS + jt + R * k2] = in[k2].y;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) o[q].y = lf[(i0 + (T / N) * q) * CS + e];
    float2 w = w0;
    if (scaled) w = cscale<float>(w, scale);
    float2* yb = y + ob + e;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const float2 a = (d.psh || scaled) ? cmul<float>(o[q], w) : o[q];
      yb[static_cast<int64_t>(i0 + (T / N) * q) * d.ostride] = a;
    }
  } else {
    float2* yb = y + ob + c + static_cast<int64_t>(jt) * d.ostride;
    if (d.psh) {
      const float2 ws = scaled ? cscale<float>(w0, scale) : w0;
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const float2 val = cmul<float>(in[u], u ? cmul<float>(ws, lsu[u]) : ws);
        yb[static_cast<int64_t>(R) * u * d.ostride] = val;
      }
    } else {
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const float2 val = scaled ? cscale<float>(in[u], scale) : in[u];
        yb[static_cast<int64_t>(R) * u * d.ostride] = val;
      }
    }
  }
}

template <int DIR, int R, int C>
hipError_t launch_rc(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  if (d.rows)
    hipLaunchKernelGGL((k_lean_sq<DIR, R, C, 1>), dim3(d.groups), dim3(C * R), 0, st, d,
                       static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
  else
    hipLaunchKernelGGL((k_lean_sq<DIR, R, C, 0>), dim3(d.groups), dim3(C * R), 0, st, d,
                       static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
  return hipGetLastError();
}

template <int DIR>
hipError_t launch_sq_dir(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  if (d.n == 100 && d.cols == 50) return launch_rc<DIR, 10, 50>(d, x, y, tws, st);
  return hipErrorInvalidValue;
}

}  // namespace

int lean_sq_cols(int n) { return n == 100 ? 50 : 0; }

hipError_t launch_lean_sq(int dir, const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  return dir < 0 ? launch_sq_dir<-1>(d, x, y, tws, st) : launch_sq_dir<1>(d, x, y, tws, st);
}

}  // namespace fftjh
