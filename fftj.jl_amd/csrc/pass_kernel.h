// The pass kernel template shared by the plain chain (stockham.hip) and the
// split-precision MFMA chain (stockham_mat.hip).
#pragma once
#include "butterflies.h"
#include "common.h"
#include "passcommon.h"
#include "stages.h"
#include "stages_mat.h"

namespace fftjh {

// First stage fused with the HBM load (ns = 1, no twiddles): the thread's
// butterflies j = jt + b*jstep read their inputs j + r*n/R straight from HBM
// into registers and scatter the outputs j*R + r to LDS.  Saves the load
// phase's LDS round trip and barrier.
template <typename T, int DIR, int R>
__device__ __noinline__ void first_stage(const int n_, const int C_, TwPtr<T> xb, const int64_t ist_, const bool valid,
                                         LdsPtr<T> lds_, const int c, const int jt, const int jstep_) {
  constexpr int BPT = kElemsPerThread / R;
  // (n, C, ist, lds and jstep are uniform: scalar registers)
  const int n = uni(n_), C = uni(C_), jstep = uni(jstep_);
  const int64_t ist = uni64(ist_);
  const LdsPtr<T> lds = uni_lds<T>(lds_);
  const int nr = n / R;
  const int64_t rstep = static_cast<int64_t>(nr) * ist;  // between a butterfly's inputs
  V2<T> v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    const bool on = valid && j < nr;
    TwPtr<T> p = xb + static_cast<int64_t>(j) * ist;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[b][r] = on ? ldv<T>(p, 0) : mk<T>(T(0), T(0));
      p += rstep;
    }
  }
  // outputs j*R + r of column c: flat slots (j*R)*C + c + r*C, affine in r
  // when C is a multiple of 16
  const bool lin = (C & 15) == 0;
  const int pc = uni(C + (C >> 4));
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    if (j < nr) {
      dft<T, DIR, R>(v[b]);
      const int f = static_cast<int>(__umul24(static_cast<unsigned>(j * R), static_cast<unsigned>(C))) + c;
      if (lin) {
        const LdsPtr<T> q = lds + lpad(f);
#pragma unroll
        for (int r = 0; r < R; ++r) stv(q, r * pc, v[b][r]);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) stv(lds, lpad(f + r * C), v[b][r]);
      }
    }
  }
  __syncthreads();
}

// Last stage fused with the HBM store (ns = n/R): butterfly j's outputs are
// the transform's outputs j + r*ns, written from registers with the
// multi-pass twiddle (started at j, stepped by ns) and the scale.  Saves the
// store phase's LDS round trip.
// The multi-pass twiddle of outputs j + r*ns, j = jt + b*jstep, comes in as
// the caller's FP64 start w0 = w(jt) and the steps over r (ns) and over b
// (jstep), each as separate arguments (in registers; a struct would go
// through the stack): FP64 recurrences here, one rounding per output.
template <typename T, int DIR, int R>
__device__ __noinline__ void last_stage(const int n_, const int C_, const StageDesc s, LdsPtr<T> lds_, TwPtr<T> tw_,
                                        GPtr<T> yb, const int64_t ost_, const bool valid, const T scale,
                                        const bool scaled, const bool twon, const double2 w0, const double2 Sr,
                                        const double2 Sb, const int c, const int jt, const int jstep_) {
  constexpr int BPT = kElemsPerThread / R;
  const int n = uni(n_), C = uni(C_), jstep = uni(jstep_), ns = uni(s.ns);
  const int64_t ost = uni64(ost_);
  const LdsPtr<T> lds = uni_lds<T>(lds_);
  using TwBytes = const __attribute__((address_space(1))) char*;
  const TwBytes twb = reinterpret_cast<TwBytes>(uni_ptr(tw_) + uni(s.tw_off));
  const int nr = n / R;
  const int64_t ostep = static_cast<int64_t>(ns) * ost;  // between a butterfly's outputs
  // inputs j + r*nr of column c: flat slots j*C + c + r*(nr*C)
  const int dr = nr * C;
  const bool lin = (dr & 15) == 0;
  const int pr = uni(dr + (dr >> 4));
  V2<T> v[BPT][R];
  V2<T> w[BPT][R - 1];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    if (j < nr) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const unsigned off = static_cast<unsigned>(j + (r - 1) * ns) * static_cast<unsigned>(sizeof(NV2<T>));
        const NV2<T> e = *reinterpret_cast<TwPtr<T>>(twb + off);
        w[b][r - 1] = mk<T>(e.x, e.y);
      }
      const int f = static_cast<int>(__umul24(static_cast<unsigned>(j), static_cast<unsigned>(C))) + c;
      if (lin) {
        const LdsPtr<T> q = lds + lpad(f);
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = ldv<T>(q, r * pr);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = ldv<T>(lds, lpad(f + r * dr));
      }
    }
  }
  if (!valid) return;
  if (twon) {
    double2 wb = w0;
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = jt + b * jstep;
      if (j < nr) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul<T>(v[b][r], w[b][r - 1]);
        dft<T, DIR, R>(v[b]);
        double2 wr = wb;
        GPtr<T> p = yb + static_cast<int64_t>(j) * ost;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          V2<T> val = cmul<T>(v[b][r], mk<T>(static_cast<T>(wr.x), static_cast<T>(wr.y)));
          if (scaled) val = cscale<T>(val, scale);
          stv(p, 0, val);
          p += ostep;
          wr = cmul64(wr, Sr);
        }
      }
      wb = cmul64(wb, Sb);
    }
  } else {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = jt + b * jstep;
      if (j < nr) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul<T>(v[b][r], w[b][r - 1]);
        dft<T, DIR, R>(v[b]);
        GPtr<T> p = yb + static_cast<int64_t>(j) * ost;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          stv(p, 0, scaled ? cscale<T>(v[b][r], scale) : v[b][r]);
          p += ostep;
        }
      }
    }
  }
}

template <typename T> struct MaxThreads { static constexpr int value = kMaxThreads; };
template <> struct MaxThreads<double> { static constexpr int value = kMaxThreads / 2; };

// The descriptor sits in device memory (written once at plan time): the
// stage loop indexes it with a uniform, dynamic stage number, which turns
// into scalar loads instead of a private copy of a by-value argument.
// MAT (FP32 only): kMatNone, or the split mode whose MFMA stages take the
// radices 9, 16 and 32 (operand table `mat`); the other radices run on VALU.
// (the MFMA chains keep twice the registers per lane: at most 512 threads)
constexpr int kMatMaxThreads = kMaxThreads / 2;
#ifndef FFTJH_MAT_WAVES
#define FFTJH_MAT_WAVES 1
#endif
template <typename T, int DIR, int MAT = kMatNone>
__global__ void __launch_bounds__(MAT != kMatNone ? kMatMaxThreads : MaxThreads<T>::value)
__attribute__((amdgpu_waves_per_eu(MAT != kMatNone ? FFTJH_MAT_WAVES : 1)))
k_pass(const PassDesc* __restrict__ dp, const V2<T>* x, V2<T>* y, const V2<T>* tw, const void* mat) {
  extern __shared__ __align__(16) unsigned char smem[];
  V2<T>* lds = reinterpret_cast<V2<T>*>(smem);
  const PassDesc& d = *dp;
  const int n = d.n, C = d.C;
  const int tid = static_cast<int>(threadIdx.x);
  const int c = tid & (C - 1);
  const int jt = tid >> d.logC;
  const int jstep = d.nthreads >> d.logC;
  const int64_t ist = d.in.stride, ost = d.out.stride;
  const GroupWalk walk(d);
  const T scale = static_cast<T>(d.scale);
  const bool scaled = d.scale != 1.0;

  for (int g = walk.start; g < walk.end; g += walk.step) {
    const int64_t b = static_cast<int64_t>(g) * C + c;
    const bool valid = b < d.batch;
    const BatchIndex ix(d, b);
    // load: rows jt + u*jstep of this thread's column into LDS
    {
      const V2<T>* xb = x + ix.base(d.in);
      for (int i0 = jt; i0 < n; i0 += kIoUnroll * jstep) {
        V2<T> t[kIoUnroll];
#pragma unroll
        for (int u = 0; u < kIoUnroll; ++u) {
          const int i = i0 + u * jstep;
          t[u] = (valid && i < n) ? xb[i * ist] : mk<T>(T(0), T(0));
        }
#pragma unroll
        for (int u = 0; u < kIoUnroll; ++u) {
          const int i = i0 + u * jstep;
          if (i < n) lds[lpad(i * C + c)] = t[u];
        }
      }
    }
    __syncthreads();
    for (int s = 0; s < d.nstages; ++s) {
      const StageDesc& st = d.st[s];
      if constexpr (MAT != kMatNone) {
        if (st.radix == 9 || st.radix == 16 || st.radix == 32) {
          const int lane = tid & 63, wave = tid >> 6, nwaves = d.nthreads >> 6;
          if (st.radix == 9) run_stage_mat<DIR, MAT, 9>(n, C, d.logC, st, as_lds<T>(lds), as_tw<T>(tw), mat, lane, wave, nwaves);
          else if (st.radix == 16) run_stage_mat<DIR, MAT, 16>(n, C, d.logC, st, as_lds<T>(lds), as_tw<T>(tw), mat, lane, wave, nwaves);
          else run_stage_mat<DIR, MAT, 32>(n, C, d.logC, st, as_lds<T>(lds), as_tw<T>(tw), mat, lane, wave, nwaves);
          continue;
        }
      }
      switch (st.radix) {
        case 2: run_stage<T, DIR, 2>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 3: run_stage<T, DIR, 3>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 4: run_stage<T, DIR, 4>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 5: run_stage<T, DIR, 5>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 6: run_stage<T, DIR, 6>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 9: run_stage<T, DIR, 9>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 10: run_stage<T, DIR, 10>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 12: run_stage<T, DIR, 12>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 14: run_stage<T, DIR, 14>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 15: run_stage<T, DIR, 15>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 7: run_stage<T, DIR, 7>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 8: run_stage<T, DIR, 8>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 11: run_stage<T, DIR, 11>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 13: run_stage<T, DIR, 13>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        default: run_stage<T, DIR, 16>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
      }
    }
    // store, with the multi-pass twiddle and the scale
    if (valid) {
      V2<T>* yb = y + ix.base(d.out);
      if (d.tw_period) {
        PostTwiddle pt;
        pt.init(d, ix, jt, jstep);
        for (int k = jt; k < n; k += jstep) {
          V2<T> v = cmul<T>(lds[lpad(k * C + c)], pt.value<T>());
          if (scaled) v = cscale<T>(v, scale);
          yb[k * ost] = v;
          pt.advance();
        }
      } else {
#pragma unroll 4
        for (int k = jt; k < n; k += jstep) {
          V2<T> v = lds[lpad(k * C + c)];
          if (scaled) v = cscale<T>(v, scale);
          yb[k * ost] = v;
        }
      }
    }
    __syncthreads();  // the store phase's LDS reads are done before the next load
  }
}


// The same contract for chains of two or more VALU stages, with the first
// stage fused into the HBM load and the last into the store (two LDS round
// trips and barriers fewer per group).  A kernel of its own: the stage
// calls then hold little live state, and the plain kernel keeps its
// register allocation.
template <typename T, int DIR>
__global__ void __launch_bounds__(MaxThreads<T>::value)
k_pass_fused(const PassDesc* __restrict__ dp, const V2<T>* x, V2<T>* y, const V2<T>* tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  V2<T>* lds = reinterpret_cast<V2<T>*>(smem);
  const PassDesc& d = *dp;
  const int n = d.n, C = d.C;
  const int tid = static_cast<int>(threadIdx.x);
  const int c = tid & (C - 1);
  const int jt = tid >> d.logC;
  const int jstep = d.nthreads >> d.logC;
  const int64_t ist = d.in.stride, ost = d.out.stride;
  const GroupWalk walk(d);
  const T scale = static_cast<T>(d.scale);
  const bool scaled = d.scale != 1.0;
  const int S = d.nstages;
  for (int g = walk.start; g < walk.end; g += walk.step) {
    const int64_t b = static_cast<int64_t>(g) * C + c;
    const bool valid = b < d.batch;
    {
      const TwPtr<T> xb = as_tw<T>(x + BatchIndex(d, b).base(d.in));
#define FFTJH_FIRST(R_) first_stage<T, DIR, R_>(n, C, xb, ist, valid, as_lds<T>(lds), c, jt, jstep)
      switch (d.st[0].radix) {
        case 2: FFTJH_FIRST(2); break;
        case 3: FFTJH_FIRST(3); break;
        case 4: FFTJH_FIRST(4); break;
        case 5: FFTJH_FIRST(5); break;
        case 6: FFTJH_FIRST(6); break;
        case 9: FFTJH_FIRST(9); break;
        case 10: FFTJH_FIRST(10); break;
        case 12: FFTJH_FIRST(12); break;
        case 14: FFTJH_FIRST(14); break;
        case 15: FFTJH_FIRST(15); break;
        case 7: FFTJH_FIRST(7); break;
        case 8: FFTJH_FIRST(8); break;
        case 11: FFTJH_FIRST(11); break;
        case 13: FFTJH_FIRST(13); break;
        default: FFTJH_FIRST(16); break;
      }
#undef FFTJH_FIRST
    }
    for (int s = 1; s < S - 1; ++s) {
      const StageDesc& st = d.st[s];
      switch (st.radix) {
        case 2: run_stage<T, DIR, 2>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 3: run_stage<T, DIR, 3>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 4: run_stage<T, DIR, 4>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 5: run_stage<T, DIR, 5>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 6: run_stage<T, DIR, 6>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 9: run_stage<T, DIR, 9>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 10: run_stage<T, DIR, 10>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 12: run_stage<T, DIR, 12>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 14: run_stage<T, DIR, 14>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 15: run_stage<T, DIR, 15>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 7: run_stage<T, DIR, 7>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 8: run_stage<T, DIR, 8>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 11: run_stage<T, DIR, 11>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        case 13: run_stage<T, DIR, 13>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
        default: run_stage<T, DIR, 16>(n, C, st, as_lds<T>(lds), as_tw<T>(tw), c, jt, jstep); break;
      }
    }
    {
      // (the batch index is recomputed here rather than kept live across
      // the stage calls)
      const BatchIndex ix(d, b);
      const StageDesc& sl = d.st[S - 1];
      const GPtr<T> yb = as_g<T>(y + ix.base(d.out));
      const bool twon = d.tw_period != 0;
      double2 w0 = make_double2(1.0, 0.0), Sr = w0, Sb = w0;
      if (twon) {
        PostTwiddle pt;
        pt.init(d, ix, jt, sl.ns);
        w0 = pt.w;
        Sr = pt.S;
        Sb = pt.step_for(d, jstep);
      }
#define FFTJH_LAST(R_)                                                                                         \
  last_stage<T, DIR, R_>(n, C, sl, as_lds<T>(lds), as_tw<T>(tw), yb, ost, valid, scale, scaled, twon, w0, Sr, Sb, \
                         c, jt, jstep)
      switch (sl.radix) {
        case 2: FFTJH_LAST(2); break;
        case 3: FFTJH_LAST(3); break;
        case 4: FFTJH_LAST(4); break;
        case 5: FFTJH_LAST(5); break;
        case 6: FFTJH_LAST(6); break;
        case 9: FFTJH_LAST(9); break;
        case 10: FFTJH_LAST(10); break;
        case 12: FFTJH_LAST(12); break;
        case 14: FFTJH_LAST(14); break;
        case 15: FFTJH_LAST(15); break;
        case 7: FFTJH_LAST(7); break;
        case 8: FFTJH_LAST(8); break;
        case 11: FFTJH_LAST(11); break;
        case 13: FFTJH_LAST(13); break;
        default: FFTJH_LAST(16); break;
      }
#undef FFTJH_LAST
    }
    __syncthreads();  // the last stage's LDS reads are done before the next group's first stage
  }
}

template <typename T, int DIR>
static hipError_t launch_pass_fused_t(const PassDesc& d, const PassDesc* dev_desc, const void* x, void* y,
                                      const void* tw, hipStream_t stream) {
  static bool attr_set = false;
  if (!attr_set) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass_fused<T, DIR>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    attr_set = true;
  }
  const size_t lds = static_cast<size_t>(lds_elems(d.C * d.n)) * sizeof(V2<T>);
  const int grid = d.grid > 0 ? d.grid : d.ngroups;
  hipLaunchKernelGGL((k_pass_fused<T, DIR>), dim3(grid), dim3(d.nthreads), lds, stream, dev_desc,
                     static_cast<const V2<T>*>(x), static_cast<V2<T>*>(y), static_cast<const V2<T>*>(tw));
  return hipGetLastError();
}

template <typename T, int DIR>
static int resident_groups_fused_t(int nthreads, size_t lds) {
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass_fused<T, DIR>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_pass_fused<T, DIR>),
                                                   nthreads, lds) != hipSuccess)
    return 0;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return per_cu * cus;
}

template <typename T, int DIR, int MAT>
static hipError_t launch_pass_t(const PassDesc& d, const PassDesc* dev_desc, const void* x, void* y,
                                const void* tw, const void* mat, hipStream_t stream) {
  static bool attr_set = false;
  if (!attr_set) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<T, DIR, MAT>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    attr_set = true;
  }
  const size_t lds = static_cast<size_t>(lds_elems(d.C * d.n)) * sizeof(V2<T>);
  const int grid = d.grid > 0 ? d.grid : d.ngroups;
  hipLaunchKernelGGL((k_pass<T, DIR, MAT>), dim3(grid), dim3(d.nthreads), lds, stream, dev_desc,
                     static_cast<const V2<T>*>(x), static_cast<V2<T>*>(y),
                     static_cast<const V2<T>*>(tw), mat);
  return hipGetLastError();
}

template <typename T, int DIR, int MAT>
static int resident_groups_t(int nthreads, size_t lds) {
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<T, DIR, MAT>), hipFuncAttributeMaxDynamicSharedMemorySize,
                      kLdsBytes);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_pass<T, DIR, MAT>),
                                                   nthreads, lds) != hipSuccess)
    return 0;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return per_cu * cus;
}

}  // namespace fftjh
