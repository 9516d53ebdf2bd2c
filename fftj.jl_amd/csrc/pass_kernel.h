// The pass kernel template shared by the plain chain (stockham.hip) and the
// split-precision MFMA chain (stockham_mat.hip).
#pragma once
#include "butterflies.h"
#include "common.h"
#include "passcommon.h"
#include "stages.h"
#include "stages_mat.h"

namespace fftjh {

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
