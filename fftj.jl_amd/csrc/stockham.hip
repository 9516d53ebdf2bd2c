// The pass kernel: C whole length-n transforms per workgroup, all Stockham
// stages of the chain in one launch.
//
// Per workgroup:
//   load   HBM -> LDS, element i of column c to lds[i*C + c]; a wave reads
//          C columns x 64/C consecutive rows (row transforms) or 64/C rows x
//          C adjacent columns (strided axis passes, four-step pass A), so
//          every access is a contiguous run;
//   stages LDS -> registers -> LDS, in place, two barriers per stage;
//   store  LDS -> HBM with the scale and the optional four-step twiddle
//          folded in, same access pattern as the load.
// HBM is touched exactly once per element in each direction per pass.
//
// Reference: the stage loops of stockham.jl (stage4! :26-76, stage8! :81-130,
// stagep! :351-392, stage2! :394-428) run one radix pass per sweep over split
// planes; run_chain! (plan.jl:311-321) strings them together.  Here the whole
// chain runs inside one workgroup.
#include "butterflies.h"
#include "common.h"
#include "launch.h"

namespace fftjh {

// rows a thread moves per unrolled batch of the load/store phases
constexpr int kIoUnroll = 8;

// LDS slot of flat index f = i*C + c: one pad element every 16.  With 8- or
// 16-byte elements this turns the stride-R scatter of the early stages (and
// the ns-blocked scatter of the later ones) into the minimum number of bank
// passes for a 64-lane access.
__device__ __forceinline__ int lpad(int f) { return f + (f >> 4); }

// One radix stage, LDS -> registers -> LDS.  Kept out of line: each radix
// then gets its own register allocation (inlined into the stage switch the
// union of all radices' live ranges spills; out of line none does).
template <typename T, int DIR, int R>
__device__ __noinline__ void run_stage(const int n, const int C, const StageDesc& s, V2<T>* lds,
                                          const V2<T>* __restrict__ tw, const int c, const int jt,
                                          const int jstep) {
  constexpr int BPT = kElemsPerThread / R;
  const int nr = n / R;
  V2<T> v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    if (j < nr) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = lds[lpad((j + r * nr) * C + c)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    if (j < nr) {
      int k = 0;
      if (s.ns > 1) {
        const int q = static_cast<int>(__umulhi(static_cast<unsigned>(j), s.ns_magic));
        k = j - q * s.ns;
        const V2<T>* twp = tw + (s.tw_off + k);
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul<T>(v[b][r], twp[(r - 1) * s.ns]);
      }
      dft<T, DIR, R>(v[b]);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[lpad((base + r * s.ns) * C + c)] = v[b][r];
    }
  }
  __syncthreads();
}

// logical workgroup for hardware workgroup w: the dispatcher deals
// workgroups round-robin over the 8 XCDs, so give each XCD one contiguous
// range of logical groups (adjacent columns then share that XCD's L2)
__device__ __forceinline__ int xcd_logical(int w, int G) {
  const int x = w & 7, l = w >> 3;
  const int q = G >> 3, r = G & 7;
  return x * q + min(x, r) + l;
}

template <typename T> struct MaxThreads { static constexpr int value = kMaxThreads; };
template <> struct MaxThreads<double> { static constexpr int value = kMaxThreads / 2; };

// The descriptor sits in device memory (written once at plan time): the
// stage loop indexes it with a uniform, dynamic stage number, which turns
// into scalar loads instead of a private copy of a by-value argument.
template <typename T, int DIR>
__global__ void __launch_bounds__(MaxThreads<T>::value)
k_pass(const PassDesc* __restrict__ dp, const V2<T>* x, V2<T>* y, const V2<T>* __restrict__ tw,
       const V2<T>* __restrict__ thi, const V2<T>* __restrict__ tlo) {
  extern __shared__ __align__(16) unsigned char smem[];
  V2<T>* lds = reinterpret_cast<V2<T>*>(smem);
  const PassDesc& d = *dp;
  const int n = d.n, C = d.C;
  const int g = d.xcd_remap ? xcd_logical(static_cast<int>(blockIdx.x), d.ngroups)
                            : static_cast<int>(blockIdx.x);
  const int tid = static_cast<int>(threadIdx.x);
  const int c = tid & (C - 1);
  const int jt = tid >> d.logC;
  const int jstep = d.nthreads >> d.logC;
  const int64_t b = static_cast<int64_t>(g) * C + c;
  const bool valid = b < d.batch;
  const int64_t bi = b % d.Bi, bo = b / d.Bi;

  // load
  {
    const V2<T>* xb = x + (bi * d.in.s_in + bo * d.in.s_out);
    const int64_t st = d.in.stride;
    for (int i0 = jt; i0 < n; i0 += kIoUnroll * jstep) {
      V2<T> t[kIoUnroll];
#pragma unroll
      for (int u = 0; u < kIoUnroll; ++u) {
        const int i = i0 + u * jstep;
        t[u] = (valid && i < n) ? xb[i * st] : mk<T>(T(0), T(0));
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
    switch (st.radix) {
      case 2: run_stage<T, DIR, 2>(n, C, st, lds, tw, c, jt, jstep); break;
      case 3: run_stage<T, DIR, 3>(n, C, st, lds, tw, c, jt, jstep); break;
      case 4: run_stage<T, DIR, 4>(n, C, st, lds, tw, c, jt, jstep); break;
      case 5: run_stage<T, DIR, 5>(n, C, st, lds, tw, c, jt, jstep); break;
      case 7: run_stage<T, DIR, 7>(n, C, st, lds, tw, c, jt, jstep); break;
      case 8: run_stage<T, DIR, 8>(n, C, st, lds, tw, c, jt, jstep); break;
      case 11: run_stage<T, DIR, 11>(n, C, st, lds, tw, c, jt, jstep); break;
      case 13: run_stage<T, DIR, 13>(n, C, st, lds, tw, c, jt, jstep); break;
      default: run_stage<T, DIR, 16>(n, C, st, lds, tw, c, jt, jstep); break;
    }
  }

  // store
  if (valid) {
    V2<T>* yb = y + (bi * d.out.s_in + bo * d.out.s_out);
    const int64_t st = d.out.stride;
    const T scale = static_cast<T>(d.scale);
    const bool scaled = d.scale != 1.0;
    const bool post = d.post_tw != 0;
    const uint32_t j1 = static_cast<uint32_t>(bi);
    const uint32_t lo_mask = (1u << d.lo_bits) - 1u;
    for (int k0 = jt; k0 < n; k0 += kIoUnroll * jstep) {
#pragma unroll
      for (int u = 0; u < kIoUnroll; ++u) {
        const int k = k0 + u * jstep;
        if (k < n) {
          V2<T> v = lds[lpad(k * C + c)];
          if (post) {
            // e = j1*k < N <= 2^32 for every four-step this library plans
            const uint32_t e = j1 * static_cast<uint32_t>(k);
            v = cmul<T>(v, cmul<T>(thi[e >> d.lo_bits], tlo[e & lo_mask]));
          }
          if (scaled) v = cscale<T>(v, scale);
          yb[k * st] = v;
        }
      }
    }
  }
}

template <typename T, int DIR>
static hipError_t launch_pass_t(const PassDesc& d, const PassDesc* dev_desc, const void* x, void* y,
                                const void* tw,
                                const void* thi, const void* tlo, hipStream_t stream) {
  static bool attr_set = false;
  if (!attr_set) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<T, DIR>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    attr_set = true;
  }
  const size_t lds = static_cast<size_t>(lds_elems(d.C * d.n)) * sizeof(V2<T>);
  hipLaunchKernelGGL((k_pass<T, DIR>), dim3(d.ngroups), dim3(d.nthreads), lds, stream, dev_desc,
                     static_cast<const V2<T>*>(x), static_cast<V2<T>*>(y),
                     static_cast<const V2<T>*>(tw), static_cast<const V2<T>*>(thi),
                     static_cast<const V2<T>*>(tlo));
  return hipGetLastError();
}

int max_threads(int dtype) { return dtype == 0 ? MaxThreads<float>::value : MaxThreads<double>::value; }

hipError_t launch_pass(int dtype, int dir, const PassDesc& d, const PassDesc* dev_desc, const void* x,
                       void* y, const void* tw, const void* thi, const void* tlo, hipStream_t stream) {
  if (dtype == 0)
    return dir < 0 ? launch_pass_t<float, -1>(d, dev_desc, x, y, tw, thi, tlo, stream)
                   : launch_pass_t<float, 1>(d, dev_desc, x, y, tw, thi, tlo, stream);
  return dir < 0 ? launch_pass_t<double, -1>(d, dev_desc, x, y, tw, thi, tlo, stream)
                 : launch_pass_t<double, 1>(d, dev_desc, x, y, tw, thi, tlo, stream);
}

}  // namespace fftjh
