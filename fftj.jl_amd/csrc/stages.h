// The generic LDS radix stage shared by the pass kernel and the ragged
// persistent kernel: C columns of length n in LDS (element i of column c at
// lpad(i*C + c)), butterflies j = jt + b*jstep of column c per thread.
#pragma once
#include "butterflies.h"
#include "common.h"

namespace fftjh {

// rows a thread moves per unrolled batch of the load phase
constexpr int kIoUnroll = 8;

// The out-of-line stages take LDS and the twiddle table through
// address-space-typed pointers: ds_read/ds_write and global_load with the
// base in a register.  (A generic pointer to the kernel's dynamic LDS would
// be resolved per access through the per-kernel LDS offset table, a scalar
// load and wait in front of every LDS instruction.)
// Elements behind these pointers are plain two-lane vectors (no
// constructors, so they are copied across address spaces on both compiler
// passes); ldv/stv convert to and from V2.
template <typename T> using NV2 = T __attribute__((ext_vector_type(2)));
template <typename T> using LdsPtr = __attribute__((address_space(3))) NV2<T>*;
template <typename T> using TwPtr = const __attribute__((address_space(1))) NV2<T>*;
template <typename T> using GPtr = __attribute__((address_space(1))) NV2<T>*;
template <typename T, typename P> __device__ __forceinline__ V2<T> ldv(P p, int64_t i) {
  const NV2<T> e = p[i];
  return mk<T>(e.x, e.y);
}
template <typename P, typename V> __device__ __forceinline__ void stv(P p, int64_t i, const V& v) {
  p[i] = {v.x, v.y};
}
// (the base goes through an empty asm so that it stays a run-time argument:
// constant-propagated into the stage it would turn back into the table form)
template <typename T> __device__ __forceinline__ LdsPtr<T> as_lds(V2<T>* p) {
  unsigned a = static_cast<unsigned>(reinterpret_cast<uintptr_t>((LdsPtr<T>)(p)));
  asm volatile("" : "+s"(a));
  return (LdsPtr<T>)(static_cast<uintptr_t>(a));
}
template <typename T> __device__ __forceinline__ TwPtr<T> as_tw(const V2<T>* p) { return (TwPtr<T>)(p); }
template <typename T> __device__ __forceinline__ GPtr<T> as_g(V2<T>* p) { return (GPtr<T>)(p); }

// LDS slot of flat index f = i*C + c: one pad element every 16.  With 8- or
// 16-byte elements this turns the stride-R scatter of the early stages (and
// the ns-blocked scatter of the later ones) into the minimum number of bank
// passes for a 64-lane access.
__device__ __forceinline__ int lpad(int f) { return f + (f >> 4); }

// One radix stage, LDS -> registers -> LDS.  Kept out of line: each radix
// then gets its own register allocation (inlined into the stage switch the
// union of all radices' live ranges spills; out of line none does).  The
// twiddle loads are issued first so their latency overlaps the LDS reads and
// the barrier.
template <typename T, int DIR, int R>
__device__ __noinline__ void run_stage(const int n, const int C, const StageDesc s, LdsPtr<T> lds,
                                       TwPtr<T> tw, const int c, const int jt,
                                       const int jstep) {
  constexpr int BPT = kElemsPerThread / R;
  const int nr = n / R;
  const int ns = s.ns;
  V2<T> v[BPT][R];
  V2<T> w[BPT][R > 1 ? R - 1 : 1];
  int kk[BPT];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    kk[b] = 0;
    if (j < nr && ns > 1) {
      const int q = static_cast<int>(__umulhi(static_cast<unsigned>(j), s.ns_magic));
      const int k = j - q * ns;
      kk[b] = k;
      TwPtr<T> twp = tw + (s.tw_off + k);
#pragma unroll
      for (int r = 1; r < R; ++r) w[b][r - 1] = ldv<T>(twp, (r - 1) * ns);
    }
  }
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    if (j < nr) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = ldv<T>(lds, lpad((j + r * nr) * C + c));
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    if (j < nr) {
      if (ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul<T>(v[b][r], w[b][r - 1]);
      }
      dft<T, DIR, R>(v[b]);
      const int base = (j - kk[b]) * R + kk[b];
#pragma unroll
      for (int r = 0; r < R; ++r) stv(lds, lpad((base + r * ns) * C + c), v[b][r]);
    }
  }
  __syncthreads();
}

}  // namespace fftjh
