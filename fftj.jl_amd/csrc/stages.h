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

// Values that are the same in every lane of a wave but reach an out-of-line
// function in vector registers (its arguments): reading lane 0 puts them in
// scalar registers, so the index arithmetic on them runs on the scalar unit
// and the memory instructions take a scalar base.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uni64(int64_t x) {
  const uint64_t u = static_cast<uint64_t>(x);
  const uint64_t lo = static_cast<unsigned>(uni(static_cast<int>(u & 0xffffffffu)));
  const uint64_t hi = static_cast<unsigned>(uni(static_cast<int>(u >> 32)));
  return static_cast<int64_t>(lo | (hi << 32));
}
template <typename T> __device__ __forceinline__ LdsPtr<T> uni_lds(LdsPtr<T> p) {
  const unsigned a = static_cast<unsigned>(reinterpret_cast<uintptr_t>(p));
  return (LdsPtr<T>)(static_cast<uintptr_t>(static_cast<unsigned>(uni(static_cast<int>(a)))));
}
template <typename P> __device__ __forceinline__ P uni_ptr(P p) {
  const uint64_t a = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(p));
  const uint64_t lo = static_cast<unsigned>(uni(static_cast<int>(a & 0xffffffffu)));
  const uint64_t hi = static_cast<unsigned>(uni(static_cast<int>(a >> 32)));
  return (P)(static_cast<uintptr_t>(lo | (hi << 32)));
}

// One radix stage, LDS -> registers -> LDS.  Kept out of line: each radix
// then gets its own register allocation (inlined into the stage switch the
// union of all radices' live ranges spills; out of line none does).  The
// twiddle loads are issued first so their latency overlaps the LDS reads and
// the barrier.  Every argument but c and jt is uniform over the workgroup.
// A butterfly reads flat slots f + r*dr and writes g + r*dw (dr = C*n/R,
// dw = C*ns); when a step is a multiple of 16, lpad is affine along it,
// lpad(f + r*d) = lpad(f) + r*(d + d/16), and each access costs one add.
template <typename T, int DIR, int R>
__device__ __noinline__ void run_stage(const int n_, const int C_, const StageDesc s, LdsPtr<T> lds_,
                                       TwPtr<T> tw_, const int c, const int jt,
                                       const int jstep_) {
  constexpr int BPT = kElemsPerThread / R;
  const int n = uni(n_), C = uni(C_), ns = uni(s.ns), jstep = uni(jstep_);
  const unsigned magic = static_cast<unsigned>(uni(static_cast<int>(s.ns_magic)));
  const LdsPtr<T> lds = uni_lds<T>(lds_);
  using TwBytes = const __attribute__((address_space(1))) char*;
  const TwBytes twb = reinterpret_cast<TwBytes>(uni_ptr(tw_) + uni(s.tw_off));
  const int nr = n / R;
  const int dr = nr * C, dw = ns * C;
  const bool lin_r = (dr & 15) == 0, lin_w = (dw & 15) == 0;
  // per-r offsets of the affine case, formed on the scalar unit
  int orr[R], orw[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    orr[r] = uni(r * (dr + (dr >> 4)));
    orw[r] = uni(r * (dw + (dw >> 4)));
  }
  V2<T> v[BPT][R];
  V2<T> w[BPT][R > 1 ? R - 1 : 1];
  int kk[BPT];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    kk[b] = 0;
    if (j < nr && ns > 1) {
      const int q = static_cast<int>(__umulhi(static_cast<unsigned>(j), magic));
      const int k = j - static_cast<int>(__umul24(static_cast<unsigned>(q), static_cast<unsigned>(ns)));
      kk[b] = k;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        // a 32-bit byte offset from the scalar base (the saddr form)
        const unsigned off = static_cast<unsigned>(k + (r - 1) * ns) * static_cast<unsigned>(sizeof(NV2<T>));
        const NV2<T> e = *reinterpret_cast<TwPtr<T>>(twb + off);
        w[b][r - 1] = mk<T>(e.x, e.y);
      }
    }
  }
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * jstep;
    if (j < nr) {
      const int f = static_cast<int>(__umul24(static_cast<unsigned>(j), static_cast<unsigned>(C))) + c;
      if (lin_r) {
        const LdsPtr<T> p = lds + lpad(f);
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = ldv<T>(p, orr[r]);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = ldv<T>(lds, lpad(f + r * dr));
      }
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
      const int f = static_cast<int>(__umul24(static_cast<unsigned>(base), static_cast<unsigned>(C))) + c;
      if (lin_w) {
        const LdsPtr<T> p = lds + lpad(f);
#pragma unroll
        for (int r = 0; r < R; ++r) stv(p, orw[r], v[b][r]);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) stv(lds, lpad(f + r * dw), v[b][r]);
      }
    }
  }
  __syncthreads();
}

}  // namespace fftjh
