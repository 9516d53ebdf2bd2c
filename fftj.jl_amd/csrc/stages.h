// The generic LDS radix stage shared by the pass kernel and the ragged
// persistent kernel: C columns of length n in LDS (element i of column c at
// lpad(i*C + c)), butterflies j = jt + b*jstep of column c per thread.
#pragma once
#include "butterflies.h"
#include "common.h"

namespace fftjh {

// rows a thread moves per unrolled batch of the load phase
constexpr int kIoUnroll = 8;

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
__device__ __noinline__ void run_stage(const int n, const int C, const StageDesc& s, V2<T>* lds,
                                       const V2<T>* tw, const int c, const int jt,
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
      const V2<T>* twp = tw + (s.tw_off + k);
#pragma unroll
      for (int r = 1; r < R; ++r) w[b][r - 1] = twp[(r - 1) * ns];
    }
  }
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
      if (ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul<T>(v[b][r], w[b][r - 1]);
      }
      dft<T, DIR, R>(v[b]);
      const int base = (j - kk[b]) * R + kk[b];
#pragma unroll
      for (int r = 0; r < R; ++r) lds[lpad((base + r * ns) * C + c)] = v[b][r];
    }
  }
  __syncthreads();
}

}  // namespace fftjh
