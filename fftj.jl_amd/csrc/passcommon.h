// Device helpers shared by the pass kernels: the workgroup -> group schedule,
// the three-level batch geometry and the multi-pass twiddle.
#pragma once
#include "butterflies.h"
#include "common.h"

namespace fftjh {

// Logical groups a workgroup runs.  The dispatcher deals workgroups
// round-robin over the 8 XCDs, so workgroup w sits on XCD w % 8.  With
// xcd_remap each XCD owns one contiguous range of logical groups and its
// workgroups walk that range side by side, so the groups in flight on one
// XCD are adjacent (adjacent columns then share that XCD's L2 lines).  A
// grid of one workgroup per group walks one group each.
struct GroupWalk {
  int start, step, end;
  __device__ explicit GroupWalk(const PassDesc& d) {
    const int nwg = static_cast<int>(gridDim.x), w = static_cast<int>(blockIdx.x);
    const int G = d.ngroups;
    if (d.xcd_remap && nwg >= 8) {
      const int x = w & 7, l = w >> 3;
      const int wx = (nwg - x + 7) >> 3;  // workgroups on XCD x
      int before = 0;                      // workgroups on XCDs < x
      for (int i = 0; i < x; ++i) before += (nwg - i + 7) >> 3;
      start = static_cast<int>((static_cast<int64_t>(G) * before) / nwg) + l;
      step = wx;
      end = static_cast<int>((static_cast<int64_t>(G) * (before + wx)) / nwg);
    } else {
      start = w;
      step = nwg;
      end = G;
    }
  }
};

// transform b split into (bi, bm, bo) by the pass's Bi and B2
struct BatchIndex {
  int64_t bi, bm, bo;
  __device__ BatchIndex(const PassDesc& d, int64_t b) {
    bi = b % d.Bi;
    const int64_t r = b / d.Bi;
    bm = r % d.B2;
    bo = r / d.B2;
  }
  __device__ int64_t base(const Geometry& g) const { return bi * g.s_in + bm * g.s_mid + bo * g.s_out; }
};

// exp(dir*2*pi*i*e/L) in FP64 from an exact integer numerator
__device__ __forceinline__ double2 expi64(uint64_t e, uint64_t L, int dir) {
  e %= L;
  int64_t es = static_cast<int64_t>(e);
  if (2 * e > L) es -= static_cast<int64_t>(L);
  double s, c;
  sincospi(2.0 * static_cast<double>(es) / static_cast<double>(L), &s, &c);
  return make_double2(c, dir < 0 ? -s : s);
}

__device__ __forceinline__ double2 cmul64(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}

// (a*b) mod m for a, b < m < 2^63
__device__ __forceinline__ uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) {
  return static_cast<uint64_t>((static_cast<unsigned __int128>(a) * b) % m);
}

// Multi-pass twiddle for the outputs k = k0 + u*kstep of transform b:
// w_P^{q*(acoef*bm + bcoef*k)} = A * S^u, A and S from exact exponents; the
// caller advances w *= S in FP64 and rounds each w once.
struct PostTwiddle {
  double2 w, S;
  template <int DIR>
  __device__ void init(const PassDesc& d, const BatchIndex& ix, int k0, int kstep) {
    const uint64_t P = static_cast<uint64_t>(d.tw_period);
    const uint64_t q = static_cast<uint64_t>(ix.bi / d.tw_qdiv) % P;
    const uint64_t a = (mulmod(static_cast<uint64_t>(d.tw_acoef) % P, static_cast<uint64_t>(ix.bm) % P, P) +
                        mulmod(static_cast<uint64_t>(d.tw_bcoef) % P, static_cast<uint64_t>(k0), P)) % P;
    w = expi64(mulmod(q, a, P), P, DIR);
    S = expi64(mulmod(q, mulmod(static_cast<uint64_t>(d.tw_bcoef) % P, static_cast<uint64_t>(kstep), P), P), P,
               DIR);
  }
  __device__ void advance() { w = cmul64(w, S); }
};

}  // namespace fftjh
