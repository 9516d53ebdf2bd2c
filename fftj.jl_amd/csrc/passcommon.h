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

// floor(x / D) = (x * m) >> s for 0 <= x < 2^31 and 1 <= D < 2^31, with
// s = 31 + ceil(log2 D) and m = ceil(2^s / D) < 2^32: the error of m is below
// D, so x*m/2^s stays under the next integer (division by invariant integers
// by multiplication; the planner fills m and s)
__device__ __forceinline__ uint32_t fdiv(uint32_t x, uint32_t m, int s) {
  return static_cast<uint32_t>((static_cast<uint64_t>(x) * m) >> s);
}

// transform b split into (bi, bm, bo) by the pass's Bi and B2
struct BatchIndex {
  int64_t bi, bm, bo;
  __device__ BatchIndex(const PassDesc& d, int64_t b) {
    if (d.fast_batch) {
      const uint32_t x = static_cast<uint32_t>(b);
      const uint32_t r = fdiv(x, d.bi_m, d.bi_s);
      const uint32_t o = fdiv(r, d.b2_m, d.b2_s);
      bi = x - r * static_cast<uint32_t>(d.Bi);
      bm = r - o * static_cast<uint32_t>(d.B2);
      bo = o;
    } else {
      bi = b % d.Bi;
      const int64_t r = b / d.Bi;
      bm = r % d.B2;
      bo = r / d.B2;
    }
  }
  __device__ int64_t base(const Geometry& g) const { return bi * g.s_in + bm * g.s_mid + bo * g.s_out; }
};

// (a*b) mod P for a, b < P: below 2^31 through an FP64 estimate of the
// quotient (off by at most one either way, then corrected), else in 128 bits
__device__ __forceinline__ uint64_t mulmod(uint64_t a, uint64_t b, uint64_t P, double invp) {
  if (P < (uint64_t(1) << 31)) {
    const uint64_t x = a * b;
    const uint64_t q = static_cast<uint64_t>(static_cast<double>(x) * invp);
    int64_t r = static_cast<int64_t>(x - q * P);
    if (r < 0) r += static_cast<int64_t>(P);
    if (r < 0) r += static_cast<int64_t>(P);
    if (r >= static_cast<int64_t>(P)) r -= static_cast<int64_t>(P);
    if (r >= static_cast<int64_t>(P)) r -= static_cast<int64_t>(P);
    return static_cast<uint64_t>(r);
  }
  return static_cast<uint64_t>((static_cast<unsigned __int128>(a) * b) % P);
}

__device__ __forceinline__ double2 cmul64(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}

// Multi-pass twiddle for the outputs k = k0 + u*kstep of transform b:
// w_P^e, e = q*(acoef*bm + bcoef*k) mod P.  The start value w_P^e0 and the
// step S = w_P^de, de = q*bcoef*kstep mod P, are each one product of two
// entries of FP64 tables generated at plan time (w_P^(i*2^sh), w_P^i) from
// exact integer exponents; the caller advances w *= S in FP64 and rounds
// each w once to the plan's precision.
struct PostTwiddle {
  double2 w, S;
  uint64_t q;  // the transform's multiplier q (reduced mod P)
  __device__ void init(const PassDesc& d, const BatchIndex& ix, int k0, int kstep) {
    const uint64_t P = static_cast<uint64_t>(d.tw_period);
    const double ip = d.tw_invp;
    if (d.fast_batch)
      q = fdiv(static_cast<uint32_t>(ix.bi), d.qd_m, d.qd_s);
    else
      q = static_cast<uint64_t>(ix.bi / d.tw_qdiv);
    // (the planner's passes have q, bm, k0, kstep < P: the reductions below
    // are guards that do not run)
    auto red = [P](uint64_t v) { return v < P ? v : v % P; };
    q = red(q);
    const uint64_t ar = static_cast<uint64_t>(d.tw_ar), br = static_cast<uint64_t>(d.tw_br);
    uint64_t a = mulmod(ar, red(static_cast<uint64_t>(ix.bm)), P, ip) + mulmod(br, red(static_cast<uint64_t>(k0)), P, ip);
    if (a >= P) a -= P;
    const uint64_t e0 = mulmod(q, a, P, ip);
    const uint64_t de = mulmod(q, mulmod(br, red(static_cast<uint64_t>(kstep)), P, ip), P, ip);
    const double2* t = static_cast<const double2*>(d.ptw);
    const uint64_t mask = (uint64_t(1) << d.ptw_sh) - 1;
    w = cmul64(t[e0 >> d.ptw_sh], t[d.ptw_lo + static_cast<int64_t>(e0 & mask)]);
    S = cmul64(t[de >> d.ptw_sh], t[d.ptw_lo + static_cast<int64_t>(de & mask)]);
  }
  template <typename T>
  __device__ V2<T> value() const { return mk<T>(static_cast<T>(w.x), static_cast<T>(w.y)); }
  __device__ void advance() { w = cmul64(w, S); }
  // the step of a second sequence over the same transform, w_P^(q*bcoef*kstep)
  __device__ double2 step_for(const PassDesc& d, int kstep) const {
    const uint64_t P = static_cast<uint64_t>(d.tw_period);
    const double ip = d.tw_invp;
    const uint64_t kr = static_cast<uint64_t>(kstep) < P ? static_cast<uint64_t>(kstep) : static_cast<uint64_t>(kstep) % P;
    const uint64_t de = mulmod(q, mulmod(static_cast<uint64_t>(d.tw_br), kr, P, ip), P, ip);
    const double2* t = static_cast<const double2*>(d.ptw);
    const uint64_t mask = (uint64_t(1) << d.ptw_sh) - 1;
    return cmul64(t[de >> d.ptw_sh], t[d.ptw_lo + static_cast<int64_t>(de & mask)]);
  }
};

}  // namespace fftjh
