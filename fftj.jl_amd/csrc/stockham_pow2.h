#pragma once
// Implementation of the pass kernel for lengths with a compile-time radix
// chain: the powers of two 32..16384 (radices from {2, 4, 8, 16}) and the
// listed fixed chains of other lengths (100 = 10 x 10).
//
// Each thread owns E elements of a transform (16 for the powers of two, 10
// for 100), rows jt + u*J (u < E, J = n/E); every radix R of the chain
// divides E.  With the autosort indexing a radix-R stage's butterflies
// j = jt + b*J (b < E/R) read rows j + r*n/R = jt + J*(b + r*E/R) -- the
// same E rows -- and the last stage (k == j) writes those rows too.  So:
//   stage 1   runs on the registers the HBM loads filled, writes LDS;
//   stages    LDS -> registers -> LDS (two barriers each);
//   last      LDS -> registers -> HBM, with scale and four-step twiddle.
// Two LDS round trips for a 4096-point transform instead of five, no calls,
// everything unrolled at compile time.  The kernel is persistent: each
// workgroup walks several groups (FFTJH_P2_PREFETCH=1 also loads the next
// group's registers during the middle stages; measured no faster, off).
//
// Four-step twiddle w_N^{j1*k} for k = jt + u*J: w = A * S^u with A and S
// each a product of two FP64 table entries at exact integer exponents, the
// powers formed in FP64 and rounded once (PostTwiddle, passcommon.h).
#include "butterflies.h"
#include "common.h"
#include "launch.h"
#include "passcommon.h"

namespace fftjh {

namespace {


#ifndef FFTJH_P2_MAXT
#define FFTJH_P2_MAXT 512
#endif
#ifndef FFTJH_P2_PREFETCH
#define FFTJH_P2_PREFETCH 0
#endif
template <typename T> struct MaxThreadsP2 { static constexpr int value = FFTJH_P2_MAXT; };
template <> struct MaxThreadsP2<double> { static constexpr int value = FFTJH_P2_MAXT / 2; };

__device__ __forceinline__ int lpad2(int f) { return f + (f >> 4); }

// non-temporal (streaming) element load / store
template <typename T> __device__ __forceinline__ V2<T> ld_nt(const V2<T>* p) {
  using NV = T __attribute__((ext_vector_type(2)));
  const NV v = __builtin_nontemporal_load(reinterpret_cast<const NV*>(p));
  return mk<T>(v.x, v.y);
}
template <typename T> __device__ __forceinline__ void st_nt(V2<T>* p, const V2<T>& v) {
  using NV = T __attribute__((ext_vector_type(2)));
  NV w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<NV*>(p));
}

// the chain's radices, and per stage its ns and twiddle-table offset
// (entries (r-1)*ns + k per stage)
template <int R1, int R2, int R3, int R4>
struct Radices {
  static constexpr int count = (R1 > 1) + (R2 > 1) + (R3 > 1) + (R4 > 1);
  static constexpr int get(int s) { return s == 0 ? R1 : s == 1 ? R2 : s == 2 ? R3 : R4; }
  static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * get(s - 1); }
  static constexpr int off(int s) { return s == 0 ? 0 : off(s - 1) + (get(s - 1) - 1) * ns(s - 1); }
};

// butterflies of one stage on the thread's registers: b < 16/R, rows
// j = jt + b*J, inputs v[b + r*16/R]; twiddles w[b][r-1] already applied
template <typename T, int DIR, int R, int E>
__device__ __forceinline__ void stage_regs(V2<T> (&v)[E]) {
  constexpr int BPT = E / R;
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    V2<T> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[b + r * BPT];
    dft<T, DIR, R>(a);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b + r * BPT] = a[r];
  }
}

// twiddles of stage (R, NS) for this thread's butterflies, from the table
template <typename T, int R, int NS, int E>
__device__ __forceinline__ void stage_twiddle(V2<T> (&v)[E], const V2<T>* tw, int jt, int J) {
  if constexpr (NS > 1) {
    constexpr int BPT = E / R;
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = jt + b * J;
      const int k = j % NS;
      const V2<T>* twp = tw + k;
#pragma unroll
      for (int r = 1; r < R; ++r) v[b + r * BPT] = cmul<T>(v[b + r * BPT], twp[(r - 1) * NS]);
    }
  }
}

// scatter a stage's outputs into LDS: output r of butterfly j at
// (j - k)*R + k + r*NS, k = j mod NS
template <typename T, int R, int NS, int E>
__device__ __forceinline__ void stage_scatter(const V2<T> (&v)[E], V2<T>* lds, int C, int c, int jt, int J) {
  constexpr int BPT = E / R;
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * J;
    const int k = j % NS;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[lpad2((base + r * NS) * C + c)] = v[b + r * BPT];
  }
}

// gather a stage's inputs from LDS: input r of butterfly j at j + r*N/R
template <typename T, int N, int R, int E>
__device__ __forceinline__ void stage_gather(V2<T> (&v)[E], const V2<T>* lds, int C, int c, int jt, int J) {
  constexpr int BPT = E / R;
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = jt + b * J;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b + r * BPT] = lds[lpad2((j + r * (N / R)) * C + c)];
  }
}

template <typename T, int DIR, int N, typename Ch, int S, int E>
__device__ __forceinline__ void middle_stages(V2<T> (&v)[E], V2<T>* lds, const V2<T>* tw, int C,
                                              int c, int jt) {
  if constexpr (S < Ch::count - 1) {
    constexpr int R = Ch::get(S), NS = Ch::ns(S), OFF = Ch::off(S);
    constexpr int J = N / E;
    stage_gather<T, N, R, E>(v, lds, C, c, jt, J);
    __syncthreads();
    stage_twiddle<T, R, NS, E>(v, tw + OFF, jt, J);
    stage_regs<T, DIR, R, E>(v);
    stage_scatter<T, R, NS, E>(v, lds, C, c, jt, J);
    __syncthreads();
    middle_stages<T, DIR, N, Ch, S + 1, E>(v, lds, tw, C, c, jt);
  }
}

}  // namespace

// IO: 0 = registers on both HBM sides, 1 = rows staged through LDS on the
// way in, 2 = on the way out (separate instantiations keep the plain
// kernel's register allocation small); 3 = R2C rows (register I/O, the
// real twist at the store), 4 = C2R rows (the real pre-pass at the load),
// 5 / 6 = the same edges on the last / first pass of a multi-pass
// half-length transform, partners exchanged between paired columns in LDS;
// 7 = the resampling deconvolution at the load of a first pass; 8 / 9 / 10
// = register I/O with non-temporal loads / stores / both
#ifndef FFTJH_P2_WPE
#define FFTJH_P2_WPE 0
#endif
template <typename T, int DIR, int N, int R1, int R2, int R3, int R4, int IO, int kE = 16>
__global__ void __launch_bounds__(MaxThreadsP2<T>::value)
#if FFTJH_P2_WPE > 0
__attribute__((amdgpu_waves_per_eu(FFTJH_P2_WPE)))
#endif
k_pass_pow2(const PassDesc* __restrict__ dp, const V2<T>* x, V2<T>* y, const V2<T>* tw) {
  using Ch = Radices<R1, R2, R3, R4>;
  constexpr int NST = Ch::count;
  // IO 8 / 9 / 10: register I/O (as 0) with streamed loads / stores / both
  constexpr bool kNtLoad = IO == 8 || IO == 10, kNtStore = IO == 9 || IO == 10;
  constexpr int J = N / kE;
  static_assert(NST >= 2, "fused first and last stages need two stages");
  extern __shared__ __align__(16) unsigned char smem[];
  V2<T>* lds = reinterpret_cast<V2<T>*>(smem);
  const PassDesc& d = *dp;
  const int C = d.C;
  const int tid = static_cast<int>(threadIdx.x);
  const int c = tid & (C - 1);
  const int jt = tid >> d.logC;  // < J
  const int64_t ist = d.in.stride, ost = d.out.stride;
  const GroupWalk walk(d);

  V2<T> v[kE], pf[kE];
  const int64_t istep = static_cast<int64_t>(J) * ist, ostep = static_cast<int64_t>(J) * ost;
  // transform of column cc in group gg.  Paired real edges (IO 5, 6): with
  // M = pair_M transforms per row and G1 = M/32 groups per row, group
  // bo*G1 + a holds T = 16a + i (columns i < 16) and their mirrors M - T
  // (columns 16 + i); the slot of the mirror of T = 0 holds T = M/2, the
  // other self-mirrored transform.
  const int pM = d.pair_M;
  auto bidx = [&](int gg, int cc) -> int64_t {
    if constexpr (IO == 5 || IO == 6) {
      const int G1 = pM >> 5;
      const int bo = gg / G1, a = gg - bo * G1, i = cc & 15;
      const int Tm = cc < 16 ? 16 * a + i : ((a == 0 && i == 0) ? (pM >> 1) : pM - 16 * a - i);
      return static_cast<int64_t>(bo) * pM + Tm;
    } else {
      return static_cast<int64_t>(gg) * C + cc;
    }
  };
  // column holding the mirror of this thread's transform in group gg
  auto mirror_col = [&](int gg) {
    const int a = gg % (pM >> 5);
    return (a == 0 && (c & 15) == 0) ? c : (c ^ 16);
  };
  // the real edges' twiddle w_{2h}^e (h the half length: N for the row
  // kernels, real_h for the paired ones) from the FP64 two-level table
  auto rtw_at = [&](uint64_t e) -> double2 {
    const double2* t = static_cast<const double2*>(d.rtw);
    const uint64_t mask = (uint64_t(1) << d.rtw_sh) - 1;
    return cmul64(t[e >> d.rtw_sh], t[d.rtw_hi + static_cast<int64_t>(e & mask)]);
  };
  auto load = [&](int g, V2<T> (&dst)[kE]) {
    const int64_t b = bidx(g, c);
    if (b < d.batch) {
      const V2<T>* p = x + (BatchIndex(d, b).base(d.in) + jt * ist);
      if constexpr (kNtLoad) {
#pragma unroll
        for (int u = 0; u < kE; ++u) {
          dst[u] = ld_nt<T>(p);
          p += istep;
        }
      } else {
#pragma unroll
        for (int u = 0; u < kE; ++u) {
          dst[u] = *p;
          p += istep;
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < kE; ++u) dst[u] = mk<T>(T(0), T(0));
    }
  };

  // Row-major staging for contiguous rows when C columns share a workgroup:
  // lanes run along a row (f = tid + u*nthreads -> column f / N, row f % N),
  // so every HBM access is a contiguous run, and LDS does the transposition.
  const int nthr = d.nthreads;
  auto load_rows = [&](int g) {
    const int64_t b0 = static_cast<int64_t>(g) * C;
    const V2<T>* xg = x + BatchIndex(d, b0).base(d.in);
    const int64_t left = d.batch - b0;
#pragma unroll
    for (int u = 0; u < kE; ++u) {
      const int f = tid + u * nthr;
      const int cc = f / N, i = f % N;
      V2<T> val = mk<T>(T(0), T(0));
      if (cc < left) val = xg[cc * d.in_cdelta + i * ist];
      lds[lpad2(i * C + cc)] = val;
    }
  };
  auto store_rows = [&](int g) {
    const int64_t b0 = static_cast<int64_t>(g) * C;
    V2<T>* yg = y + BatchIndex(d, b0).base(d.out);
    const int64_t left = d.batch - b0;
#pragma unroll
    for (int u = 0; u < kE; ++u) {
      const int f = tid + u * nthr;
      const int cc = f / N, k = f % N;
      // (the planner never combines rows_out with a multi-pass twiddle)
      if (cc < left) {
        V2<T> val = lds[lpad2(k * C + cc)];
        if (d.scale != 1.0) val = cscale<T>(val, static_cast<T>(d.scale));
        yg[cc * d.out_cdelta + k * ost] = val;
      }
    }
  };

  // C2R edge (IO 4, real.jl:149-193): thread (c, jt) forms its 16 inputs
  // Z_j, j = jt + u*J, of row b from the bins X_j and X_{N-j} (both read
  // from HBM; the second read of a row hits L2) in registers:
  //   Z_j = xe + i*(w_j*xd),  xe, xd = X_j +- conj(X_{N-j}),  w_j = e^{+2 pi i j/2N}
  // (irfft's pre-factor 1/2 is a power of two and rides, exactly, on the
  // store scale)
  auto load_pre = [&](int g, V2<T> (&dst)[kE]) {
    const int64_t b = static_cast<int64_t>(g) * C + c;
    if (b < d.batch) {
      const V2<T>* xr = x + BatchIndex(d, b).base(d.in);
      // w_j, j = jt + u*J: FP64 start and step
      double2 wv = rtw_at(static_cast<uint64_t>(jt));
      const double2 ws = rtw_at(static_cast<uint64_t>(J));
#pragma unroll
      for (int u = 0; u < kE; ++u) {
        const int j = jt + u * J;
        const V2<T> a = xr[j];
        const V2<T> xc = xr[N - j];
        const V2<T> cx = mk<T>(xc.x, -xc.y);
        const V2<T> xe = cadd<T>(a, cx);
        const V2<T> t = cmul<T>(csub<T>(a, cx), mk<T>(static_cast<T>(wv.x), static_cast<T>(wv.y)));
        wv = cmul64(wv, ws);
        dst[u] = mk<T>(xe.x - t.y, xe.y + t.x);
      }
    } else {
#pragma unroll
      for (int u = 0; u < kE; ++u) dst[u] = mk<T>(T(0), T(0));
    }
  };

  // C2R edge of a multi-pass first pass (IO 6): column P = T of the
  // half-length input j = P + M*j3 (j3 = jt + u*J here); the partner bin
  // X_{h-j} is element N-1-j3 of column M-P (element (N-j3) mod N of column
  // 0 for P = 0, and X_h itself for j = 0), exchanged through LDS; then the
  // same pre-pass as IO 4
  auto load_pair = [&](int gg, V2<T> (&dst)[kE]) {
    load(gg, dst);
#pragma unroll
    for (int u = 0; u < kE; ++u) lds[lpad2((jt + u * J) * C + c)] = dst[u];
    __syncthreads();
    const int64_t b = bidx(gg, c);
    const int P = static_cast<int>(b % pM);
    const int pc = mirror_col(gg);
    // w_j, j = P + M*(jt + u*J): FP64 start and step, stepped per element
    double2 wv = rtw_at(static_cast<uint64_t>(P) + static_cast<uint64_t>(pM) * jt);
    const double2 ws = rtw_at(static_cast<uint64_t>(pM) * J);
    const V2<T>* xr = x + BatchIndex(d, b).base(d.in);
#pragma unroll
    for (int u = 0; u < kE; ++u) {
      const int j3 = jt + u * J;
      const V2<T> wj = mk<T>(static_cast<T>(wv.x), static_cast<T>(wv.y));
      wv = cmul64(wv, ws);
      V2<T> xc;
      if (P == 0 && j3 == 0) {
        xc = xr[d.real_h];
      } else {
        const int jp = P == 0 ? ((N - j3) & (N - 1)) : (N - 1 - j3);
        xc = lds[lpad2(jp * C + pc)];
      }
      const V2<T> a = dst[u];
      const V2<T> cx = mk<T>(xc.x, -xc.y);
      const V2<T> xe = cadd<T>(a, cx);
      const V2<T> t = cmul<T>(csub<T>(a, cx), wj);
      dst[u] = mk<T>(xe.x - t.y, xe.y + t.x);
    }
    __syncthreads();
  };

  // resampling deconvolution at the load (IO 7): element j of this
  // transform's m-grid row, from x and the correction factors
  auto load_rs = [&](int gg, V2<T> (&dst)[kE]) {
    const int64_t b = bidx(gg, c);
    if (b < d.batch) {
      const BatchIndex ix(d, b);
      // grid indices in 32 bits (the planner keeps m below 2^31)
      const int j0 = static_cast<int>(ix.base(d.in) - ix.bo * d.in.s_out);
      const int js = static_cast<int>(ist);
      const V2<T>* xr = x + ix.bo * d.rs_n;
      const T* cr = static_cast<const T*>(d.rs_corr);
      const int hp = static_cast<int>((d.rs_n + 1) / 2), shift = static_cast<int>(d.rs_m - d.rs_n);
      const int hi0 = shift + hp;
#pragma unroll
      for (int u = 0; u < kE; ++u) {
        const int j = j0 + (jt + u * J) * js;
        dst[u] = mk<T>(T(0), T(0));
        if (j < hp) {
          dst[u] = cscale<T>(xr[j], cr[j]);
        } else if (j >= hi0) {
          dst[u] = cscale<T>(xr[j - shift], cr[j - shift]);
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < kE; ++u) dst[u] = mk<T>(T(0), T(0));
    }
  };

  int g = walk.start;
  if (g >= walk.end) return;
  if constexpr (IO == 1) {
    load_rows(g);
    __syncthreads();
    stage_gather<T, N, R1, kE>(v, lds, C, c, jt, J);
    __syncthreads();
  } else if constexpr (IO == 4) {
    load_pre(g, v);
  } else if constexpr (IO == 6) {
    load_pair(g, v);
  } else if constexpr (IO == 7) {
    load_rs(g, v);
  } else {
    load(g, v);
  }
  const T scale = static_cast<T>(d.scale);
  const bool scaled = d.scale != 1.0;
  while (true) {
    // stage 1 on the loaded registers (ns = 1: no twiddles)
    stage_regs<T, DIR, R1, kE>(v);
    stage_scatter<T, R1, 1, kE>(v, lds, C, c, jt, J);
    __syncthreads();
    const int gn = g + walk.step;
    const bool more = gn < walk.end;
    if (FFTJH_P2_PREFETCH && more) load(gn, pf);
    middle_stages<T, DIR, N, Ch, 1, kE>(v, lds, tw, C, c, jt);
    // last stage: LDS -> registers -> HBM
    {
      constexpr int S = NST - 1;
      constexpr int R = Ch::get(S), NS = Ch::ns(S), OFF = Ch::off(S);
      const int64_t b = bidx(g, c);
      const BatchIndex ix(d, b);
      // the multi-pass twiddle's base and step, formed while v is dead
      PostTwiddle pt;
      if ((IO <= 1 || IO >= 6) && d.tw_period) pt.init(d, ix, jt, J);
      stage_gather<T, N, R, kE>(v, lds, C, c, jt, J);
      __syncthreads();
      stage_twiddle<T, R, NS, kE>(v, tw + OFF, jt, J);
      stage_regs<T, DIR, R, kE>(v);
      if constexpr (IO == 2) {
        // last stage: k == j, so output r of butterfly j goes to row j + r*N/R
        stage_scatter<T, R, N / R, kE>(v, lds, C, c, jt, J);
        __syncthreads();
        store_rows(g);
        __syncthreads();
      } else if constexpr (IO == 3) {
        // R2C edge (real.jl:88-131): v[u] = Z_k, k = jt + u*J; the partner
        // bins Z_{N-k} come through LDS (same column, lanes contiguous)
        //   y_0, y_N = Re Z_0 +- Im Z_0,  y_k = xe + w_k*(-i*xo),
        //   xe, xo = (Z_k +- conj(Z_{N-k}))/2,  w_k = e^{-2 pi i k/2N}
#pragma unroll
        for (int u = 0; u < kE; ++u) lds[lpad2((jt + u * J) * C + c)] = v[u];
        __syncthreads();
        if (b < d.batch) {
          V2<T>* yr = y + ix.base(d.out);
          double2 wv = rtw_at(static_cast<uint64_t>(jt));
          const double2 ws = rtw_at(static_cast<uint64_t>(J));
#pragma unroll
          for (int u = 0; u < kE; ++u) {
            const int k = jt + u * J;
            const V2<T> z = v[u];
            const V2<T> wk = mk<T>(static_cast<T>(wv.x), static_cast<T>(wv.y));
            wv = cmul64(wv, ws);
            if (u == 0 && jt == 0) {
              yr[0] = mk<T>(z.x + z.y, T(0));
              yr[N] = mk<T>(z.x - z.y, T(0));
            } else {
              const V2<T> zc = lds[lpad2((N - k) * C + c)];
              const V2<T> cz = mk<T>(zc.x, -zc.y);
              const T half = T(0.5);
              const V2<T> xe = cscale<T>(cadd<T>(z, cz), half);
              const V2<T> xo = cscale<T>(csub<T>(z, cz), half);
              yr[k] = cadd<T>(xe, cmul<T>(mk<T>(xo.y, -xo.x), wk));
            }
          }
        }
        __syncthreads();
      } else if constexpr (IO == 5) {
        // R2C edge of a multi-pass last pass: v[u] = Z_k, k = T + M*k1,
        // k1 = jt + u*J; Z_{h-k} is element N-1-k1 of transform M-T
        // ((N-k1) mod N of transform 0 for T = 0), from the mirror column
        // through LDS; the twist as IO 3, y_k at T + M*k1 of the h+1 bins
#pragma unroll
        for (int u = 0; u < kE; ++u) lds[lpad2((jt + u * J) * C + c)] = v[u];
        __syncthreads();
        const int Tm = static_cast<int>(b % pM);
        const int pc = mirror_col(g);
        V2<T>* yr = y + ix.base(d.out);
        double2 wv = rtw_at(static_cast<uint64_t>(Tm) + static_cast<uint64_t>(pM) * jt);
        const double2 ws = rtw_at(static_cast<uint64_t>(pM) * J);
#pragma unroll
        for (int u = 0; u < kE; ++u) {
          const int k1 = jt + u * J;
          const V2<T> z = v[u];
          const V2<T> wk = mk<T>(static_cast<T>(wv.x), static_cast<T>(wv.y));
          wv = cmul64(wv, ws);
          if (Tm == 0 && k1 == 0) {
            yr[0] = mk<T>(z.x + z.y, T(0));
            yr[d.real_h] = mk<T>(z.x - z.y, T(0));
          } else {
            const int kp = Tm == 0 ? ((N - k1) & (N - 1)) : (N - 1 - k1);
            const V2<T> zc = lds[lpad2(kp * C + pc)];
            const V2<T> cz = mk<T>(zc.x, -zc.y);
            const T half = T(0.5);
            const V2<T> xe = cscale<T>(cadd<T>(z, cz), half);
            const V2<T> xo = cscale<T>(csub<T>(z, cz), half);
            yr[k1 * ost] = cadd<T>(xe, cmul<T>(mk<T>(xo.y, -xo.x), wk));
          }
        }
        __syncthreads();
      } else {
       if (b < d.batch) {
        V2<T>* yb = y + (ix.base(d.out) + jt * ost);
        if (d.tw_period) {
#pragma unroll
          for (int u = 0; u < kE; ++u) {
            V2<T> val = cmul<T>(v[u], pt.value<T>());
            if (scaled) val = cscale<T>(val, scale);
            *yb = val;
            yb += ostep;
            pt.advance();
          }
        } else if constexpr (kNtStore) {
#pragma unroll
          for (int u = 0; u < kE; ++u) {
            st_nt<T>(yb, scaled ? cscale<T>(v[u], scale) : v[u]);
            yb += ostep;
          }
        } else {
#pragma unroll
          for (int u = 0; u < kE; ++u) {
            *yb = scaled ? cscale<T>(v[u], scale) : v[u];
            yb += ostep;
          }
        }
       }
      }
    }
    if (!more) break;
    if (FFTJH_P2_PREFETCH) {
#pragma unroll
      for (int u = 0; u < kE; ++u) v[u] = pf[u];
    } else if constexpr (IO == 1) {
      load_rows(gn);
      __syncthreads();
      stage_gather<T, N, R1, kE>(v, lds, C, c, jt, J);
      __syncthreads();
    } else if constexpr (IO == 4) {
      load_pre(gn, v);
    } else if constexpr (IO == 6) {
      load_pair(gn, v);
    } else if constexpr (IO == 7) {
      load_rs(gn, v);
    } else {
      load(gn, v);
    }
    g = gn;
  }
}

namespace {

template <typename T, int DIR, int N, int R1, int R2, int R3, int R4, int IO, int E = 16>
hipError_t launch_p2_io(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                        hipStream_t st) {
  static bool attr = false;
  auto* fn = &k_pass_pow2<T, DIR, N, R1, R2, R3, R4, IO, E>;
  if (!attr) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    attr = true;
  }
  const size_t lds = static_cast<size_t>(lds_elems(d.C * N)) * sizeof(V2<T>);
  const int grid = d.grid > 0 ? d.grid : d.ngroups;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(d.nthreads), lds, st, dd, static_cast<const V2<T>*>(x),
                     static_cast<V2<T>*>(y), static_cast<const V2<T>*>(tw));
  return hipGetLastError();
}

template <typename T, int DIR, int N, int R1, int R2, int R3, int R4>
hipError_t launch_p2(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                     hipStream_t st) {
  if (d.real_io) {
    // the real edges exist in one direction each: R2C forward, C2R inverse;
    // the paired ones need 32 columns of N points in one workgroup
    constexpr bool paired = 2 * N <= MaxThreadsP2<T>::value;
    if constexpr (DIR == -1) {
      if (d.real_io == 1) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 3>(d, dd, x, y, tw, st);
      if constexpr (paired)
        if (d.real_io == 3) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 5>(d, dd, x, y, tw, st);
    } else {
      if (d.real_io == 2) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 4>(d, dd, x, y, tw, st);
      if constexpr (paired)
        if (d.real_io == 4) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 6>(d, dd, x, y, tw, st);
    }
    return hipErrorInvalidValue;
  }
  if (d.rs_io) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 7>(d, dd, x, y, tw, st);
  if (d.rows_in) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 1>(d, dd, x, y, tw, st);
  if (d.rows_out) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 2>(d, dd, x, y, tw, st);
  if (d.nt_load && d.nt_store) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 10>(d, dd, x, y, tw, st);
  if (d.nt_load) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 8>(d, dd, x, y, tw, st);
  if (d.nt_store) return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 9>(d, dd, x, y, tw, st);
  return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 0>(d, dd, x, y, tw, st);
}

template <typename T, int DIR, int N, int R1, int R2, int R3, int R4, int IO, int E = 16>
int resident_p2_io(const PassDesc& d) {
  auto* fn = &k_pass_pow2<T, DIR, N, R1, R2, R3, R4, IO, E>;
  hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
  const size_t lds = static_cast<size_t>(lds_elems(d.C * N)) * sizeof(V2<T>);
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), d.nthreads, lds) !=
      hipSuccess)
    return 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return per_cu * cus;
}

template <typename T, int DIR, int N, int R1, int R2, int R3, int R4>
int resident_p2(const PassDesc& d) {
  if (d.real_io) {
    constexpr bool paired = 2 * N <= MaxThreadsP2<T>::value;
    if constexpr (DIR == -1) {
      if (d.real_io == 1) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 3>(d);
      if constexpr (paired)
        if (d.real_io == 3) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 5>(d);
    } else {
      if (d.real_io == 2) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 4>(d);
      if constexpr (paired)
        if (d.real_io == 4) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 6>(d);
    }
    return 0;
  }
  if (d.rs_io) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 7>(d);
  if (d.rows_in) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 1>(d);
  if (d.rows_out) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 2>(d);
  if (d.nt_load && d.nt_store) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 10>(d);
  if (d.nt_load) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 8>(d);
  if (d.nt_store) return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 9>(d);
  return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 0>(d);
}

// fixed chains of other lengths (register I/O only): E elements per thread,
// every radix dividing E, N/E threads per transform
template <typename T, int DIR, int N, int R1, int R2, int R3, int R4, int E>
hipError_t launch_fixed(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                        hipStream_t st) {
  if (d.real_io || d.rs_io || d.rows_in || d.rows_out) return hipErrorInvalidValue;
  return launch_p2_io<T, DIR, N, R1, R2, R3, R4, 0, E>(d, dd, x, y, tw, st);
}
template <typename T, int DIR, int N, int R1, int R2, int R3, int R4, int E>
int resident_fixed(const PassDesc& d) {
  return resident_p2_io<T, DIR, N, R1, R2, R3, R4, 0, E>(d);
}

// the compile-time chains: n -> radices (16s first, the remainder last)
#define FFTJH_POW2_CHAINS(X)  \
  X(32, 16, 2, 1, 1)          \
  X(64, 16, 4, 1, 1)          \
  X(128, 16, 8, 1, 1)         \
  X(256, 16, 16, 1, 1)        \
  X(512, 16, 16, 2, 1)        \
  X(1024, 16, 16, 4, 1)       \
  X(2048, 16, 16, 8, 1)       \
  X(4096, 16, 16, 16, 1)      \
  X(8192, 16, 16, 16, 2)      \
  X(16384, 16, 16, 16, 4)
// n -> radices, elements per thread (non-power-of-two lengths of the
// multi-pass plans: 10^6 = 100^3; 81 for the passes of 3^k lengths whose
// load is strided)
#define FFTJH_FIXED_CHAINS(X) \
  X(100, 10, 10, 1, 1, 10)    \
  X(81, 9, 9, 1, 1, 27)

template <typename T, int DIR>
hipError_t dispatch_launch(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                           hipStream_t st) {
  switch (d.n) {
#define X(N, A, B, C_, D_) \
  case N:                  \
    return launch_p2<T, DIR, N, A, B, C_, D_>(d, dd, x, y, tw, st);
    FFTJH_POW2_CHAINS(X)
#undef X
#define X(N, A, B, C_, D_, E) \
  case N:                     \
    return launch_fixed<T, DIR, N, A, B, C_, D_, E>(d, dd, x, y, tw, st);
    FFTJH_FIXED_CHAINS(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <typename T, int DIR>
int dispatch_resident(const PassDesc& d) {
  switch (d.n) {
#define X(N, A, B, C_, D_) \
  case N:                  \
    return resident_p2<T, DIR, N, A, B, C_, D_>(d);
    FFTJH_POW2_CHAINS(X)
#undef X
#define X(N, A, B, C_, D_, E) \
  case N:                     \
    return resident_fixed<T, DIR, N, A, B, C_, D_, E>(d);
    FFTJH_FIXED_CHAINS(X)
#undef X
  }
  return 0;
}

}  // namespace

}  // namespace fftjh
