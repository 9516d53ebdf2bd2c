// Lean column-pass kernel for the passes of three-pass power-of-two plans
// (FP32): 16 adjacent transforms ("columns") of N = 16*R2 points
// (N = 64, 128, 256) per workgroup of N threads, one group per workgroup.
//
// Everything the generic power-of-two kernel decides at run time is fixed
// here at compile time or made workgroup-uniform, so that the address of
// every HBM access is a scalar base (the group's offset plus u times the
// element step, in SGPRs) and one 32-bit lane offset: no per-lane 64-bit
// index arithmetic, no batch split per lane, no persistent loop whose next
// loads wait behind its stores.
//
// A column of N = 16*R2 points, element m = m_lo + R2*m_hi:
//   stage 1 (thread m_lo = jt, registers u = m_hi): Y[k1] = DFT16 over m_hi,
//     times w_N^(jt*k1) (stage twiddle, FP64-generated table);
//   exchange through LDS;
//   stage 2 (thread jt, for k1 = jt + R2*b): DFT_R2 over m_lo, giving
//     X[k1 + 16*k2] = X[jt + R2*u], u = b + (16/R2)*k2.
// So thread jt loads elements jt + R2*u and ends holding outputs jt + R2*u
// in v[u] (the Stockham autosort order of a two-stage chain).
//
// Edges:
//   ROWS = 0: outputs go back to the same column layout (possibly in
//     place): y[out_base + c + (jt + R2*u)*ostride];
//   ROWS = 1: outputs are transposed through LDS so the group writes its 16
//     columns as 16 contiguous rows: thread t stores element t of each
//     column i at y[out_base + i*ocol + t].
// Multi-pass twiddle (power-of-two period P = 2^psh): output k of column c
// in a group with digits (d0, d1, d2) is multiplied by w_P^e,
//   e = d1*(a0*(16*d0 + c) + b*k) mod P,
// exact in wrapping 32-bit integers (P divides 2^32); w_P^e comes from two
// FP64 tables (w_P^(i*2^sh), w_P^i) as one product rounded to FP32.  With
// ROWS = 1 (a0 = 0) every thread needs one value per group (k = t, the same
// for all 16 columns); with ROWS = 0 a start value per lane times the
// group's step powers S^u, which 16 threads form and stage in LDS.
// The stage table w_N^i is staged in LDS too: table reads from global memory
// (15 per thread) cost as much issue time as the data loads.
//
// Measured on one MI355X (2^24 forward + inverse, 6 launches): the lean
// layout 0.366 -> 0.301 ms against the generic three-pass one; tables in
// LDS 0.287; the same three passes as plain copies (no arithmetic, no
// exchange) take 0.244 ms, the floor of this layout.
#include "butterflies.h"
#include "common.h"
#include "launch.h"
#include "passcommon.h"
#include "stages_mat.h"

#include <cstdlib>

namespace fftjh {

namespace {

__device__ __forceinline__ int lp(int f) { return f + (f >> 4); }

// Group of workgroup b.  Workgroups are dispatched to the 8 XCDs round-robin
// by index and each XCD has its own L2; with `xcd` set, XCD x runs the
// contiguous groups [start(x), start(x + 1)), so neighbouring column groups
// (whose 128-byte runs share lines when the rows are not line-aligned) run on
// one XCD at nearly the same time.  A bijection on [0, groups).
__device__ __forceinline__ uint32_t lean_group(uint32_t b, uint32_t groups, int xcd) {
  if (!xcd) return b;
  const uint32_t q = groups >> 3, r = groups & 7u, x = b & 7u, s = b >> 3;
  return x * q + (x < r ? x : r) + s;
}

// non-temporal (streaming) element load, and the store of an output
// (non-temporal when bit 1 of NT is set)
__device__ __forceinline__ float2 ld_nt2(const float2* p) {
  using NV = float __attribute__((ext_vector_type(2)));
  const NV v = __builtin_nontemporal_load(reinterpret_cast<const NV*>(p));
  return make_float2(v.x, v.y);
}
template <int NT> __device__ __forceinline__ void st_out(float2* p, const float2 v) {
  if constexpr (NT & 2) {
    using NV = float __attribute__((ext_vector_type(2)));
    NV w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<NV*>(p));
  } else {
    *p = v;
  }
}

// SPLIT = 1: the exchanges move the real parts, then the imaginary parts,
// through an LDS array of floats: half the LDS of a float2 exchange (twice
// the groups per CU) for two more barriers per exchange.
// One group per workgroup: a persistent grid that prefetched the next
// group's data while transforming the current one (138-151 VGPRs, 3 waves
// per SIMD) measured slower on the 2^24 plan (0.285 -> 0.303 ms per
// forward + inverse), as did half the occupancy (LDS padding).
// XLANE = 1 (N = 64: one wave, four threads per column in lanes c, c + 16,
// c + 32, c + 48): the exchange between the stages is a 4 x 4 transpose of
// 4-element blocks between the wave's rows of 16 lanes, done in registers by
// v_permlane32_swap (rows r <-> r ^ 2) and v_permlane16_swap (r <-> r ^ 1),
// with no LDS traffic and no barrier.  The pairs (k1 = 4b + s, 4b + 2 + s)
// swap across halves and (4b, 4b + 1), (4b + 2, 4b + 3) across neighbouring
// rows, after which register q holds the stage-2 input q of the LDS path.
// NT: bit 0 loads x non-temporally (streaming: the plan does not read it
// again), bit 1 stores y non-temporally.
// RIN = 1: the 16 transforms of a group are contiguous rows (row i at
// istride*i from the group's base): thread t loads element t of every row
// (whole 256-byte-and-up runs per wave) and an LDS transposition hands
// thread (c, jt) its elements jt + R2*u of row c.
// RS = 1: the loads go through the deconvolution of a resampled length
// (LeanDesc::rs): grid element j of transform d2 is x[d2*rs_n + j'] *
// corr[j'] on the two kept ranges, zero between.
// TWQ = 1 (columns out): the multi-pass twiddle with the column as the
// multiplier (LeanDesc::twq); each lane's start value w_P^(col*(a0*d1 +
// b*jt)) times its column's step powers S_col^u, S_col = w_P^(col*b*R2),
// a 16 x 16 table in LDS.
// REAL = 1 / 2 (with RIN and ROWS): the real edges of an even-length real
// transform of 2N reals whose half-length transform is this pass, as
// LeanDesc::real.  R2C: y_0 = Re Z_0 + Im Z_0, y_N = Re Z_0 - Im Z_0,
// y_k = xe + w_k*(-i*xo), xe, xo = (Z_k +- conj(Z_{N-k}))/2, read from the
// rows-out transposition (both partners are in LDS).  C2R: Z_k = xe +
// i*(w_k*xd), xe, xd = X_k +- conj(X_{N-k}), formed through LDS from the
// rows-in loads (X_N by thread 0) before the rows-in transposition; the 1/2
// and 1/N ride on the pass's scale.
template <int DIR, int N, int ROWS, int SPLIT, int XLANE = 0, int NT = 0, int RIN = 0, int RS = 0, int TWQ = 0,
          int REAL = 0>
__global__ void __launch_bounds__(256) k_lean(LeanDesc d, const float2* __restrict__ x, float2* y,
                                              const float2* __restrict__ tws) {
  constexpr int R2 = N / 16;   // threads per column
  constexpr int BPT = 16 / R2; // stage-2 butterflies per thread
  constexpr int LDS_ELEMS = 16 * N + N + 1;
  // rows-out transposition: column c at stride N + 4 words, so the 64 lanes
  // of a wave (16 columns x 4 elements) write 64 different banks
  constexpr int CS = N + 4;
  static_assert(16 * CS <= LDS_ELEMS, "transposition fits the exchange buffer");
  __shared__ __align__(16) float lf[SPLIT ? LDS_ELEMS : 2 * LDS_ELEMS];
  __shared__ float2 ltw[N];     // the stage table w_N^i
  __shared__ float2 lsu[16];    // the group's twiddle steps S^u (columns out)
  __shared__ float2 lsq[TWQ ? 256 : 1];  // TWQ: S_col^u at [u*16 + c]
  float2* lds = reinterpret_cast<float2*>(lf);
  const int tid = static_cast<int>(threadIdx.x);
  const int c = tid & 15, jt = tid >> 4;
  const int groups = d.groups;

  // the multi-pass twiddle from the FP64 tables -- rows out: w_P^(d1*b*t)
  // for this thread's element t; columns out: the lane's start
  // w_P^(d1*(a + b*jt)) and, in threads u < 16, the group's powers S^u of
  // the step S = w_P^(d1*b*R2), staged in LDS, so that each output takes one
  // FP32 product w0 * S^u (three roundings in all)
  const double2* ptw = static_cast<const double2*>(d.ptw);
  const uint32_t pmask = d.psh >= 32 ? 0xffffffffu : ((1u << d.psh) - 1u);
  const uint32_t lmask = (1u << d.ptw_sh) - 1u;
  auto tw_at = [&](uint32_t e) -> double2 {
    e &= pmask;
    return cmul64(ptw[e >> d.ptw_sh], ptw[d.ptw_lo + (e & lmask)]);
  };
  struct Grp {
    uint32_t d0, d1, d2;
    int64_t ib, ob;
  };
  auto group = [&](uint32_t g) {
    Grp r;
    r.d0 = g % static_cast<uint32_t>(d.D0);
    const uint32_t gr = g / static_cast<uint32_t>(d.D0);
    r.d1 = gr % static_cast<uint32_t>(d.D1);
    r.d2 = gr / static_cast<uint32_t>(d.D1);
    r.ib = r.d0 * d.in_s0 + r.d1 * d.in_s1 + r.d2 * d.in_s2;
    r.ob = r.d0 * d.out_s0 + r.d1 * d.out_s1 + r.d2 * d.out_s2;
    return r;
  };
  // twiddle gathers first, then the data loads (waits on the gathers then
  // do not wait on the data)
  auto fetch = [&](const Grp& G, float2 (&v)[16], double2& w0d, double2& sud) {
    w0d = make_double2(1.0, 0.0);
    sud = make_double2(1.0, 0.0);
    if (d.psh) {
      const uint32_t bb = static_cast<uint32_t>(d.tw_b);
      if constexpr (TWQ) {
        const uint32_t col = 16u * G.d0 + static_cast<uint32_t>(c);
        w0d = tw_at(col * (static_cast<uint32_t>(d.tw_a0) * G.d1 + bb * static_cast<uint32_t>(jt)));
      } else if constexpr (ROWS) {
        w0d = tw_at(G.d1 * bb * static_cast<uint32_t>(tid));
      } else {
        const uint32_t a = static_cast<uint32_t>(d.tw_a0) * (16u * G.d0 + static_cast<uint32_t>(c));
        w0d = tw_at(G.d1 * (a + bb * static_cast<uint32_t>(jt)));
        if (tid < 16) sud = tw_at(G.d1 * bb * static_cast<uint32_t>(R2) * static_cast<uint32_t>(tid));
      }
    }
    if constexpr (RS) {
      // grid indices in 32 bits (the planner keeps m below 2^31)
      const int j0 = static_cast<int>(G.d0 * d.in_s0 + G.d1 * d.in_s1) + c + jt * static_cast<int>(d.istride);
      const float2* xr = x + static_cast<int64_t>(G.d2) * d.rs_n;
      const float* cr = static_cast<const float*>(d.rs_corr);
      const int hp = static_cast<int>((d.rs_n + 1) / 2), shift = static_cast<int>(d.rs_m - d.rs_n);
      const int hi0 = shift + hp;
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = j0 + R2 * u * static_cast<int>(d.istride);
        v[u] = make_float2(0.f, 0.f);
        if (j < hp) {
          v[u] = cscale<float>(xr[j], cr[j]);
        } else if (j >= hi0) {
          v[u] = cscale<float>(xr[j - shift], cr[j - shift]);
        }
      }
      return;
    }
    const float2* xb = x + G.ib;
    if constexpr (RIN) {
      // element tid of row u (transposed after the loads)
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const float2* p = xb + static_cast<int64_t>(u) * d.istride + tid;
        v[u] = (NT & 1) ? ld_nt2(p) : *p;
      }
      return;
    }
    const int lo = c + jt * static_cast<int>(d.istride);
    // (a partial column group: the absent columns load zeros)
    const bool here = !d.ncol || 16 * static_cast<int>(G.d0) + c < d.ncol;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const float2* p = xb + static_cast<int64_t>(R2) * u * d.istride + lo;
      v[u] = !here ? make_float2(0.f, 0.f) : (NT & 1) ? ld_nt2(p) : *p;
    }
  };

  if (static_cast<int>(blockIdx.x) >= groups) return;
  const uint32_t g = lean_group(blockIdx.x, static_cast<uint32_t>(groups), d.xcd);
  const float2 tw_own = tws[tid];
  const Grp G = group(g);
  float2 v[16];
  double2 w0d, sud;
  fetch(G, v, w0d, sud);
  if constexpr (TWQ) {
    if (d.psh) {
      const uint32_t bb = static_cast<uint32_t>(d.tw_b);
      for (int i = tid; i < 256; i += N) {
        const uint32_t coli = 16u * G.d0 + static_cast<uint32_t>(i & 15);
        const double2 sq = tw_at(coli * bb * static_cast<uint32_t>(R2) * static_cast<uint32_t>(i >> 4));
        lsq[i] = make_float2(static_cast<float>(sq.x), static_cast<float>(sq.y));
      }
    }
  }
  if constexpr (REAL == 2) {
    // C2R pre-pass: thread tid holds X_i[tid] of the 16 rows; X_i[N] (for
    // Z_0) comes from thread 0; the partners X_i[N - tid] through LDS
    float2 xn[16];
    if (tid == 0) {
      const float2* xb = x + G.ib;
#pragma unroll
      for (int i = 0; i < 16; ++i) xn[i] = xb[static_cast<int64_t>(i) * d.istride + N];
    }
    const float2 w = static_cast<const float2*>(d.rtw)[tid];
    float br[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) lf[i * CS + tid] = v[i].x;
    if (tid == 0) {
#pragma unroll
      for (int i = 0; i < 16; ++i) lf[i * CS + N] = xn[i].x;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) br[i] = lf[i * CS + N - tid];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) lf[i * CS + tid] = v[i].y;
    if (tid == 0) {
#pragma unroll
      for (int i = 0; i < 16; ++i) lf[i * CS + N] = xn[i].y;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float2 a = v[i];
      const float2 c = make_float2(br[i], -lf[i * CS + N - tid]);  // conj(X_{N-k})
      const float2 xe = make_float2(a.x + c.x, a.y + c.y);
      const float2 xd = make_float2(a.x - c.x, a.y - c.y);
      const float2 t2 = cmul<float>(xd, w);
      v[i] = make_float2(xe.x - t2.y, xe.y + t2.x);  // xe + i*(w*xd)
    }
    __syncthreads();
  }
  if constexpr (RIN) {
    static_assert(SPLIT == 1, "the rows-in transposition uses the float exchange buffer");
    // v[i] = element tid of row i -> v[u] = element jt + R2*u of row c
    // (rows at column stride N + 4 words, as the rows-out transposition)
#pragma unroll
    for (int i = 0; i < 16; ++i) lf[i * CS + tid] = v[i].x;
    __syncthreads();
    float re[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) re[u] = lf[c * CS + jt + R2 * u];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) lf[i * CS + tid] = v[i].y;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = make_float2(re[u], lf[c * CS + jt + R2 * u]);
    __syncthreads();
  }
  ltw[tid] = tw_own;
  const float scale = static_cast<float>(d.scale);
  const bool scaled = d.scale != 1.0;
  {
    if (!ROWS && tid < 16) lsu[tid] = make_float2(static_cast<float>(sud.x), static_cast<float>(sud.y));
    const float2 w0 = make_float2(static_cast<float>(w0d.x), static_cast<float>(w0d.y));
    __syncthreads();
    // stage 1: DFT16 over m_hi, then w_N^(jt*k1); LDS slot [k1][m_lo][c]
    float2 t[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) t[u] = v[u];
    dft_pow2<float, DIR, 16>(t);
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) t[k1] = cmul<float>(t[k1], ltw[jt * k1]);
    // exchange: stage-2 input q = b*R2 + m of this thread is slot
    // [jt + R2*b][m][c]
    auto wslot = [&](int k1) { return lp((k1 * R2 + jt) * 16 + c); };
    auto rslot = [&](int q) { return lp(((jt + R2 * (q / R2)) * R2 + q % R2) * 16 + c); };
    float2 in[16];
    if constexpr (XLANE) {
      static_assert(N == 64, "the register exchange is for one-wave groups");
      auto swap32 = [](float& a, float& b) {
        const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a),
                                                        __builtin_bit_cast(unsigned, b), false, false);
        // (the elements are copied out before the casts: a bit cast of
        // r[1] itself compiled to element 0 at -O3 with this toolchain)
        const unsigned r0 = r[0], r1 = r[1];
        a = __uint_as_float(r0);
        b = __uint_as_float(r1);
      };
      auto swap16 = [](float& a, float& b) {
        const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, a),
                                                        __builtin_bit_cast(unsigned, b), false, false);
        // (the elements are copied out before the casts: a bit cast of
        // r[1] itself compiled to element 0 at -O3 with this toolchain)
        const unsigned r0 = r[0], r1 = r[1];
        a = __uint_as_float(r0);
        b = __uint_as_float(r1);
      };
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
          swap32(t[4 * b + s2].x, t[4 * b + 2 + s2].x);
          swap32(t[4 * b + s2].y, t[4 * b + 2 + s2].y);
        }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        swap16(t[4 * b].x, t[4 * b + 1].x);
        swap16(t[4 * b].y, t[4 * b + 1].y);
        swap16(t[4 * b + 2].x, t[4 * b + 3].x);
        swap16(t[4 * b + 2].y, t[4 * b + 3].y);
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) in[q] = t[q];
    } else if constexpr (SPLIT) {
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) lf[wslot(k1)] = t[k1].x;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) in[q].x = lf[rslot(q)];
      __syncthreads();
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) lf[wslot(k1)] = t[k1].y;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) in[q].y = lf[rslot(q)];
    } else {
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) lds[wslot(k1)] = t[k1];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) in[q] = lds[rslot(q)];
    }
    // stage 2: for k1 = jt + R2*b, DFT_R2 over m_lo; output k2 to t[b + BPT*k2]
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      float2 a[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) a[m] = in[b * R2 + m];
      dft_pow2<float, DIR, R2>(a);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) t[b + BPT * k2] = a[k2];
    }
    if constexpr (REAL == 1) {
      // R2C: the rows-out transposition reads Z_k and Z_{N-k} of every row
      static_assert(ROWS == 1 && SPLIT == 1, "the twist rides on the rows-out transposition");
      float ar[16], br[16];
      const int mk = (N - tid) & (N - 1);
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 16; ++u) lf[c * CS + jt + R2 * u] = t[u].x;
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        ar[i] = lf[i * CS + tid];
        br[i] = lf[i * CS + mk];
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 16; ++u) lf[c * CS + jt + R2 * u] = t[u].y;
      __syncthreads();
      const float2 w = static_cast<const float2*>(d.rtw)[tid];
      float2* yb = y + G.ob;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const float2 a = make_float2(ar[i], lf[i * CS + tid]);
        float2 out;
        if (tid == 0) {
          out = make_float2(a.x + a.y, 0.f);
          float2 last = make_float2(a.x - a.y, 0.f);
          if (scaled) last = cscale<float>(last, scale);
          yb[static_cast<int64_t>(i) * d.ostride + N] = last;
        } else {
          const float2 cc = make_float2(br[i], -lf[i * CS + mk]);  // conj(Z_{N-k})
          const float2 xe = make_float2(0.5f * (a.x + cc.x), 0.5f * (a.y + cc.y));
          const float2 xo = make_float2(0.5f * (a.x - cc.x), 0.5f * (a.y - cc.y));
          const float2 mixo = make_float2(xo.y, -xo.x);  // -i * xo
          const float2 m = cmul<float>(mixo, w);
          out = make_float2(xe.x + m.x, xe.y + m.y);
        }
        if (scaled) out = cscale<float>(out, scale);
        yb[static_cast<int64_t>(i) * d.ostride + tid] = out;
      }
    } else if constexpr (ROWS) {
      // transpose through LDS: column c, element k = jt + R2*u at [c][k];
      // thread t then holds element t of every column i
      float2 o[16];
      __syncthreads();
      if constexpr (SPLIT) {
#pragma unroll
        for (int u = 0; u < 16; ++u) lf[c * CS + jt + R2 * u] = t[u].x;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) o[i].x = lf[i * CS + tid];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 16; ++u) lf[c * CS + jt + R2 * u] = t[u].y;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) o[i].y = lf[i * CS + tid];
      } else {
#pragma unroll
        for (int u = 0; u < 16; ++u) lds[c * CS + jt + R2 * u] = t[u];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) o[i] = lds[i * CS + tid];
      }
      // one twiddle w_P^(d1*b*t) for the thread's element t of every column
      float2 w = w0;
      if (scaled) w = cscale<float>(w, scale);
      float2* yb = y + G.ob;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const float2 a = (d.psh || scaled) ? cmul<float>(o[i], w) : o[i];
        st_out<NT>(yb + i * d.ostride + tid, a);
      }
    } else {
      float2* yb = y + G.ob;
      const int lo = c + jt * static_cast<int>(d.ostride);
      if (d.ncol && 16 * static_cast<int>(G.d0) + c >= d.ncol) return;  // an absent column
      if (TWQ && d.psh) {
        // e_u = col*(a0*d1 + b*(jt + R2*u)) = e_0 + u*col*b*R2
        const float2 ws = scaled ? cscale<float>(w0, scale) : w0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const float2 val = cmul<float>(t[u], u ? cmul<float>(ws, lsq[u * 16 + c]) : ws);
          st_out<NT>(yb + static_cast<int64_t>(R2) * u * d.ostride + lo, val);
        }
      } else if (d.psh) {
        // e_u = d1*(a0*(16*d0 + c) + b*(jt + R2*u)) = e_0 + u*d1*b*R2
        const float2 ws = scaled ? cscale<float>(w0, scale) : w0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const float2 val = cmul<float>(t[u], u ? cmul<float>(ws, lsu[u]) : ws);
          st_out<NT>(yb + static_cast<int64_t>(R2) * u * d.ostride + lo, val);
        }
      } else {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const float2 val = scaled ? cscale<float>(t[u], scale) : t[u];
          st_out<NT>(yb + static_cast<int64_t>(R2) * u * d.ostride + lo, val);
        }
      }
    }
  }
}

// Split-precision MFMA variant (FP32 plans with precision f16, bf16 or
// bf16x3): the radix-16 stages -- stage 1 always, stage 2 when R2 = 16 --
// run as 16x16x32 MFMA products of the split DFT-16 matrix (stages_mat.h),
// 16 butterflies per tile, a tile being the 16 columns at one index.
// Lane l of wave w (g = l/16, col = l%16) loads, for tile t (stage-1 index
// jt = 4w + t), points u = 4g..4g+3 of column col: the instruction's B
// layout, with the loads still 16 columns wide.  It receives outputs
// k1 = 4g..4g+3, which go to the VALU kernel's LDS slots.  Stage 2 with
// R2 = 16 takes tile t = output k1 = 4w + t over m_lo = 4g..4g+3 and leaves
// lane l holding X[(4w + t) + 16*(4g + i)], i < 4; with R2 = 8 or 4 it is
// the VALU stage of k_lean.  The multi-pass twiddle of output k = jt' + R2*u'
// is w0(jt') * S^u' as in k_lean, with one start value per tile.
template <int DIR, int N, int ROWS, int MODE>
__global__ void __launch_bounds__(256) k_lean_mat(LeanDesc d, const float2* __restrict__ x, float2* y,
                                                  const float2* __restrict__ tws) {
  using Ops = MatOps<MODE>;
  using V = typename Ops::V;
  constexpr int T = Ops::terms;
  constexpr int R2 = N / 16;
  constexpr int BPT = 16 / R2;
  constexpr bool MAT2 = R2 == 16;
  constexpr int LDS_ELEMS = 16 * N + N + 1;
  constexpr int CS = N + 4;
  static_assert(16 * CS <= LDS_ELEMS, "transposition fits the exchange buffer");
  __shared__ __align__(16) float lf[LDS_ELEMS];
  __shared__ float2 ltw[N];
  __shared__ float2 lsu[16];
  const int tid = static_cast<int>(threadIdx.x);
  const int lane = tid & 63, wave = tid >> 6;
  const int g4 = lane >> 4, col = lane & 15;
  const int jt = tid >> 4;  // the VALU stage's thread role (column col)
  if (static_cast<int>(blockIdx.x) >= d.groups) return;
  const uint32_t g = lean_group(blockIdx.x, static_cast<uint32_t>(d.groups), d.xcd);

  const double2* ptw = static_cast<const double2*>(d.ptw);
  const uint32_t pmask = d.psh >= 32 ? 0xffffffffu : ((1u << d.psh) - 1u);
  const uint32_t lmask = (1u << d.ptw_sh) - 1u;
  auto tw_at = [&](uint32_t e) -> double2 {
    e &= pmask;
    return cmul64(ptw[e >> d.ptw_sh], ptw[d.ptw_lo + (e & lmask)]);
  };
  const uint32_t d0 = g % static_cast<uint32_t>(d.D0);
  const uint32_t gr = g / static_cast<uint32_t>(d.D0);
  const uint32_t d1 = gr % static_cast<uint32_t>(d.D1);
  const uint32_t d2 = gr / static_cast<uint32_t>(d.D1);
  const int64_t ib = d0 * d.in_s0 + d1 * d.in_s1 + d2 * d.in_s2;
  const int64_t ob = d0 * d.out_s0 + d1 * d.out_s1 + d2 * d.out_s2;

  using GV = const __attribute__((address_space(1))) V*;
  const GV A = (GV)(d.matp) + mat_base(16, T);
  V a1[2][T];
#pragma unroll
  for (int comp = 0; comp < 2; ++comp)
#pragma unroll
    for (int t = 0; t < T; ++t) a1[comp][t] = A[(comp * T + t) * 64 + lane];

  // multi-pass twiddle starts (gathers before the data loads)
  constexpr int NW0 = (!ROWS && MAT2) ? 4 : 1;
  float2 w0[NW0];
  double2 sud = make_double2(1.0, 0.0);
#pragma unroll
  for (int t = 0; t < NW0; ++t) w0[t] = make_float2(1.f, 0.f);
  if (d.psh) {
    const uint32_t bb = static_cast<uint32_t>(d.tw_b);
    auto f2 = [](double2 a) { return make_float2(static_cast<float>(a.x), static_cast<float>(a.y)); };
    if constexpr (ROWS) {
      w0[0] = f2(tw_at(d1 * bb * static_cast<uint32_t>(tid)));
    } else {
      const uint32_t a = static_cast<uint32_t>(d.tw_a0) * (16u * d0 + static_cast<uint32_t>(col));
      if constexpr (MAT2) {
#pragma unroll
        for (int t = 0; t < 4; ++t) w0[t] = f2(tw_at(d1 * (a + bb * static_cast<uint32_t>(4 * wave + t))));
      } else {
        w0[0] = f2(tw_at(d1 * (a + bb * static_cast<uint32_t>(jt))));
      }
      if (tid < 16) sud = tw_at(d1 * bb * static_cast<uint32_t>(R2) * static_cast<uint32_t>(tid));
    }
  }
  const float2 tw_own = tws[tid];
  // loads: tile t, point u = 4*g4 + q of column col is element
  // (4w + t + R2*q) + R2*4*g4
  float2 v[4][4];
  {
    const float2* xb = x + ib;
    const int lo = col + 4 * g4 * R2 * static_cast<int>(d.istride);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) v[t][q] = (xb + static_cast<int64_t>(4 * wave + t + R2 * q) * d.istride)[lo];
  }
  ltw[tid] = tw_own;
  if (!ROWS && tid < 16) lsu[tid] = make_float2(static_cast<float>(sud.x), static_cast<float>(sud.y));
  __syncthreads();

  // one radix-16 MFMA stage on tile values p[4] (points 4*g4..+3 of
  // butterfly col): outputs 4*g4..+3
  auto dft16 = [&](const float2 (&p)[4], float2 (&r)[4]) {
    int ex = 0;
    if constexpr (Ops::scaled) {
      float mx = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) mx = fmaxf(mx, fmaxf(fabsf(p[q].x), fabsf(p[q].y)));
      mx = fmaxf(mx, __shfl_xor(mx, 16));
      mx = fmaxf(mx, __shfl_xor(mx, 32));
      if (mx > 0.f) (void)frexpf(mx, &ex);
    }
    float xs[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      xs[2 * q] = Ops::scaled ? ldexpf(p[q].x, -ex) : p[q].x;
      xs[2 * q + 1] = Ops::scaled ? ldexpf(p[q].y, -ex) : p[q].y;
    }
    V b[T];
    Ops::split(xs, b);
    mat_f4 zr = {0.f, 0.f, 0.f, 0.f}, zi = {0.f, 0.f, 0.f, 0.f};
    mat_dot2<Ops>(a1[0], a1[1], b, zr, zi);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      r[i] = Ops::scaled ? make_float2(ldexpf(zr[i], ex), ldexpf(zi[i], ex)) : make_float2(zr[i], zi[i]);
  };

  // stage 1 and its twiddle w_N^(jt*k1), jt = 4w + t, k1 = 4*g4 + i
  float2 o[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    dft16(v[t], o[t]);
#pragma unroll
    for (int i = 0; i < 4; ++i) o[t][i] = cmul<float>(o[t][i], ltw[(4 * wave + t) * (4 * g4 + i)]);
  }
  // exchange (real parts, then imaginary parts) through slots [k1][jt][col]
  auto wslot = [&](int t, int i) { return lp(((4 * g4 + i) * R2 + 4 * wave + t) * 16 + col); };
  float2 in[16];
  auto rslot = [&](int q) {
    if constexpr (MAT2) return lp(((4 * wave + (q >> 2)) * 16 + 4 * g4 + (q & 3)) * 16 + col);
    return lp(((jt + R2 * (q / R2)) * R2 + q % R2) * 16 + col);
  };
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) lf[wslot(t, i)] = o[t][i].x;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) in[q].x = lf[rslot(q)];
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) lf[wslot(t, i)] = o[t][i].y;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) in[q].y = lf[rslot(q)];

  // stage 2: out[u] holds output k(u) of column col, with
  //   MAT2: u = 4t + i, k = (4w + t) + 16*(4*g4 + i);
  //   else: k = jt + R2*u (the VALU kernel's order)
  float2 out[16];
  if constexpr (MAT2) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float2 p[4] = {in[4 * t], in[4 * t + 1], in[4 * t + 2], in[4 * t + 3]};
      float2 r[4];
      dft16(p, r);
#pragma unroll
      for (int i = 0; i < 4; ++i) out[4 * t + i] = r[i];
    }
  } else {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      float2 a[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) a[m] = in[b * R2 + m];
      dft_pow2<float, DIR, R2>(a);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) out[b + BPT * k2] = a[k2];
    }
  }
  auto kof = [&](int u) {
    if constexpr (MAT2) return (4 * wave + (u >> 2)) + 16 * (4 * g4 + (u & 3));
    return jt + R2 * u;
  };
  const float scale = static_cast<float>(d.scale);
  const bool scaled = d.scale != 1.0;
  float2* yb = y + ob;
  if constexpr (ROWS) {
    // transpose through LDS: column col, element k at [col][k]
    float2 oo[16];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) lf[col * CS + kof(u)] = out[u].x;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) oo[i].x = lf[i * CS + tid];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) lf[col * CS + kof(u)] = out[u].y;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) oo[i].y = lf[i * CS + tid];
    float2 w = w0[0];
    if (scaled) w = cscale<float>(w, scale);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float2 a = (d.psh || scaled) ? cmul<float>(oo[i], w) : oo[i];
      (yb + i * d.ostride)[tid] = a;
    }
  } else if constexpr (MAT2) {
    // k = (4w + t + 16 i) + 64*g4: scalar row part, lane part col + 64*g4*ostride
    const int lo = col + 64 * g4 * static_cast<int>(d.ostride);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float2 ws = w0[t];
      if (scaled) ws = cscale<float>(ws, scale);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float2 val = out[4 * t + i];
        if (d.psh) {
          val = cmul<float>(val, cmul<float>(ws, lsu[4 * g4 + i]));
        } else if (scaled) {
          val = cscale<float>(val, scale);
        }
        (yb + static_cast<int64_t>(4 * wave + t + 16 * i) * d.ostride)[lo] = val;
      }
    }
  } else {
    const int lo = col + jt * static_cast<int>(d.ostride);
    float2 ws = w0[0];
    if (scaled) ws = cscale<float>(ws, scale);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      float2 val = out[u];
      if (d.psh) {
        val = cmul<float>(val, u ? cmul<float>(ws, lsu[u]) : ws);
      } else if (scaled) {
        val = cscale<float>(val, scale);
      }
      (yb + static_cast<int64_t>(R2) * u * d.ostride)[lo] = val;
    }
  }
}

// FFTJH_LEAN_SPLIT: 1 = split exchanges (float LDS), 0 = float2 exchanges
static int lean_split() {
  static const int v = [] {
    const char* e = std::getenv("FFTJH_LEAN_SPLIT");
    return e ? std::atoi(e) : 1;
  }();
  return v;
}

// FFTJH_LEAN_XLANE: 1 = the register exchange for N = 64, 0 = through LDS
static int lean_xlane() {
  static const int v = [] {
    const char* e = std::getenv("FFTJH_LEAN_XLANE");
    return e ? std::atoi(e) : 1;
  }();
  return v;
}

template <int DIR, int N, int ROWS, int SPLIT, int XLANE = 0>
hipError_t launch_k(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  // streaming variants for the split-exchange kernel only
  if constexpr (SPLIT == 1) {
    if (d.nt == 1) {
      hipLaunchKernelGGL((k_lean<DIR, N, ROWS, SPLIT, XLANE, 1>), dim3(d.groups), dim3(N), 0, st, d,
                         static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
      return hipGetLastError();
    }
    if (d.nt == 2) {
      hipLaunchKernelGGL((k_lean<DIR, N, ROWS, SPLIT, XLANE, 2>), dim3(d.groups), dim3(N), 0, st, d,
                         static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_lean<DIR, N, ROWS, SPLIT, XLANE>), dim3(d.groups), dim3(N), 0, st, d,
                     static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
  return hipGetLastError();
}

template <int DIR, int N>
hipError_t launch_n(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  if (d.twq) {
    // a generic layout's middle pass (columns out, the column as the
    // twiddle's multiplier)
    if (d.rows) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_lean<DIR, N, 0, 1, 0, 0, 0, 0, 1>), dim3(d.groups), dim3(N), 0, st, d,
                       static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
    return hipGetLastError();
  }
  if (d.rs) {
    // the first pass of a resampled length's grid transform (rows out)
    if (!d.rows) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_lean<DIR, N, 1, 1, 0, 0, 0, 1>), dim3(d.groups), dim3(N), 0, st, d,
                       static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
    return hipGetLastError();
  }
  if (d.real) {
    // the real edges of rows of 2N reals (rows in and out)
    if (!d.rin || !d.rows) return hipErrorInvalidValue;
    if (d.real == 1)
      hipLaunchKernelGGL((k_lean<DIR, N, 1, 1, 0, 0, 1, 0, 0, 1>), dim3(d.groups), dim3(N), 0, st, d,
                         static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
    else
      hipLaunchKernelGGL((k_lean<DIR, N, 1, 1, 0, 0, 1, 0, 0, 2>), dim3(d.groups), dim3(N), 0, st, d,
                         static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
    return hipGetLastError();
  }
  if (d.rin) {
    // rows in, rows out (the split exchange; non-temporal loads when asked),
    // or rows in, columns out (a generic layout's last pass)
    if (!d.rows) {
      hipLaunchKernelGGL((k_lean<DIR, N, 0, 1, 0, 0, 1>), dim3(d.groups), dim3(N), 0, st, d,
                         static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
      return hipGetLastError();
    }
    if (d.nt == 1) {
      hipLaunchKernelGGL((k_lean<DIR, N, 1, 1, 0, 1, 1>), dim3(d.groups), dim3(N), 0, st, d,
                         static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
      return hipGetLastError();
    }
    hipLaunchKernelGGL((k_lean<DIR, N, 1, 1, 0, 0, 1>), dim3(d.groups), dim3(N), 0, st, d,
                       static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
    return hipGetLastError();
  }
  if constexpr (N == 64) {
    if (lean_xlane())
      return d.rows ? launch_k<DIR, N, 1, 1, 1>(d, x, y, tws, st) : launch_k<DIR, N, 0, 1, 1>(d, x, y, tws, st);
  }
  if (lean_split())
    return d.rows ? launch_k<DIR, N, 1, 1>(d, x, y, tws, st) : launch_k<DIR, N, 0, 1>(d, x, y, tws, st);
  return d.rows ? launch_k<DIR, N, 1, 0>(d, x, y, tws, st) : launch_k<DIR, N, 0, 0>(d, x, y, tws, st);
}

template <int DIR, int N, int MODE>
hipError_t launch_mat_m(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  if (d.rows)
    hipLaunchKernelGGL((k_lean_mat<DIR, N, 1, MODE>), dim3(d.groups), dim3(N), 0, st, d,
                       static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
  else
    hipLaunchKernelGGL((k_lean_mat<DIR, N, 0, MODE>), dim3(d.groups), dim3(N), 0, st, d,
                       static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
  return hipGetLastError();
}

template <int DIR, int N>
hipError_t launch_mat_n(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  switch (d.mat) {
    case kMatF16: return launch_mat_m<DIR, N, kMatF16>(d, x, y, tws, st);
    case kMatBF16: return launch_mat_m<DIR, N, kMatBF16>(d, x, y, tws, st);
    case kMatBF16x3: return launch_mat_m<DIR, N, kMatBF16x3>(d, x, y, tws, st);
  }
  return hipErrorInvalidValue;
}

template <int DIR>
hipError_t launch_dir(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  if (d.mat != kMatNone) {
    switch (d.n) {
      case 64: return launch_mat_n<DIR, 64>(d, x, y, tws, st);
      case 128: return launch_mat_n<DIR, 128>(d, x, y, tws, st);
      case 256: return launch_mat_n<DIR, 256>(d, x, y, tws, st);
    }
    return hipErrorInvalidValue;
  }
  switch (d.n) {
    case 64: return launch_n<DIR, 64>(d, x, y, tws, st);
    case 128: return launch_n<DIR, 128>(d, x, y, tws, st);
    case 256: return launch_n<DIR, 256>(d, x, y, tws, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace

bool lean_length(int n) { return n == 64 || n == 128 || n == 256; }

hipError_t launch_lean(int dir, const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  if (d.cols) return launch_lean_sq(dir, d, x, y, tws, st);  // square lengths (stockham_lean_sq.hip)
  if (d.real >= 3) return launch_lean_pair(dir, d, x, y, tws, st);  // paired real edges
  return dir < 0 ? launch_dir<-1>(d, x, y, tws, st) : launch_dir<1>(d, x, y, tws, st);
}

}  // namespace fftjh
