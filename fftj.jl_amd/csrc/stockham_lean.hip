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

#include <cstdlib>

namespace fftjh {

namespace {

__device__ __forceinline__ int lp(int f) { return f + (f >> 4); }

// SPLIT = 1: the exchanges move the real parts, then the imaginary parts,
// through an LDS array of floats: half the LDS of a float2 exchange (twice
// the groups per CU) for two more barriers per exchange.
// One group per workgroup: a persistent grid that prefetched the next
// group's data while transforming the current one (138-151 VGPRs, 3 waves
// per SIMD) measured slower on the 2^24 plan (0.285 -> 0.303 ms per
// forward + inverse), as did half the occupancy (LDS padding).
template <int DIR, int N, int ROWS, int SPLIT>
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
    uint32_t d0, d1;
    int64_t ib, ob;
  };
  auto group = [&](uint32_t g) {
    Grp r;
    r.d0 = g % static_cast<uint32_t>(d.D0);
    const uint32_t gr = g / static_cast<uint32_t>(d.D0);
    r.d1 = gr % static_cast<uint32_t>(d.D1);
    const uint32_t d2 = gr / static_cast<uint32_t>(d.D1);
    r.ib = r.d0 * d.in_s0 + r.d1 * d.in_s1 + d2 * d.in_s2;
    r.ob = r.d0 * d.out_s0 + r.d1 * d.out_s1 + d2 * d.out_s2;
    return r;
  };
  // twiddle gathers first, then the data loads (waits on the gathers then
  // do not wait on the data)
  auto fetch = [&](const Grp& G, float2 (&v)[16], double2& w0d, double2& sud) {
    w0d = make_double2(1.0, 0.0);
    sud = make_double2(1.0, 0.0);
    if (d.psh) {
      const uint32_t bb = static_cast<uint32_t>(d.tw_b);
      if constexpr (ROWS) {
        w0d = tw_at(G.d1 * bb * static_cast<uint32_t>(tid));
      } else {
        const uint32_t a = static_cast<uint32_t>(d.tw_a0) * (16u * G.d0 + static_cast<uint32_t>(c));
        w0d = tw_at(G.d1 * (a + bb * static_cast<uint32_t>(jt)));
        if (tid < 16) sud = tw_at(G.d1 * bb * static_cast<uint32_t>(R2) * static_cast<uint32_t>(tid));
      }
    }
    const float2* xb = x + G.ib;
    const int lo = c + jt * static_cast<int>(d.istride);
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = (xb + static_cast<int64_t>(R2) * u * d.istride)[lo];
  };

  const uint32_t g = blockIdx.x;
  if (static_cast<int>(g) >= groups) return;
  const float2 tw_own = tws[tid];
  const Grp G = group(g);
  float2 v[16];
  double2 w0d, sud;
  fetch(G, v, w0d, sud);
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
    if constexpr (SPLIT) {
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
    if constexpr (ROWS) {
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
        (yb + i * d.ostride)[tid] = a;
      }
    } else {
      float2* yb = y + G.ob;
      const int lo = c + jt * static_cast<int>(d.ostride);
      if (d.psh) {
        // e_u = d1*(a0*(16*d0 + c) + b*(jt + R2*u)) = e_0 + u*d1*b*R2
        const float2 ws = scaled ? cscale<float>(w0, scale) : w0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const float2 val = cmul<float>(t[u], u ? cmul<float>(ws, lsu[u]) : ws);
          (yb + static_cast<int64_t>(R2) * u * d.ostride)[lo] = val;
        }
      } else {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const float2 val = scaled ? cscale<float>(t[u], scale) : t[u];
          (yb + static_cast<int64_t>(R2) * u * d.ostride)[lo] = val;
        }
      }
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

template <int DIR, int N, int ROWS, int SPLIT>
hipError_t launch_k(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  hipLaunchKernelGGL((k_lean<DIR, N, ROWS, SPLIT>), dim3(d.groups), dim3(N), 0, st, d, static_cast<const float2*>(x),
                     static_cast<float2*>(y), static_cast<const float2*>(tws));
  return hipGetLastError();
}

template <int DIR, int N>
hipError_t launch_n(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  if (lean_split())
    return d.rows ? launch_k<DIR, N, 1, 1>(d, x, y, tws, st) : launch_k<DIR, N, 0, 1>(d, x, y, tws, st);
  return d.rows ? launch_k<DIR, N, 1, 0>(d, x, y, tws, st) : launch_k<DIR, N, 0, 0>(d, x, y, tws, st);
}

template <int DIR>
hipError_t launch_dir(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
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
  return dir < 0 ? launch_dir<-1>(d, x, y, tws, st) : launch_dir<1>(d, x, y, tws, st);
}

}  // namespace fftjh
