// Lean kernel for the paired real edges of a multi-pass half-length
// transform (FP32): the last pass of an R2C plan (MODE 3: the twist at the
// store) or the first pass of a C2R plan (MODE 4: the pre-pass at the load),
// N = 64 or 128 points.  The half-length transform of h = N*M points has M
// transforms T per row (real transform); output (R2C) or input (C2R) bin
// k = T + M*k1 needs its partner h - k = (M - T) + M*(N - 1 - k1), element
// N - 1 - k1 of transform M - T (for T = 0: element (N - k1) of T itself;
// T = M/2 is its own mirror).  A workgroup holds two sub-groups of 16
// columns: A the transforms T = 16a + c, B their mirrors M - 16a - 15 + c
// (adjacent in memory, so both load and store 16-column runs), except that
// for a = 0 B's column 15 (which would be M) holds M/2.  Column c of A pairs
// with column 15 - c of B; T = 0 and M/2 pair with themselves.  Each
// sub-group runs the lean kernel's two radix stages (16 x R2, the exchange
// through its own LDS buffer); the partners meet in LDS.
//
// Sides: transform T of row r sits at (T % pbi)*s0 + ((T / pbi) % pb2)*s1 +
// r*s2 on either side (the pass's batch split, pbi*pb2 = M), elements at the
// side's stride.  MODE 3 reads 16 contiguous rows (istride 1) per sub-group,
// transposed through LDS as the lean rows-in form, and stores columns of the
// h + 1 bins; MODE 4 loads columns of the h + 1 bins and stores columns with
// the multi-pass twiddle w_P^e, e = q*(a0*bm + b*k), q = (T % pbi)/qdiv,
// bm = (T / pbi) % pb2, each lane's start value times its column's step
// powers (a 2 x 16 x 16 table in LDS).
//
// The real edges (real.jl:88-131, 149-193): R2C y_0, y_h = Re Z_0 +- Im Z_0,
// y_k = xe + w_k*(-i*xo), xe, xo = (Z_k +- conj(Z_{h-k}))/2; C2R
// Z_j = xe + i*(w_j*xd), xe, xd = X_j +- conj(X_{h-j}) (its 1/2 rides on the
// scale); w_k = exp(dir*2*pi*i*k/(2h)) from an FP64 two-level table, stepped
// in FP64 along k1 and rounded once per output.
#include "butterflies.h"
#include "common.h"
#include "launch.h"
#include "passcommon.h"

namespace fftjh {

namespace {

__device__ __forceinline__ int lpp(int f) { return f + (f >> 4); }

// XCD-local group order (as the lean kernel's): XCD x runs the contiguous
// groups [start(x), start(x + 1)); a bijection on [0, groups)
__device__ __forceinline__ uint32_t pair_group(uint32_t b, uint32_t groups, int xcd) {
  if (!xcd) return b;
  const uint32_t q = groups >> 3, r = groups & 7u, x = b & 7u, s = b >> 3;
  return x * q + (x < r ? x : r) + s;
}

template <int DIR, int N, int MODE>
__global__ void __launch_bounds__(2 * N) k_lean_pair(LeanDesc d, const float2* __restrict__ x, float2* y,
                                                     const float2* __restrict__ tws) {
  static_assert(MODE == 3 || MODE == 4, "paired R2C last pass or C2R first pass");
  constexpr int R2 = N / 16;           // threads per column
  constexpr int BPT = 16 / R2;         // stage-2 butterflies per thread
  constexpr int GT = N;                // threads per sub-group
  constexpr int CS = N + 4;            // column stride (words) of the pairing / transposition buffers
  constexpr int XE = 16 * N + N + 1;   // one sub-group's exchange buffer (floats)
  static_assert(16 * CS <= XE, "the pairing buffer fits the exchange buffer");
  __shared__ __align__(16) float lf[2 * XE];
  __shared__ float2 ltw[N];
  __shared__ float2 lsq[MODE == 4 ? 512 : 1];
  __shared__ int64_t lbase[2][2][16];  // [side: in, out][sub-group][column]
  const int tid = static_cast<int>(threadIdx.x);
  const int sub = tid / GT, st = tid - sub * GT;
  const int c = st & 15, jt = st >> 4;
  float* const own = lf + sub * XE;

  if (static_cast<int>(blockIdx.x) >= d.groups) return;
  const uint32_t g = pair_group(blockIdx.x, static_cast<uint32_t>(d.groups), d.xcd);
  const uint32_t M = static_cast<uint32_t>(d.pm), Ga = M >> 5;
  const uint32_t a = g % Ga, row = g / Ga;
  const uint32_t pbi = static_cast<uint32_t>(d.pbi), pb2 = static_cast<uint32_t>(d.pb2);
  auto col_T = [&](int s, int cc) -> uint32_t {
    if (s == 0) return 16u * a + static_cast<uint32_t>(cc);
    if (a == 0 && cc == 15) return M >> 1;
    return M - 16u * a - 15u + static_cast<uint32_t>(cc);
  };
  if (tid < 64) {
    const int side = tid >> 5, s = (tid >> 4) & 1, cc = tid & 15;
    const uint32_t T = col_T(s, cc);
    const int64_t bi = T % pbi, bm = (T / pbi) % pb2;
    lbase[side][s][cc] = side == 0 ? bi * d.in_s0 + bm * d.in_s1 + static_cast<int64_t>(row) * d.in_s2
                                   : bi * d.out_s0 + bm * d.out_s1 + static_cast<int64_t>(row) * d.out_s2;
  }
  if (sub == 0) ltw[st] = tws[st];
  const uint32_t T = col_T(sub, c);
  // the partner of element k of this column: element N - 1 - k of the other
  // sub-group's column 15 - c; T = 0: element (N - k) of itself; T = M/2:
  // element N - 1 - k of itself
  const bool zero_col = sub == 0 && a == 0 && c == 0;
  const bool half_col = sub == 1 && a == 0 && c == 15;
  const float* const pbuf = (zero_col || half_col) ? own : lf + (1 - sub) * XE;
  const int pcol = (zero_col || half_col) ? c : 15 - c;

  const double2* rtw = static_cast<const double2*>(d.rtw);
  const uint64_t rmask = (uint64_t(1) << d.rtw_sh) - 1;
  auto rtw_at = [&](uint64_t e) -> double2 { return cmul64(rtw[e >> d.rtw_sh], rtw[d.rtw_hi + static_cast<int64_t>(e & rmask)]); };
  const double2* ptw = static_cast<const double2*>(d.ptw);
  const uint32_t pmask = d.psh >= 32 ? 0xffffffffu : ((1u << d.psh) - 1u);
  const uint32_t lmask = (1u << d.ptw_sh) - 1u;
  auto tw_at = [&](uint32_t e) -> double2 {
    e &= pmask;
    return cmul64(ptw[e >> d.ptw_sh], ptw[d.ptw_lo + (e & lmask)]);
  };
  __syncthreads();  // lbase, ltw

  const float scale = static_cast<float>(d.scale);
  const bool scaled = d.scale != 1.0;
  float2 v[16];
  double2 w0d = make_double2(1.0, 0.0);
  if constexpr (MODE == 3) {
    // rows in: thread st loads element st of the sub-group's 16 rows, then
    // the rows-in transposition hands thread (c, jt) elements jt + R2*u of row c
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = x[lbase[0][sub][i] + st];
#pragma unroll
    for (int i = 0; i < 16; ++i) own[i * CS + st] = v[i].x;
    __syncthreads();
    float re[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) re[u] = own[c * CS + jt + R2 * u];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) own[i * CS + st] = v[i].y;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = make_float2(re[u], own[c * CS + jt + R2 * u]);
    __syncthreads();
  } else {
    // columns in (the h + 1 bins): element i = jt + R2*u of column T; the
    // column T = 0 also needs X_h (element N), loaded by its thread jt = 0
    const float2* xc = x + lbase[0][sub][c];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = xc[static_cast<int64_t>(jt + R2 * u) * d.istride];
    float2 xh = make_float2(0.f, 0.f);
    if (zero_col && jt == 0) xh = xc[static_cast<int64_t>(N) * d.istride];
    // the multi-pass twiddle: this lane's start and both sub-groups' step
    // powers S_col^u (S_col = w_P^(q*b*R2)), formed while the loads fly
    if (d.psh) {
      const uint32_t bb = static_cast<uint32_t>(d.tw_b), aa = static_cast<uint32_t>(d.tw_a0);
      const uint32_t qd = static_cast<uint32_t>(d.pqdiv);
      const uint32_t q = (T % pbi) / qd, bm = (T / pbi) % pb2;
      w0d = tw_at(q * (aa * bm + bb * static_cast<uint32_t>(jt)));
      for (int i = tid; i < 512; i += 2 * GT) {
        const uint32_t Ti = col_T(i >> 8, i & 15);
        const uint32_t qi = (Ti % pbi) / qd;
        const double2 sq = tw_at(qi * bb * static_cast<uint32_t>(R2) * static_cast<uint32_t>((i >> 4) & 15));
        lsq[i] = make_float2(static_cast<float>(sq.x), static_cast<float>(sq.y));
      }
    }
    // pre-pass: the partners through LDS (real parts, then imaginary parts)
    float pr[16];
    auto pidx = [&](int i) { return zero_col ? N - i : N - 1 - i; };
#pragma unroll
    for (int u = 0; u < 16; ++u) own[c * CS + jt + R2 * u] = v[u].x;
    if (zero_col && jt == 0) own[N] = xh.x;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) pr[u] = pbuf[pcol * CS + pidx(jt + R2 * u)];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) own[c * CS + jt + R2 * u] = v[u].y;
    if (zero_col && jt == 0) own[N] = xh.y;
    __syncthreads();
    double2 wv = rtw_at(static_cast<uint64_t>(T) + static_cast<uint64_t>(M) * static_cast<uint64_t>(jt));
    const double2 ws = rtw_at(static_cast<uint64_t>(M) * R2);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const float2 xa = v[u];
      const float2 xp = make_float2(pr[u], -pbuf[pcol * CS + pidx(jt + R2 * u)]);  // conj(X_{h-j})
      const float2 xe = make_float2(xa.x + xp.x, xa.y + xp.y);
      const float2 xd = make_float2(xa.x - xp.x, xa.y - xp.y);
      const float2 t2 = cmul<float>(xd, make_float2(static_cast<float>(wv.x), static_cast<float>(wv.y)));
      v[u] = make_float2(xe.x - t2.y, xe.y + t2.x);  // xe + i*(w*xd)
      wv = cmul64(wv, ws);
    }
    __syncthreads();
  }

  // the lean kernel's two stages on the sub-group's buffer:
  // stage 1: DFT16 over m_hi, then w_N^(jt*k1); stage 2: DFT_R2 over m_lo
  float2 t[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) t[u] = v[u];
  dft_pow2<float, DIR, 16>(t);
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) t[k1] = cmul<float>(t[k1], ltw[jt * k1]);
  auto wslot = [&](int k1) { return lpp((k1 * R2 + jt) * 16 + c); };
  auto rslot = [&](int q) { return lpp(((jt + R2 * (q / R2)) * R2 + q % R2) * 16 + c); };
  float2 in[16];
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) own[wslot(k1)] = t[k1].x;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) in[q].x = own[rslot(q)];
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) own[wslot(k1)] = t[k1].y;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) in[q].y = own[rslot(q)];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    float2 av[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) av[m] = in[b * R2 + m];
    dft_pow2<float, DIR, R2>(av);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) t[b + BPT * k2] = av[k2];
  }
  // t[u] = output k1 = jt + R2*u of column T

  if constexpr (MODE == 3) {
    // the twist: the partners through LDS, the h + 1 bins out as columns
    __syncthreads();
    auto pidx = [&](int k) { return zero_col ? ((N - k) & (N - 1)) : N - 1 - k; };
    float pr[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) own[c * CS + jt + R2 * u] = t[u].x;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) pr[u] = pbuf[pcol * CS + pidx(jt + R2 * u)];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) own[c * CS + jt + R2 * u] = t[u].y;
    __syncthreads();
    float2* yc = y + lbase[1][sub][c];
    double2 wv = rtw_at(static_cast<uint64_t>(T) + static_cast<uint64_t>(M) * static_cast<uint64_t>(jt));
    const double2 ws = rtw_at(static_cast<uint64_t>(M) * R2);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int k1 = jt + R2 * u;
      const float2 z = t[u];
      if (zero_col && k1 == 0) {
        float2 lo = make_float2(z.x + z.y, 0.f), hi = make_float2(z.x - z.y, 0.f);
        if (scaled) {
          lo = cscale<float>(lo, scale);
          hi = cscale<float>(hi, scale);
        }
        yc[0] = lo;
        yc[static_cast<int64_t>(N) * d.ostride] = hi;
      } else {
        const float2 cz = make_float2(pr[u], -pbuf[pcol * CS + pidx(k1)]);  // conj(Z_{h-k})
        const float2 xe = make_float2(0.5f * (z.x + cz.x), 0.5f * (z.y + cz.y));
        const float2 xo = make_float2(0.5f * (z.x - cz.x), 0.5f * (z.y - cz.y));
        const float2 m = cmul<float>(make_float2(xo.y, -xo.x), make_float2(static_cast<float>(wv.x), static_cast<float>(wv.y)));
        float2 out = make_float2(xe.x + m.x, xe.y + m.y);
        if (scaled) out = cscale<float>(out, scale);
        yc[static_cast<int64_t>(k1) * d.ostride] = out;
      }
      wv = cmul64(wv, ws);
    }
  } else {
    // columns out with the multi-pass twiddle and the scale
    float2* yc = y + lbase[1][sub][c];
    const float2 w0 = make_float2(static_cast<float>(w0d.x), static_cast<float>(w0d.y));
    const float2 wsc = scaled ? cscale<float>(w0, scale) : w0;
    if (d.psh) {
      __syncthreads();  // lsq
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const float2 w = u ? cmul<float>(wsc, lsq[sub * 256 + u * 16 + c]) : wsc;
        yc[static_cast<int64_t>(jt + R2 * u) * d.ostride] = cmul<float>(t[u], w);
      }
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u)
        yc[static_cast<int64_t>(jt + R2 * u) * d.ostride] = scaled ? cscale<float>(t[u], scale) : t[u];
    }
  }
}

template <int DIR, int N, int MODE>
hipError_t launch_pair_t(const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  hipLaunchKernelGGL((k_lean_pair<DIR, N, MODE>), dim3(d.groups), dim3(2 * N), 0, st, d,
                     static_cast<const float2*>(x), static_cast<float2*>(y), static_cast<const float2*>(tws));
  return hipGetLastError();
}

}  // namespace

hipError_t launch_lean_pair(int dir, const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t st) {
  if (d.groups <= 0) return hipSuccess;
  if (d.real == 3 && dir < 0) {
    if (d.n == 64) return launch_pair_t<-1, 64, 3>(d, x, y, tws, st);
    if (d.n == 128) return launch_pair_t<-1, 128, 3>(d, x, y, tws, st);
  } else if (d.real == 4 && dir > 0) {
    if (d.n == 64) return launch_pair_t<1, 64, 4>(d, x, y, tws, st);
    if (d.n == 128) return launch_pair_t<1, 128, 4>(d, x, y, tws, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace fftjh
