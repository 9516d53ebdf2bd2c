// Odd non-smooth lengths by Gaussian resampling onto a smooth t > n (the
// reference's HvdhPlan, hvdh.jl:1-19 after Harvey & van der Hoeven), FP32:
//
//   permute  xp[Delta l mod n] = x_l, Delta = t - n  (hvdh.jl:266-333)  -> k_hvdh_permute
//   spread   v[q] = sum_o W_o xp[(q + (o-1) Delta) mod n], o = 1-h .. h,
//            W_o = exp(-b (delta - o)^2)/alpha, delta = (t - q)/t,
//            b = pi/alpha^2  (hvdh.jl:355-470)                        -> k_hvdh_spread
//   Y      = DFT_t(v)                                                -> the pass kernels
//   solve    (n/t) Y[q] = sum_l exp(-a (rho l - q)^2) x_l,  x_l = X_(Delta l mod n),
//            on the row q nearest each unknown (the others identities), a
//            banded system indexed by q (entries exp(-a (d + beta_j)^2),
//            |d| <= B = 2), solved by LU without pivoting in chunks of L core
//            rows with H halo rows per side (hvdh.jl:474-686)        -> k_hvdh_solve
//   permute  y[Delta l mod n] = scale x_l                             -> k_hvdh_permute
//
// The reference solves G*K chunks of about 256 rows with SIMD lanes across
// chunks; here one lane solves one chunk of L = 16 core rows, so that a
// transform of 2^20 rows has 65,536 independent lanes and the rows of the
// back substitution stay in registers (the sweep is a sequential recurrence
// of ~40 instructions per row).  The inverse decays: core rows are exact to
// rounding when H is the probed halo (hvdh.jl:118-145, done by the planner).
//
// Measured on one MI355X (F32, x 4, forward): 1,000,003 takes 213 us (the
// launches 26 / 41 / 59 for the three grid passes / 73 / 27 us), 65,537
// 59 us, against 115 and 34 us by resampling; the planner keeps resampling
// and the route runs when asked for.  Gathering the spread's taps by k
// instead of through the permuted input took 374 us for the spread alone.
#include <algorithm>

#include "butterflies.h"
#include "common.h"
#include "launch.h"

namespace fftjh {

namespace {

// dst[Delta*l mod n] = scale * src[l] for the `batch` rows of n (the
// reference's permuted scatter, hvdh.jl:266-333): with c = Delta^-1 mod n,
// r = i*Delta + j (j < Delta, i < I = n / Delta) is Delta*l mod n for
// l = i + j*c mod n, so a tile of 16 i x 64 j reads 64 runs of x along i
// and writes 16 runs along j, transposed in LDS; the rows r >= I*Delta take
// l = r*c mod n singly.  Both the spread's input (x into Delta-permuted
// order) and the solve's output (unknowns x_l to y) are this permutation.
// src is rows of n; dst element r of transform b is at (b % o_bi)*o_in +
// (b / o_bi)*o_out + r*o_stride, so the solve's output can land at a stride
// (the length-p transforms of a split length, each at stride s).
constexpr int kPermI = 16, kPermJ = 64;
__global__ void __launch_bounds__(256) k_hvdh_permute(HvdhDesc d, const float2* __restrict__ src, float2* dst,
                                                      float scale) {
  __shared__ float2 tile[kPermJ][kPermI + 1];
  const int64_t n = d.n, D = d.delta, cinv = d.cinv;
  const int64_t I = n / D;
  const int64_t ti = (I + kPermI - 1) / kPermI, tj = (D + kPermJ - 1) / kPermJ;
  const int64_t tiles = ti * tj, extra = (n - I * D + 255) / 256;
  const int64_t per = tiles + extra;
  const int64_t b = blockIdx.x / per, blk = blockIdx.x - b * per;
  const float2* sb = src + b * n;
  const int64_t bi = b % d.o_bi;
  float2* db = dst + bi * d.o_in + (b - bi) / d.o_bi * d.o_out;
  const int64_t os = d.o_stride;
  const int tid = static_cast<int>(threadIdx.x);
  if (blk >= tiles) {
    const int64_t r = I * D + (blk - tiles) * 256 + tid;
    if (r < n) {
      const float2 v = sb[(r * cinv) % n];
      db[r * os] = make_float2(v.x * scale, v.y * scale);
    }
    return;
  }
  const int64_t ib = (blk / tj) * kPermI, jb = (blk - (blk / tj) * tj) * kPermJ;
#pragma unroll
  for (int e = 0; e < kPermI * kPermJ / 256; ++e) {
    const int idx = tid + 256 * e;
    const int ii = idx % kPermI, jj = idx / kPermI;
    const int64_t i = ib + ii, j = jb + jj;
    if (i < I && j < D) {
      int64_t l = i + (j * cinv) % n;
      if (l >= n) l -= n;
      tile[jj][ii] = sb[l];
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kPermI * kPermJ / 256; ++e) {
    const int idx = tid + 256 * e;
    const int jj = idx % kPermJ, ii = idx / kPermJ;
    const int64_t i = ib + ii, j = jb + jj;
    if (i < I && j < D) {
      const float2 v = tile[jj][ii];
      db[(i * D + j) * os] = make_float2(v.x * scale, v.y * scale);
    }
  }
}

// v[q] = sum_o W_o xp[(q + (o-1) Delta) mod n] over the Delta-permuted input
// xp: grid point q is output k = Delta^-1 q mod t, at k n/t = s + delta with
// delta = (t - q)/t in (0, 1], and its tap at grid s + o is x_l with
// Delta l = q + (o - 1) Delta (mod n) (hvdh.jl:266-272).  Reads and writes
// run along q.  W_o = exp(-b (delta - o)^2)/alpha = E G^o exp(-b o^2),
// E = exp(-b delta^2)/alpha, G = exp(2 b delta), in FP64 (hvdh.jl:360-380).
__global__ void __launch_bounds__(256) k_hvdh_spread(HvdhDesc d, const float2* __restrict__ xp, float2* v) {
  const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (idx >= d.batch * d.t) return;
  const int64_t b = idx / d.t, q = idx - b * d.t;
  const int64_t n = d.n, t = d.t;
  const float2* xb = xp + b * n;
  const double dl = static_cast<double>(t - q) / static_cast<double>(t);
  const double E = exp(-d.b * dl * dl) * d.ialpha;
  const double Gp = exp(2.0 * d.b * dl), Gm = 1.0 / Gp;
  const int64_t m0 = q >= n ? q - n : q;
  float ar = 0.f, ai = 0.f;
  double P = E;
  for (int o = 0; o >= 1 - d.h; --o) {  // o = 0, -1, ...: weights E G^o
    if (o < 0) P *= Gm;
    const float W = static_cast<float>(P * d.kc[-o]);
    int64_t m = m0 + d.tap[o + d.h];
    if (m >= n) m -= n;
    const float2 a = xb[m];
    ar = fmaf(W, a.x, ar);
    ai = fmaf(W, a.y, ai);
  }
  P = E;
  for (int o = 1; o <= d.h; ++o) {
    P *= Gp;
    const float W = static_cast<float>(P * d.kc[o]);
    int64_t m = m0 + d.tap[o + d.h];
    if (m >= n) m -= n;
    const float2 a = xb[m];
    ar = fmaf(W, a.x, ar);
    ai = fmaf(W, a.y, ai);
  }
  v[b * t + q] = make_float2(ar, ai);
}

// One lane per chunk of L core rows (rows q0 .. q0 + R - 1, q0 = c*L - H,
// R = L + 2H <= RMAX).  Index state of row q: e = 2t l - 2n q + n in [0, 2t)
// for the first unknown l with rho l >= q - 1/2 (hvdh.jl:147-151); the row
// holds x_l iff e < 2n, with beta = (e - n)/(2n); the next row is e - 2n
// (plus 2t, and l + 1, when that is negative).
template <int L, int HMAX>
__global__ void __launch_bounds__(64) k_hvdh_solve(HvdhDesc d, const float2* __restrict__ Y, float2* y) {
  constexpr int RMAX = L + 2 * HMAX;
  // the wave's 64 chunks read rows c0*L - H .. (c0+64)*L + H of one
  // transform: staged through LDS with coalesced loads (the lanes' own rows
  // are 16 apart, so direct loads would touch 64 lines per instruction)
  constexpr int SPAN = 64 * L + 2 * HMAX + 8;
  __shared__ float2 ys[SPAN];
  // the unknowns of the wave's core rows are consecutive, l_base .. l_base
  // + cnt - 1 (one per occupied row, at most one per row): gathered here and
  // stored in l order
  __shared__ float2 xs[64 * L];
  const int64_t per = (d.chunks + 63) / 64;  // waves per transform
  const int64_t wv = blockIdx.x, bw = wv / per, cw0 = (wv - bw * per) * 64;
  const int64_t c = cw0 + threadIdx.x, b = bw;
  {
    const int64_t qs = cw0 * L - d.H;
    const int span = 64 * L + 2 * d.H + 4;
    const float2* Ybw = Y + bw * d.t;
    for (int i = threadIdx.x; i < span; i += 64) {
      int64_t q = qs + i;
      if (q < 0) q += d.t;
      if (q >= d.t) q -= d.t;
      ys[i] = Ybw[q];
    }
  }
  __syncthreads();
  const int64_t n = d.n, t = d.t, n2 = 2 * n, t2 = 2 * t;
  // first unknown at or after row q: ceil((2n q - n) / (2t)) (hvdh.jl:147-151)
  auto first_l = [&](int64_t q) {
    const int64_t nm = n2 * q - n;
    return nm >= 0 ? (nm + t2 - 1) / t2 : -((-nm) / t2);
  };
  const int64_t l_base = first_l(cw0 * L);
  const int64_t q_end = std::min<int64_t>((cw0 + 64) * L, t);
  const int64_t cnt = first_l(q_end) - l_base;
  if (c < d.chunks) {
  const int H = d.H, R = L + 2 * H;
  const int64_t q0 = c * L - H;
  // index state at row q0 (the division once per lane; rows step after)
  int64_t l = first_l(q0);
  int64_t e = t2 * l - (n2 * q0 - n);
  const float a = d.a, nt = d.nt, c1 = d.c1, c2 = d.c2;
  // per-row values of the five rows q .. q + 4 ahead of the pivot:
  // Ap = exp(-a beta^2) (0 for an empty row), g = exp(-2 a beta), 1/g, rhs
  float wAp[5], wg[5], wgi[5], wr[5], wi[5];
  bool wocc[5];
  int64_t e_last = 0, l_last = 0;
  int j_next = 0;  // next row index the iterator produces
  auto advance = [&](int slot) {
    const int j = j_next++;
    const bool inside = j < R;
    const bool occ = inside && e < n2;
    if (j == R - 1) {
      e_last = e;
      l_last = l;
    }
    float Ap = 0.f, g = 1.f, gi = 1.f, rr = 0.f, ri = 0.f;
    if (occ) {
      const float beta = static_cast<float>(e - n) / static_cast<float>(n2);
      Ap = expf(-a * beta * beta);
      g = expf(-2.f * a * beta);
      gi = 1.f / g;
      // row q0 + j of the staged span (q0 = c*L - H, the span starts at
      // cw0*L - H)
      const float2 v = ys[(c - cw0) * L + j];
      rr = nt * v.x;
      ri = nt * v.y;
    }
    wAp[slot] = Ap;
    wg[slot] = g;
    wgi[slot] = gi;
    wr[slot] = rr;
    wi[slot] = ri;
    wocc[slot] = occ;
    // step to the next row
    e -= n2;
    if (e < 0) {
      e += t2;
      ++l;
    }
  };
  // entry of column j (window slot sj) in row i = j - m
  auto entry = [&](int sj, int m) -> float {
    const float A = wAp[sj];
    if (m == 0) return A;
    if (m == 1) return A * wg[sj] * c1;
    if (m == 2) return A * wg[sj] * wg[sj] * c2;
    if (m == -1) return A * wgi[sj] * c1;
    return A * wgi[sj] * wgi[sj] * c2;
  };
  // rows 0 and 1 need columns 0..3
#pragma unroll
  for (int s0 = 0; s0 < 4; ++s0) advance(s0);
  // pivot row k: columns k..k+2 (+ rhs); row k+1: columns k..k+3 (+ rhs)
  float p0, p1, p2, pr, pi;
  float r0, r1, r2, r3, rr1, ri1;
  {
    // row 0 (slot 0): columns 0, 1, 2 (slots 0, 1, 2)
    if (wocc[0]) {
      p0 = entry(0, 0);
      p1 = entry(1, 1);
      p2 = entry(2, 2);
      pr = wr[0];
      pi = wi[0];
    } else {
      p0 = 1.f;
      p1 = p2 = pr = pi = 0.f;
    }
    // row 1 (slot 1): columns 0..3
    if (wocc[1]) {
      r0 = entry(0, -1);
      r1 = entry(1, 0);
      r2 = entry(2, 1);
      r3 = entry(3, 2);
      rr1 = wr[1];
      ri1 = wi[1];
    } else {
      r1 = 1.f;
      r0 = r2 = r3 = rr1 = ri1 = 0.f;
    }
    if (R < 2) r0 = r1 = r2 = r3 = rr1 = ri1 = 0.f;
  }
  float U1[RMAX], U2[RMAX], Xr[RMAX], Xi[RMAX];
#pragma unroll
  for (int k = 0; k < RMAX; ++k) {
    if (k < R) {
      // row k + 2 from columns k..k+4 (slots (k .. k+4) mod 5); the
      // iterator first produces row k + 4
      advance((k + 4) % 5);
      float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f, sr = 0.f, si = 0.f;
      const int sm = (k + 2) % 5;
      if (k + 2 < R) {
        if (wocc[sm]) {
          s0 = entry(k % 5, -2);
          s1 = entry((k + 1) % 5, -1);
          s2 = entry(sm, 0);
          s3 = entry((k + 3) % 5, 1);
          s4 = entry((k + 4) % 5, 2);
          sr = wr[sm];
          si = wi[sm];
        } else {
          s2 = 1.f;
        }
      }
      // normalised pivot row k
      const float inv = 1.f / p0;
      const float u1 = p1 * inv, u2 = p2 * inv, xr = pr * inv, xi = pi * inv;
      U1[k] = u1;
      U2[k] = u2;
      Xr[k] = xr;
      Xi[k] = xi;
      // eliminate column k from rows k+1 and k+2
      const float f1 = r0;
      const float n0 = r1 - f1 * u1, n1 = r2 - f1 * u2, n2v = r3;
      const float nr = rr1 - f1 * xr, ni = ri1 - f1 * xi;
      const float f2 = s0;
      const float m0 = s1 - f2 * u1, m1 = s2 - f2 * u2, m2 = s3, m3 = s4;
      const float mr = sr - f2 * xr, mi = si - f2 * xi;
      p0 = n0;
      p1 = n1;
      p2 = n2v;
      pr = nr;
      pi = ni;
      r0 = m0;
      r1 = m1;
      r2 = m2;
      r3 = m3;
      rr1 = mr;
      ri1 = mi;
    }
  }
  // back substitution from row R-1, the index state stepping backwards
  float x1r = 0.f, x1i = 0.f, x2r = 0.f, x2i = 0.f;  // rows k+1, k+2
  int64_t eb = e_last, lb = l_last;
#pragma unroll
  for (int k = RMAX - 1; k >= 0; --k) {
    if (k < R) {
      const float xr = Xr[k] - U1[k] * x1r - U2[k] * x2r;
      const float xi = Xi[k] - U1[k] * x1i - U2[k] * x2i;
      if (k >= H && k < H + L && q0 + k < t && eb < n2) xs[lb - l_base] = make_float2(xr, xi);
      x2r = x1r;
      x2i = x1i;
      x1r = xr;
      x1i = xi;
      eb += n2;
      if (eb >= t2) {
        eb -= t2;
        --lb;
      }
    }
  }
  }
  __syncthreads();
  float2* yb = y + b * n + l_base;
  for (int i = threadIdx.x; i < cnt; i += 64) yb[i] = xs[i];
}

}  // namespace

int hvdh_max_halo() { return 16; }
int hvdh_chunk_rows() { return 16; }

hipError_t launch_hvdh_permute(const HvdhDesc& d, const void* src, void* dst, float scale, hipStream_t st) {
  const int64_t I = d.n / d.delta;
  const int64_t tiles = ((I + kPermI - 1) / kPermI) * ((d.delta + kPermJ - 1) / kPermJ);
  const int64_t extra = (d.n - I * d.delta + 255) / 256;
  hipLaunchKernelGGL(k_hvdh_permute, dim3(static_cast<unsigned>(d.batch * (tiles + extra))), dim3(256), 0, st, d,
                     static_cast<const float2*>(src), static_cast<float2*>(dst), scale);
  return hipGetLastError();
}

hipError_t launch_hvdh_spread(const HvdhDesc& d, const void* xp, void* v, hipStream_t st) {
  const int64_t threads = d.batch * d.t;
  hipLaunchKernelGGL(k_hvdh_spread, dim3(static_cast<unsigned>((threads + 255) / 256)), dim3(256), 0, st, d,
                     static_cast<const float2*>(xp), static_cast<float2*>(v));
  return hipGetLastError();
}

hipError_t launch_hvdh_solve(const HvdhDesc& d, const void* Yg, void* y, hipStream_t st) {
  if (d.H > 16 || d.B != 2) return hipErrorInvalidValue;
  const int64_t waves = d.batch * ((d.chunks + 63) / 64);
  hipLaunchKernelGGL((k_hvdh_solve<16, 16>), dim3(static_cast<unsigned>(waves)), dim3(64), 0, st, d,
                     static_cast<const float2*>(Yg), static_cast<float2*>(y));
  return hipGetLastError();
}

}  // namespace fftjh
