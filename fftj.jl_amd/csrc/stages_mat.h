// Split-precision DFT-matrix stages on CDNA4 MFMA (FP32 plans only).
//
// A radix-R Stockham stage is, per butterfly, y = F_R * a with a the R
// twiddled inputs.  Written over reals with the inputs interleaved
// (re0, im0, re1, im1, ...), Re(y) = [Re F, -Im F] a and Im(y) = [Im F, Re F] a:
// two real 16 x 32 matrices per K-tile of 16 points, 16 butterflies per
// v_mfma_f32_16x16x32_{f16,bf16} (gfx950), FP32 accumulation.  R = 9 pads the
// tile with zeros; R = 32 takes 2 x 2 tiles.
//
// Operands are split so that FP32-level accuracy survives the 16-bit inputs
// (the split-operand compensation of Markidis et al. 2018 and Ootomo &
// Yokota 2022, cited as methods):
//   kMatF16    x = hi + lo in FP16 after a per-butterfly power-of-two scale
//              (FP16's exponent range is narrow); products hi*hi + hi*lo + lo*hi;
//   kMatBF16   x = hi + lo in BF16 (16 significant bits; the same 3 products);
//   kMatBF16x3 x = hi + mid + lo in BF16 (24 bits; 6 products, the terms
//              below 2^-24 dropped).
// The matrix operands come split the same way from a table generated in FP64
// (twiddle.hip, k_mat_table); the stage twiddles are applied in FP32 first.
//
// Lane layout of the 16x16x32 instruction (checked on the device): lane l
// feeds A[l % 16][8*(l/16) + e] and B[8*(l/16) + e][l % 16], e < 8, and gets
// D[4*(l/16) + i][l % 16], i < 4.  So lane l holds points 4*(l/16)..+3 of
// butterfly column l % 16 and receives outputs 4*(l/16)..+3 of that column.
#pragma once
#include "butterflies.h"
#include "common.h"
#include "stages.h"

namespace fftjh {

enum MatMode { kMatNone = 0, kMatF16 = 1, kMatBF16 = 2, kMatBF16x3 = 3 };

__host__ __device__ constexpr int mat_terms(int mode) { return mode == kMatBF16x3 ? 3 : 2; }
__host__ __device__ constexpr bool mat_radix(int R) { return R == 9 || R == 16 || R == 32; }
__host__ __device__ constexpr int mat_kt(int R) { return R > 16 ? 2 : 1; }
// tiles of 16 butterflies one wave holds at once (16 complex values per lane
// for R = 16 and 32; R = 9 fills only 9 of 16 rows, so twice as many tiles)
__host__ __device__ constexpr int mat_tiles_per_wave(int R) { return R == 32 ? 2 : 4; }

// Operand table of one mode, in 16-byte entries (8 values): radix R's block
// starts at mat_base(R, terms); entry (((h*KT + t)*2 + comp)*terms + term)*64
// + lane holds what `lane` feeds as A for output tile h, K-tile t, component
// comp (0 = real part, 1 = imaginary part) and split term (0 = hi).
__host__ __device__ constexpr int mat_base(int R, int terms) { return R == 9 ? 0 : R == 16 ? 128 * terms : 256 * terms; }
__host__ __device__ constexpr int mat_entries(int terms) { return 768 * terms; }

typedef _Float16 mat_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 mat_b8 __attribute__((ext_vector_type(8)));
typedef float mat_f4 __attribute__((ext_vector_type(4)));

template <int MODE> struct MatOps;

template <> struct MatOps<kMatF16> {
  using V = mat_h8;
  static constexpr int terms = 2;
  static constexpr bool scaled = true;
  __device__ static void split(const float (&x)[8], V (&b)[terms]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const _Float16 h = static_cast<_Float16>(x[e]);
      b[0][e] = h;
      b[1][e] = static_cast<_Float16>(x[e] - static_cast<float>(h));
    }
  }
  __device__ static mat_f4 mma(V a, V b, mat_f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

template <int MODE> struct MatBf16 {
  using V = mat_b8;
  static constexpr int terms = mat_terms(MODE);
  static constexpr bool scaled = false;
  __device__ static void split(const float (&x)[8], V (&b)[terms]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float r = x[e];
#pragma unroll
      for (int t = 0; t < terms; ++t) {
        const __bf16 h = static_cast<__bf16>(r);
        b[t][e] = h;
        r -= static_cast<float>(h);  // exact: h is r rounded to 8 bits
      }
    }
  }
  __device__ static mat_f4 mma(V a, V b, mat_f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <> struct MatOps<kMatBF16> : MatBf16<kMatBF16> {};
template <> struct MatOps<kMatBF16x3> : MatBf16<kMatBF16x3> {};

// zr += Ar*B, zi += Ai*B over the split terms, smallest products first; the
// two components alternate so that consecutive MFMAs are independent
template <typename Ops>
__device__ __forceinline__ void mat_dot2(const typename Ops::V (&ar)[Ops::terms], const typename Ops::V (&ai)[Ops::terms],
                                         const typename Ops::V (&b)[Ops::terms], mat_f4& zr, mat_f4& zi) {
  if constexpr (Ops::terms == 2) {
    constexpr int ia[3] = {1, 0, 0}, ib[3] = {0, 1, 0};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      zr = Ops::mma(ar[ia[q]], b[ib[q]], zr);
      zi = Ops::mma(ai[ia[q]], b[ib[q]], zi);
    }
  } else {
    constexpr int ia[6] = {2, 0, 1, 1, 0, 0}, ib[6] = {0, 2, 1, 0, 1, 0};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      zr = Ops::mma(ar[ia[q]], b[ib[q]], zr);
      zi = Ops::mma(ai[ia[q]], b[ib[q]], zi);
    }
  }
}

// One radix-R stage (R in {9, 16, 32}) over the C columns in LDS, the same
// Stockham step as run_stage: butterfly beta = j*C + c of the workgroup is
// column beta % 16 of tile beta / 16; wave w takes tiles w, w + nwaves, ...
// All 64 lanes run every MFMA (inactive columns carry zeros).
template <int DIR, int MODE, int R>
__device__ __noinline__ void run_stage_mat(const int n, const int C, const int logC, const StageDesc s,
                                           LdsPtr<float> lds, TwPtr<float> tw, const void* mat, const int lane,
                                           const int wave, const int nwaves) {
  using Ops = MatOps<MODE>;
  using V = typename Ops::V;
  constexpr int KT = mat_kt(R);
  constexpr int P = 4 * KT;
  constexpr int TPW = mat_tiles_per_wave(R);
  constexpr int T = Ops::terms;
  const int nr = n / R;
  const int ns = s.ns;
  const int nb = nr * C;
  const int g = lane >> 4, col = lane & 15;
  const __attribute__((address_space(1))) V* A = (const __attribute__((address_space(1))) V*)(mat) + mat_base(R, T);
  // R <= 16: one K-tile, both components' operands held for the whole stage
  // (their loads overlap the gather and the barrier)
  V a1[2][T];
  if constexpr (KT == 1) {
#pragma unroll
    for (int comp = 0; comp < 2; ++comp)
#pragma unroll
      for (int t = 0; t < T; ++t) a1[comp][t] = A[(comp * T + t) * 64 + lane];
  }
  V2<float> v[TPW][P];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int beta = (wave + u * nwaves) * 16 + col;
#pragma unroll
    for (int q = 0; q < P; ++q) v[u][q] = mk<float>(0.f, 0.f);
    if (beta < nb) {
      const int c = beta & (C - 1), j = beta >> logC;
      const int k = ns > 1 ? j - static_cast<int>(__umulhi(static_cast<unsigned>(j), s.ns_magic)) * ns : 0;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int r = 16 * (q >> 2) + 4 * g + (q & 3);
        if (r < R) {
          V2<float> a = ldv<float>(lds, lpad((j + r * nr) * C + c));
          if (r > 0 && ns > 1) a = cmul<float>(a, ldv<float>(tw, s.tw_off + (r - 1) * ns + k));
          v[u][q] = a;
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int tile = wave + u * nwaves;
    if (tile * 16 >= nb) break;  // uniform across the wave
    int ex = 0;
    if constexpr (Ops::scaled) {
      float mx = 0.f;
#pragma unroll
      for (int q = 0; q < P; ++q) mx = fmaxf(mx, fmaxf(fabsf(v[u][q].x), fabsf(v[u][q].y)));
      mx = fmaxf(mx, __shfl_xor(mx, 16));
      mx = fmaxf(mx, __shfl_xor(mx, 32));
      if (mx > 0.f) (void)frexpf(mx, &ex);  // mx = f * 2^ex, f in [0.5, 1)
    }
    V b[KT][T];
#pragma unroll
    for (int t = 0; t < KT; ++t) {
      float xs[8];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        xs[2 * e] = Ops::scaled ? ldexpf(v[u][4 * t + e].x, -ex) : v[u][4 * t + e].x;
        xs[2 * e + 1] = Ops::scaled ? ldexpf(v[u][4 * t + e].y, -ex) : v[u][4 * t + e].y;
      }
      Ops::split(xs, b[t]);
    }
    const int beta = tile * 16 + col;
    const int c = beta & (C - 1), j = beta >> logC;
    const int k = ns > 1 ? j - static_cast<int>(__umulhi(static_cast<unsigned>(j), s.ns_magic)) * ns : 0;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int h = 0; h < KT; ++h) {
      mat_f4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
      if constexpr (KT == 1) {
        mat_dot2<Ops>(a1[0], a1[1], b[0], acc[0], acc[1]);
      } else {
        // R = 32: the operands of one output tile at a time are read (L1
        // hits) -- all four blocks live would not fit the register budget
        asm volatile("" ::: "memory");
#pragma unroll
        for (int t = 0; t < KT; ++t) {
          V ar[T], ai[T];
#pragma unroll
          for (int tt = 0; tt < T; ++tt) {
            ar[tt] = A[(((h * KT + t) * 2 + 0) * T + tt) * 64 + lane];
            ai[tt] = A[(((h * KT + t) * 2 + 1) * T + tt) * 64 + lane];
          }
          mat_dot2<Ops>(ar, ai, b[t], acc[0], acc[1]);
        }
      }
      if (beta < nb) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int m = 16 * h + 4 * g + i;
          if (m < R) {
            const float re = Ops::scaled ? ldexpf(acc[0][i], ex) : acc[0][i];
            const float im = Ops::scaled ? ldexpf(acc[1][i], ex) : acc[1][i];
            stv(lds, lpad((base + m * ns) * C + c), mk<float>(re, im));
          }
        }
      }
    }
  }
  __syncthreads();
}

}  // namespace fftjh
