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

// acc += A*B over the split terms, smallest products first
template <typename Ops>
__device__ __forceinline__ mat_f4 mat_dot(const typename Ops::V (&a)[Ops::terms], const typename Ops::V (&b)[Ops::terms],
                                          mat_f4 acc) {
  if constexpr (Ops::terms == 2) {
    acc = Ops::mma(a[1], b[0], acc);
    acc = Ops::mma(a[0], b[1], acc);
    acc = Ops::mma(a[0], b[0], acc);
  } else {
    acc = Ops::mma(a[2], b[0], acc);
    acc = Ops::mma(a[0], b[2], acc);
    acc = Ops::mma(a[1], b[1], acc);
    acc = Ops::mma(a[1], b[0], acc);
    acc = Ops::mma(a[0], b[1], acc);
    acc = Ops::mma(a[0], b[0], acc);
  }
  return acc;
}

// One radix-R stage (R in {9, 16, 32}) over the C columns in LDS, the same
// Stockham step as run_stage: butterfly beta = j*C + c of the workgroup is
// column beta % 16 of tile beta / 16; wave w takes tiles w, w + nwaves, ...
// All 64 lanes run every MFMA (inactive columns carry zeros).
template <int DIR, int MODE, int R>
__device__ __noinline__ void run_stage_mat(const int n, const int C, const int logC, const StageDesc& s,
                                           V2<float>* lds, const V2<float>* tw, const void* mat, const int lane,
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
          V2<float> a = lds[lpad((j + r * nr) * C + c)];
          if (r > 0 && ns > 1) a = cmul<float>(a, tw[s.tw_off + (r - 1) * ns + k]);
          v[u][q] = a;
        }
      }
    }
  }
  __syncthreads();
  const V* A = static_cast<const V*>(mat) + mat_base(R, T);
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int tile = wave + u * nwaves;
    if (tile * 16 >= nb) break;  // uniform across the wave
    // the operands are read again for every tile (L1 hits): kept live across
    // tiles they would not fit the register budget of a 1024-thread group
    asm volatile("" ::: "memory");
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
      mat_f4 acc[2];
#pragma unroll
      for (int comp = 0; comp < 2; ++comp) {
        if constexpr (KT > 1) asm volatile("" ::: "memory");  // one (h, comp) block of operands live at a time
        mat_f4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < KT; ++t) {
          V at[T];
#pragma unroll
          for (int tt = 0; tt < T; ++tt) at[tt] = A[(((h * KT + t) * 2 + comp) * T + tt) * 64 + lane];
          z = mat_dot<Ops>(at, b[t], z);
        }
        acc[comp] = z;
      }
      if (beta < nb) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int m = 16 * h + 4 * g + i;
          if (m < R) {
            const float re = Ops::scaled ? ldexpf(acc[0][i], ex) : acc[0][i];
            const float im = Ops::scaled ? ldexpf(acc[1][i], ex) : acc[1][i];
            lds[lpad((base + m * ns) * C + c)] = mk<float>(re, im);
          }
        }
      }
    }
  }
  __syncthreads();
}

}  // namespace fftjh
