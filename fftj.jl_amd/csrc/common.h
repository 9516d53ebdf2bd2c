// Descriptors shared by the host planner and the HIP kernels.
//
// One "pass" is one kernel launch that runs whole length-n transforms, a group
// of C of them per workgroup, entirely in LDS/registers.  Every plan (direct,
// four-step, N-d axis passes, the m-length transform inside resampling, the
// half-length transform inside real plans) is a short list of passes.
#pragma once
#include <stdint.h>

namespace fftjh {

constexpr int kMaxStages = 16;
// LDS per workgroup on gfx950 (160 KiB, verified by compiling a 160 KiB array)
constexpr int kLdsBytes = 160 * 1024;
constexpr int kMaxThreads = 1024;
// complex elements held in registers per thread inside one stage
constexpr int kElemsPerThread = 16;

// One Stockham radix stage (autosort form, twiddles before the butterfly):
// for butterfly j < n/R with k = j mod ns,
//   a_r = in[j + r*(n/R)] * w^{r*k},   w = exp(sign*2*pi*i/(ns*R))
//   out[(j-k)*R + k + r*ns] = DFT_R(a)_r
// ns is the product of the radices of the earlier stages.
struct StageDesc {
  int32_t radix;
  int32_t ns;
  int32_t tw_off;     // complex offset of this stage's table: entry (r-1)*ns + k
  uint32_t ns_magic;  // floor(2^32/ns)+1: j/ns == umulhi(j, magic) for j*ns < 2^32
};

// element i of transform b lives at
//   (b % Bi)*s_in + ((b / Bi) % B2)*s_mid + (b / (Bi*B2))*s_out + i*stride
// (Bi and B2 belong to the pass; kernels outside the pass kernels use the
// two-level form (b % Bi)*s_in + (b / Bi)*s_out)
struct Geometry {
  int64_t stride;
  int64_t s_in;
  int64_t s_out;
  int64_t s_mid;
};

struct PassDesc {
  int32_t n;
  int32_t nstages;
  StageDesc st[kMaxStages];
  int64_t batch;     // transforms in the launch
  int64_t Bi;        // inner batch count of the three-level batch index
  int64_t B2;        // middle batch count
  int32_t C;         // transforms per workgroup (power of two)
  int32_t logC;
  int32_t nthreads;  // threads per workgroup (multiple of 64 and of C)
  int32_t ngroups;   // workgroups in the launch
  Geometry in;
  Geometry out;
  // multi-pass twiddle at the store: output k of transform b is multiplied
  // by w_P^e, e = q*(tw_acoef*((b / Bi) % B2) + tw_bcoef*k), q = (b % Bi) / tw_qdiv,
  // P = tw_period (0 = none); the exponent is reduced exactly in integers,
  // the value read from the tables below
  int64_t tw_period;
  int64_t tw_qdiv;
  int64_t tw_acoef;
  int64_t tw_bcoef;
  double scale;      // applied at the store
  int32_t xcd_remap;  // map consecutive logical groups onto one XCD
  int32_t grid;       // workgroups launched (persistent: each walks several groups)
  int32_t rows_in;    // power-of-two kernel: stage HBM rows through LDS on the way in
  int32_t rows_out;   // ... and on the way out (row-major lanes, contiguous runs)
  // with rows_in / rows_out: the C transforms of a group start at
  // base(g*C) + c*cdelta (the planner checks that this holds)
  int64_t in_cdelta;
  int64_t out_cdelta;
  // batch < 2^31: the batch split in 32-bit arithmetic, x / D = (x * m) >> s
  // (FastDiv, passcommon.h) for D = Bi, B2
  int32_t fast_batch;
  uint32_t bi_m, b2_m;
  int32_t bi_s, b2_s;
  // multi-pass twiddle start values from two FP64 tables (double2):
  // w_P^e = ptw[e >> ptw_sh] * ptw[ptw_lo + (e & (2^ptw_sh - 1))]
  int32_t ptw_sh;
  int64_t ptw_lo;
  const void* ptw;
  uint32_t qd_m;     // FastDiv by tw_qdiv
  int32_t qd_s;
  int64_t tw_ar;     // tw_acoef mod P
  int64_t tw_br;     // tw_bcoef mod P
  double tw_invp;    // 1 / P
  // power-of-two row passes fused with the real edges of an even-length
  // real transform of n = 2*N reals (N the pass length): 1 = R2C (twist of
  // the half-length spectrum into N+1 bins at the store), 2 = C2R (pre-pass
  // of N+1 bins into the half-length input at the load); rtw holds
  // w_k = exp(dir*2*pi*i*k/n) as an FP64 two-level table:
  // w = rtw[e >> rtw_sh] * rtw[rtw_hi + (e & (2^rtw_sh - 1))]
  int32_t real_io;
  const void* rtw;
  // real_io 3 / 4: the last (R2C) or first (C2R) pass of a multi-pass
  // half-length transform of length real_h (n = 2*real_h), with pair_M
  // transforms per row; a group holds 16 transforms T and their 16 mirrors
  // pair_M - T
  int32_t pair_M;
  int64_t real_h;
  int32_t rtw_sh;
  int64_t rtw_hi;
  // rs_io = 1: the first pass of a resampled transform reads its m-grid
  // straight from contiguous rows of rs_n inputs: g[j] = x[j]*corr[j] for
  // j < hp = (rs_n+1)/2, x[j-(rs_m-rs_n)]*corr[j-(rs_m-rs_n)] for
  // j >= rs_m - rs_n + hp, 0 between (the deconvolution, fused)
  int32_t rs_io;
  int64_t rs_n, rs_m;
  const void* rs_corr;
  // generic (non-power-of-two) pass: first stage fused into the HBM load,
  // last into the store (chains of two or more stages)
  int32_t fuse_io;
  // streaming hints for data the plan does not touch again: the pass's HBM
  // loads (its input is read once) / stores (its output is the plan's result)
  int32_t nt_load;
  int32_t nt_store;
};

// One launch of the lean column-pass kernel (stockham_lean.hip): `groups`
// workgroups, each 16 adjacent transforms of n = 64, 128 or 256 points.
// Group g has digits d0 = g % D0, d1 = (g / D0) % D1, d2 = g / (D0*D1); its
// input starts at d0*in_s0 + d1*in_s1 + d2*in_s2, column c at +c, element m
// at +m*istride.  Output: rows == 0 the same form with out_s* and ostride;
// rows == 1 column c at out_base + c*ostride, elements contiguous.
// Multi-pass twiddle w_P^e, P = 2^psh (psh = 0: none),
// e = d1*(tw_a0*(16*d0 + c) + tw_b*k) mod P, from the FP64 tables ptw
// (w_P^(i*2^ptw_sh) for i < ptw_lo, then w_P^i).
struct LeanDesc {
  int32_t n;
  int32_t rows;     // outputs as contiguous rows (an LDS transposition)
  int32_t groups;
  int32_t D0, D1;
  int32_t rin;      // inputs as contiguous rows at stride istride (an LDS transposition)
  int32_t ncol;     // columns per block (column form): group columns 16*d0 + c >= ncol are
                    // absent (no load, no store); 0: every group is whole
  int32_t xcd;      // 1: XCD-local group order (workgroup b runs group start(b % 8) + b / 8)
  int64_t in_s0, in_s1, in_s2, istride;
  int64_t out_s0, out_s1, out_s2, ostride;
  int32_t psh;
  int32_t ptw_sh;
  int64_t ptw_lo;
  int64_t tw_a0, tw_b;
  const void* ptw;
  double scale;
  int32_t mat;      // split-precision MFMA mode of the radix-16 stages (kMatNone: VALU)
  int32_t nt;       // bit 0: non-temporal loads of x, bit 1: non-temporal stores of y
  const void* matp; // that mode's operand table for the pass's direction
  // square lengths N = R*R (k_lean_sq): columns per group, and the twiddle
  // period when it is not a power of two (0: P = 2^psh) with 1/P
  int32_t cols;
  // rs = 1 (rows-out passes): the pass reads the m-point grid of a resampled
  // length straight from rows of rs_n inputs through the deconvolution,
  // g[j] = x[j]*corr[j] (j < (rs_n+1)/2), x[j-(rs_m-rs_n)]*corr[j-(rs_m-rs_n)]
  // (j >= rs_m - rs_n + (rs_n+1)/2), 0 between; transform d2 at x + d2*rs_n
  int32_t rs;
  int64_t pmod;
  // twq = 1 (columns out): the multi-pass twiddle's multiplier is the column,
  // e = (16*d0 + c)*(tw_a0*d1 + tw_b*k) mod 2^psh (the generic passes' middle
  // form, the column's digit times the group's)
  int32_t twq;
  // real = 1 / 2 (rows in and out, half length N): the R2C twist at the
  // rows-out store (N + 1 bins per row) / the C2R pre-pass at the rows-in
  // load (N + 1 bins per row in); rtw = w_k = exp(dir*2*pi*i*k/(2N)), k <= N
  int32_t real;
  const void* rtw;
  double pinv;
  int64_t rs_n, rs_m;
  const void* rs_corr;
  // real = 3 / 4 (k_lean_pair): the paired edges of a multi-pass
  // half-length transform of real_h = n*pm points per row, pm transforms T
  // per row; transform T of row r at (T % pbi)*s0 + ((T / pbi) % pb2)*s1 +
  // r*s2 on each side (pbi*pb2 = pm), elements at istride / ostride; rtw the
  // FP64 two-level table of w_{2h} (rtw_sh, rtw_hi as the pass kernel's);
  // the C2R pass's twiddle multiplier q = (T % pbi) / pqdiv
  int32_t pm;
  int32_t rtw_sh;
  int64_t pbi, pb2, pqdiv, real_h, rtw_hi;
};

// HvdH route (hvdh.hip) of `batch` transforms of odd non-smooth n, FP32:
// the input permuted to Delta order (cinv = Delta^-1 mod n), spread onto the
// smooth grid t (Delta = t - n; tap o reads permuted element
// (q + tap[o + h]) mod n, tap[o + h] = (o - 1)*Delta mod n), the grid
// transform, `chunks` banded solves per transform of 16 core rows with H halo
// rows per side, and the unknowns permuted back into y.
struct HvdhDesc {
  int64_t n, t, delta, cinv;
  int64_t batch, chunks;
  int32_t h;         // spread taps o = 1-h .. h
  int32_t B;         // band half width of the solve (2)
  int32_t H;         // halo rows per side
  int32_t pad_;
  int64_t tap[16];   // (o - 1)*Delta mod n for o = 1-h .. h, at o + h
  double b, ialpha;  // spread weights exp(-b (delta - o)^2) / alpha
  double kc[8];      // exp(-b o^2), o = 0..h
  float a;           // solve entries exp(-a (d + beta)^2)
  float nt;          // n / t
  float c1, c2;      // exp(-a), exp(-4a)
  float scale;
  // where the output permutation writes: transform b, element r at
  // (b % o_bi)*o_in + (b / o_bi)*o_out + r*o_stride (rows: 1, 1, n, 0)
  int64_t o_stride, o_bi, o_out, o_in;
};

// LDS elements a pass with C*n = cn needs (one pad slot per 16, see lpad)
inline int lds_elems(int cn) { return cn + (cn >> 4) + 1; }

}  // namespace fftjh
