// Host-side launchers of the HIP kernels (one per kernel family).
// dtype: 0 = float32, 1 = float64.  All launches are asynchronous on `stream`.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "common.h"

namespace fftjh {

// d: host copy (launch geometry); dev_desc: the same descriptor in device memory
hipError_t launch_pass(int dtype, int dir, const PassDesc& d, const PassDesc* dev_desc, const void* x,
                       void* y, const void* tw, hipStream_t stream);
// largest workgroup the pass kernel is compiled for
int max_threads(int dtype);
// workgroups of this pass that fit the device at once (0 if unknown)
int pass_resident_groups(int dtype, int dir, const PassDesc& d);

// compile-time-chain kernels for power-of-two lengths 32..16384
bool pow2_chain(int n, int* radices, int* count);
// ... and for the listed other lengths: the chain and its elements per
// thread (16 for the powers of two)
bool fixed_chain(int n, int* radices, int* count, int* elems);
hipError_t launch_pass_pow2(int dtype, int dir, const PassDesc& d, const PassDesc* dev_desc, const void* x,
                            void* y, const void* tw, hipStream_t stream);
int pass_pow2_resident_groups(int dtype, int dir, const PassDesc& d);
int pow2_max_threads(int dtype);

// lean column-pass kernel (FP32, n = 64, 128, 256); tws: the table
// w_n^i, i < n, of the pass's direction
bool lean_length(int n);
// square lengths on the lean kernel (stockham_lean_sq.hip): columns per
// group for length n (0: none)
int lean_sq_cols(int n);
hipError_t launch_lean_sq(int dir, const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t stream);
hipError_t launch_lean(int dir, const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t stream);
// the paired real edges of a multi-pass half-length transform (stockham_lean_pair.hip)
hipError_t launch_lean_pair(int dir, const LeanDesc& d, const void* x, void* y, const void* tws, hipStream_t stream);

// HvdH route kernels (FP32): the Delta permutation dst[Delta*l mod n] =
// scale*src[l], the spread of the permuted input onto grid rows of t, the
// chunked banded solve grid -> unknowns in l order; the largest halo and
// the core rows per chunk the solve kernel is built for
hipError_t launch_hvdh_permute(const HvdhDesc& d, const void* src, void* dst, float scale, hipStream_t stream);
hipError_t launch_hvdh_spread(const HvdhDesc& d, const void* xp, void* v, hipStream_t stream);
hipError_t launch_hvdh_solve(const HvdhDesc& d, const void* Yg, void* y, hipStream_t stream);
int hvdh_max_halo();
int hvdh_chunk_rows();

// split-precision MFMA pass kernels (FP32): mode is a MatMode (stages_mat.h),
// mat the operand table launch_mat_table wrote for that mode and direction
hipError_t launch_pass_mat(int mode, int dir, const PassDesc& d, const PassDesc* dev_desc, const void* x, void* y,
                           const void* tw, const void* mat, hipStream_t stream);
int pass_mat_resident_groups(int mode, int dir, const PassDesc& d);
// bytes of the operand table of a mode; its generator (FP64 on the device)
int64_t mat_table_bytes(int mode);
hipError_t launch_mat_table(int mode, int dir, void* out, hipStream_t stream);

hipError_t launch_tw_stage(int dtype, void* out, int ns, int R, int dir, hipStream_t stream);
hipError_t launch_tw_pow(int dtype, void* out, int64_t count, int64_t mult, int64_t L, int dir,
                         hipStream_t stream);

hipError_t launch_r2c_twist(int dtype, const void* Z, void* y, const void* w, int64_t h, int64_t batch,
                            int64_t zs, int64_t ys, hipStream_t st);
hipError_t launch_c2r_pre(int dtype, const void* X, void* Z, const void* w, int64_t h, int64_t batch,
                          int64_t xs, int64_t zs, double pre, hipStream_t st);
hipError_t launch_real_to_complex(int dtype, const void* x, void* z, int64_t n, int64_t batch,
                                  hipStream_t st);
hipError_t launch_copy_prefix(int dtype, const void* z, void* y, int64_t n, int64_t keep, int64_t batch,
                              hipStream_t st);
hipError_t launch_hermitian_fill(int dtype, const void* X, void* z, int64_t n, int64_t batch,
                                 hipStream_t st);
hipError_t launch_take_real(int dtype, const void* z, void* y, int64_t count, hipStream_t st);

hipError_t launch_rs_corr(int dtype, void* corr, int64_t n, int64_t m, int w, double beta, const double* zq,
                          const double* wq, int nq, double scale, hipStream_t st);
// deconv reads transforms at geometry gx (batch index split by Bi) and
// writes contiguous rows of m; interp reads rows of m and writes geometry gy
hipError_t launch_rs_deconv(int dtype, const void* x, void* g, const void* corr, int64_t n, int64_t m,
                            int64_t batch, int64_t Bi, Geometry gx, hipStream_t st);
hipError_t launch_rs_interp(int dtype, const void* G, void* y, const void* coef, int pieces, int degree, int64_t n,
                            int64_t m, int w, int64_t batch, int64_t Bi, Geometry gy, hipStream_t st);
hipError_t launch_strided_copy(int dtype, const void* src, void* dst, int64_t n, int64_t batch, int64_t Bi,
                               Geometry gs, Geometry gd, hipStream_t st);

}  // namespace fftjh
