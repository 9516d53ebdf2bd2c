/* fftjh -- C-ABI of the MI355X (gfx950) FFT engine that stands in for the
 * plan/apply API of FFTJ.jl.
 *
 * The reference has no FFI: its "operator API" is the set of Julia plan
 * constructors and plan application methods.  Each entry point below names
 * the reference interface it replaces; a Julia package binds them with
 * ccall (the binding is shown in INTEGRATION.md of this repository).
 *
 * Conventions (all mirror the reference):
 *   - sign: -1 forward, +1 inverse (the reference's D, kernels.jl:3);
 *   - scale: multiplied into the output; the reference's defaults are 1 for
 *     fft and bfft and 1/n for ifft (plan.jl:97, 207-213);
 *   - data: interleaved complex (re, im) of the plan's precision; N-d arrays
 *     column-major, dims[0] fastest (ndplan.jl:7-31);
 *   - execution is alias safe: y == x is an in-place transform
 *     (api.jl:11-13);
 *   - errors are reported before any work is queued; codes 1 and 2 map to
 *     Julia's ArgumentError and DimensionMismatch (plan.jl:171, 361-362).
 *
 * Pointers passed to execute are device pointers (hipMalloc'd memory on the
 * plan's device).  Execution is asynchronous on the given stream, allocates
 * nothing, and may be captured into a hipGraph.  A plan owns its scratch, so
 * one plan must not run on two streams at once (the reference clones plans
 * for the same reason, fourstep.jl:332-336); distinct plans may.
 */
#ifndef FFTJH_H
#define FFTJH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFTJH_OK 0
#define FFTJH_ERR_ARGUMENT 1    /* Julia ArgumentError */
#define FFTJH_ERR_DIMENSION 2   /* Julia DimensionMismatch */
#define FFTJH_ERR_HIP 3         /* a HIP runtime call failed */
#define FFTJH_ERR_UNSUPPORTED 4 /* valid request outside what this library plans */

#define FFTJH_F32 0 /* Float32 / ComplexF32 */
#define FFTJH_F64 1 /* Float64 / ComplexF64 */

#define FFTJH_KIND_C2C 0
#define FFTJH_KIND_R2C 1
#define FFTJH_KIND_C2R 2
#define FFTJH_KIND_RAGGED 3

/* Plan flags (complex FP32 plans; at most one).  Radix-9, -16 and -32 stages
 * run as DFT-matrix products on gfx950 MFMA with split operands and FP32
 * accumulation; no reference counterpart (the reference is CPU-only):
 *   FFTJH_FLAG_MFMA_F16     FP16 hi + lo, per-butterfly power-of-two scale
 *   FFTJH_FLAG_MFMA_BF16    BF16 hi + lo (16 significant bits)
 *   FFTJH_FLAG_MFMA_BF16X3  BF16 hi + mid + lo (24 significant bits) */
#define FFTJH_FLAG_MFMA_F16 0x1u
#define FFTJH_FLAG_MFMA_BF16 0x2u
#define FFTJH_FLAG_MFMA_BF16X3 0x4u

/* Route flags for a 1-D length that is not 13-smooth, n = p * s with s its
 * 13-smooth part (at most one; 0 = the planner's choice):
 *   FFTJH_FLAG_ROUTE_COMPOSITE  one exact Cooley-Tukey step over s, then the
 *                               length-p transforms resampled (the
 *                               reference's CompositePlan, composite.jl:1-101);
 *                               applies when s > 1 and one pass holds s
 *   FFTJH_FLAG_ROUTE_RESAMPLE   the whole length resampled onto a smooth m
 *                               (ResamplePlan, resample.jl:258-320) */
#define FFTJH_FLAG_ROUTE_COMPOSITE 0x10u
#define FFTJH_FLAG_ROUTE_RESAMPLE 0x20u
/*   FFTJH_FLAG_ROUTE_HVDH       odd lengths: Gaussian resampling onto a smooth
 *                               t coprime to n and chunked banded solves (the
 *                               reference's HvdhPlan, hvdh.jl:154-241); an
 *                               even length (no grid coprime to it): one
 *                               exact step over s around HvdH transforms of
 *                               length p (SplitPlan, composite.jl:103-200);
 *                               FP32, contiguous rows */
#define FFTJH_FLAG_ROUTE_HVDH 0x40u

typedef struct fftjh_plan_s* fftjh_plan;
/* hipStream_t, passed as an opaque pointer (NULL = the null stream) */
typedef void* fftjh_stream;

/* Complex plan.  rank 1 with dims[0] = n replaces
 *   plan_fft / plan_ifft / plan_bfft(::Type{T}, n)        plan.jl:207-213
 * rank >= 2 replaces the array-based N-d constructors
 *   plan_fft / plan_ifft / plan_bfft(x::Array{Complex{T},N})  api.jl:107-124
 * (all dims transformed, dims[0] fastest).  batch > 1 runs that many
 * independent transforms stored back to back (an extension: the reference
 * has no batch API and loops over 1-D plans).  flags: 0, or one
 * FFTJH_FLAG_MFMA_* bit and/or one FFTJH_FLAG_ROUTE_* bit (the reference
 * accepts and ignores FFTW planner flags, api.jl:107-108; these are this
 * library's own). */
int fftjh_plan_c2c(fftjh_plan* plan, int dtype, int rank, const int64_t* dims, int64_t batch, int sign,
                   double scale, unsigned flags);

/* Search-space overrides for one plan (zero fields keep the planner's
 * choice, so a zero-initialized struct is the default plan).  The exhaustive planner and the autotuner enumerate
 * these; the reference has no counterpart (its routing is a fixed tree of
 * host thresholds, plan.jl:183-213, and FFTW flags are ignored). */
typedef struct fftjh_pass_tuning {
  int32_t C;            /* transforms per workgroup, a power of two <= 64 */
  int32_t nradix;       /* > 0: use radices[0..nradix) as the pass's stage chain */
  int32_t radices[16];  /* each in {2,3,4,5,7,8,11,13,16} (and 9, 32 with an
                           MFMA flag), product = pass length */
  int32_t xcd_remap;    /* 0 planner, 1 off, 2 on */
} fftjh_pass_tuning;

typedef struct fftjh_tuning {
  int64_t splits[4];            /* multi-pass factors of a 1-D plan: splits[0] the
                                   contiguous row pass (run last), then outward;
                                   zero-terminated, product = n */
  fftjh_pass_tuning pass[4];    /* by launch order of the pass kernels */
  uint32_t flags;               /* FFTJH_FLAG_* as for fftjh_plan_c2c */
  uint32_t pad_;
} fftjh_tuning;

/* fftjh_plan_c2c with explicit search-space choices (NULL = planner's). */
int fftjh_plan_c2c_tuned(fftjh_plan* plan, int dtype, int rank, const int64_t* dims, int64_t batch, int sign,
                         double scale, const fftjh_tuning* tune);

/* Ragged batch: `count` independent complex transforms of lengths sizes[i]
 * at complex offsets offsets[i] of x and y (NULL = packed back to back), run
 * by at most four launches: the short and the resampled transforms on
 * persistent item-granular launches, the long ones as two tile-granular
 * passes.  The item-granular launches run on a stream the plan owns, forked
 * from the execute call's stream by an event and joined back to it by
 * another, so the call stays ordered on that stream (and capturable in a
 * graph); fftjh_execute_static keeps all four on the caller's stream.
 * sign -1 forward; +1 inverse, scaled by
 * 1/sizes[i] when normalized != 0.  Lengths up to 65536 (FP32) / 32768
 * (FP64).  No reference counterpart: FFTJ has no batch or ragged API, and a
 * caller loops over 1-D plans (plan.jl:207-213); this stands in for that loop. */
int fftjh_plan_ragged(fftjh_plan* plan, int dtype, int64_t count, const int64_t* sizes, const int64_t* offsets,
                      int sign, int normalized, unsigned flags);

/* Ragged search-space overrides (zero fields keep the planner's choice). */
typedef struct fftjh_ragged_tuning {
  int32_t groups;        /* persistent workgroups (0 = all resident) */
  int32_t radix_pref;    /* 16 or 8: radix of the power-of-two tower */
  int32_t tile_max;      /* largest length done in one LDS tile, 256..4096 */
  int32_t pass_min;      /* smooth transforms longer than this run as two
                            tile-granular passes (0 = planner, -1 = never) */
  const double* costs;   /* per-transform predicted cost for the workgroup
                            assignment (NULL = the analytic model) */
} fftjh_ragged_tuning;

int fftjh_plan_ragged_tuned(fftjh_plan* plan, int dtype, int64_t count, const int64_t* sizes,
                            const int64_t* offsets, int sign, int normalized, const fftjh_ragged_tuning* tune);

/* Real-to-complex forward plan: n reals -> n/2+1 complex per transform.
 * Replaces plan_rfft(::Type{T}, n)                        real.jl:41-52 */
int fftjh_plan_r2c(fftjh_plan* plan, int dtype, int64_t n, int64_t batch, unsigned flags);

/* Complex-to-real inverse plan: n/2+1 complex -> n reals per transform.
 * normalized != 0 replaces plan_irfft(::Type{T}, n)        real.jl:57-70
 * normalized == 0 replaces plan_brfft(::Type{T}, n)        real.jl:72-73 */
int fftjh_plan_c2r(fftjh_plan* plan, int dtype, int64_t n, int64_t batch, int normalized, unsigned flags);

/* N-d real-to-complex plan over all dims of the column-major real array
 * dims[0..rank-1] (dims[0] fastest, halved to dims[0]/2+1 complex), `batch`
 * such arrays back to back.  Real rows along dims[0], then the complex
 * passes of the other dims in place on y.  rank 1 is fftjh_plan_r2c.
 * Replaces plan_rfft(X::Array{T,N})                      ndplan.jl:355-418,
 * api.jl:131-136 */
int fftjh_plan_r2c_nd(fftjh_plan* plan, int dtype, int rank, const int64_t* dims, int64_t batch, unsigned flags);

/* N-d complex-to-real plan: (dims[0]/2+1, dims[1], ...) complex -> dims
 * reals; the complex passes of dims[1..] run from x into plan scratch (x is
 * not written), then the real rows along dims[0].  normalized != 0 scales by
 * 1/prod(dims) (irfft), normalized == 0 by 1 (brfft).
 * Replaces plan_irfft(X::Array{Complex{T},N}, n1) / plan_brfft(X, n1)
 *                                                        ndplan.jl:435-521,
 * api.jl:137-154 */
int fftjh_plan_c2r_nd(fftjh_plan* plan, int dtype, int rank, const int64_t* dims, int64_t batch, int normalized,
                      unsigned flags);

/* Apply the plan: y = P x.  Replaces p(y, x) / mul!(y, p, x)
 * (plan.jl:359-396, fourstep.jl:338-356, real.jl:80-211, ndplan.jl:329-335,
 * api.jl:34-55).  No length checks (see fftjh_execute_checked). */
int fftjh_execute(fftjh_plan plan, void* y, const void* x, fftjh_stream stream);

/* Bytes of the fixed per-device scratch array of fftjh_execute_static. */
#define FFTJH_STATIC_SCRATCH_BYTES (64ll << 20)

/* As fftjh_execute, but allocation-free by construction: the plan's scratch
 * (the intermediates of multi-pass plans and the ragged per-workgroup
 * workspace) is carved out of one fixed array of FFTJH_STATIC_SCRATCH_BYTES
 * that the library holds on each GPU it runs on, from module load to process
 * end, instead of the buffers the plan allocated.  The call makes no
 * hipMalloc, hipFree or other allocating call (its first call on a device
 * looks up that device's array).  Plans needing more scratch than the array
 * holds are rejected with FFTJH_ERR_UNSUPPORTED before any work is queued.
 * Every call on one GPU shares that GPU's array and it keeps the last call's
 * contents (no plan relies on scratch starting at zero), so calls that
 * overlap in time on one GPU (two streams, two host threads) race on it and
 * must be serialised by the caller.  The current device must be the plan's.
 * No counterpart in the reference (its plans own their scratch, plan.jl:52-55,
 * and allocate nothing when nthreads = 1): this entry point is for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc. */
int fftjh_execute_static(fftjh_plan plan, void* y, const void* x, fftjh_stream stream);

/* As fftjh_execute, after checking the element counts of y and x against
 * the plan (complex elements for complex sides, reals for real sides);
 * returns FFTJH_ERR_DIMENSION before queuing anything on a mismatch, as
 * the reference's call methods do (plan.jl:361-362, real.jl:82-84). */
int fftjh_execute_checked(fftjh_plan plan, void* y, int64_t y_elems, const void* x, int64_t x_elems,
                          fftjh_stream stream);

/* Shape queries.  Replaces size(p) (api.jl:57-60, 166-168): kind, dtype,
 * sign, rank, dims, batch, scale and the element counts of one whole call's
 * input and output. */
int fftjh_plan_info(fftjh_plan plan, int* kind, int* dtype, int* sign, int* rank, int64_t* dims /* [8] */,
                    int64_t* batch, double* scale, int64_t* in_elems, int64_t* out_elems);

/* Human-readable schedule of the plan (kernel launches, radices, groups). */
int fftjh_plan_describe(fftjh_plan plan, char* buf, int64_t len);

/* Same description for a plan request, computed on the host only (no GPU
 * needed); kind is FFTJH_KIND_*; for C2R, normalized is taken as 1. */
int fftjh_describe_request(int kind, int dtype, int rank, const int64_t* dims, int64_t batch, int sign,
                           char* buf, int64_t len);

/* The same for a complex request with search-space choices (host only); the
 * planners score candidate schedules from it without building them.  No
 * reference counterpart (the reference has no measuring planner). */
int fftjh_describe_tuned(int dtype, int rank, const int64_t* dims, int64_t batch, int sign,
                         const fftjh_tuning* tune, char* buf, int64_t len);

/* Release the plan's scratch and its hold on the shared tables. */
int fftjh_plan_destroy(fftjh_plan plan);

/* Message of the last error on the calling thread ("" if none). */
const char* fftjh_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int fftjh_version(void);

/* Timing support for the benchmark: runs `iters` applications of the plan
 * on `stream` with HIP events around every kernel launch, and writes the
 * mean duration in milliseconds of launch i into ms[i] (up to *nlaunch
 * entries; *nlaunch is set to the plan's launch count).  Host-synchronous. */
int fftjh_profile_launches(fftjh_plan plan, void* y, const void* x, fftjh_stream stream, int iters,
                           double* ms, int* nlaunch);

/* ---- learned planner: the cost model's network ---------------------------
 * A 3-layer MLP (in -> hidden -> hidden -> 1, ReLU, FP32) that predicts the
 * log run time of a plan from its features, trained on timings the GPU plans
 * produce.  Forward, backward and Adam run in this library's HIP kernels.
 * No reference counterpart: FFTJ routes by fixed host thresholds
 * (plan.jl:183-213) and has no measuring planner.  X is a device array of
 * batch x in_features floats (row-major), targets/outputs device arrays of
 * batch floats.  Parameter layout: W1[h][in], b1[h], W2[h][h], b2[h],
 * W3[1][h], b3[1]. */
typedef struct fftjh_mlp_s* fftjh_mlp;
int fftjh_mlp_create(fftjh_mlp* model, int in_features, int hidden);
int fftjh_mlp_destroy(fftjh_mlp model);
int64_t fftjh_mlp_param_count(fftjh_mlp model);
int fftjh_mlp_set_params(fftjh_mlp model, const float* host_params);
int fftjh_mlp_get_params(fftjh_mlp model, float* host_params);
int fftjh_mlp_predict(fftjh_mlp model, const float* X, float* out, int batch, fftjh_stream stream);
/* gradients of L = sum_r (out_r - t_r)^2 / (2 batch) into host_grads (may be NULL) and the loss */
int fftjh_mlp_gradients(fftjh_mlp model, const float* X, const float* target, int batch, float* host_grads,
                        float* loss, fftjh_stream stream);
/* one Adam step (beta 0.9 / 0.999) on that loss; writes the loss before the step */
int fftjh_mlp_train_step(fftjh_mlp model, const float* X, const float* target, int batch, float lr, float* loss,
                         fftjh_stream stream);
const char* fftjh_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FFTJH_H */
