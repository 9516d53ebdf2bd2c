/* CPU oracle for the fftjh GPU engine -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of FFTJ.jl's CPU algorithms on the GPU hot path,
 * used (a) by tests/ as the parity checker and (b) by bench.py's
 * cpu_baseline leg as the "port" baseline (FFTJ itself needs Julia, which
 * this environment does not have).  Nothing in the product links or calls
 * it.  Each function cites the reference code it follows:
 *
 *   chain_init / build_chain   plan.jl:20-34, 66-95, 105-151 (factor_radices,
 *                              make_stage, _tower, build_stages)
 *   stage                      stockham.jl:26-130, 351-428; kernels.jl:93-228
 *   chain_run / fft_chain      plan.jl:311-321, 359-396
 *   fourstep                   fourstep.jl:272-495 (pass A / pass B)
 *   resample                   resample.jl:12-21, 33-55, 202-276, 392-424
 *   hvdh                       hvdh.jl:1-330, 560-688 (Gaussian resampling,
 *                              banded chunk solves)
 *   split                      composite.jl:103-262 (SplitPlan: length-p
 *                              HvdH sub-transforms, radix-s combine)
 *   oracle_rfft / irfft        real.jl:26-39, 80-211
 *   oracle_ndfft               ndplan.jl:57-118, 270-327
 *   chunk_range / baselines    threads.jl:15-38
 *   dft_naive                  the DFT definition (pinning only)
 *
 * Built three times through oracle_core.h: float (_f32, the CPU baseline's
 * arithmetic), double (_f64, the parity oracle: "the reference's own
 * Float64 output") and long double (_ld, for pinning).  Twiddles come from
 * exact integer angle numerators evaluated in long double, as the reference
 * evaluates them with an exact numerator in Float64 (plan.jl:76-77).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------- helpers */

/* exp(D*2*pi*i*e/L) in long double: e reduced mod L in integers, then to the
 * first octant, so the only rounding is that of sinl/cosl on |a| <= pi/4 */
static void expi_exact(int64_t e, int64_t L, int D, long double* c, long double* s) {
  e %= L;
  if (e < 0) e += L;
  /* 8e/L in octants: angle = 2*pi*e/L */
  __int128 e8 = (__int128)e * 8;
  int64_t oct = (int64_t)(e8 / L);          /* 0..7 */
  __int128 rem = e8 - (__int128)oct * L;    /* angle within octant: (pi/4)*rem/L */
  long double a;
  long double sn, cs;
  const long double PI4 = 0.785398163397448309615660845819875721L;
  /* quadrant q = oct/2; inside the quadrant use the nearer end */
  int q = (int)(oct / 2);
  int odd = (int)(oct & 1);
  if (!odd) {
    a = PI4 * (long double)rem / (long double)L;
    sn = sinl(a);
    cs = cosl(a);
  } else {
    a = PI4 * (long double)((__int128)L - rem) / (long double)L;
    sn = cosl(a);
    cs = sinl(a);
  }
  long double rc, rs;
  switch (q) {
    case 0: rc = cs; rs = sn; break;
    case 1: rc = -sn; rs = cs; break;
    case 2: rc = -cs; rs = -sn; break;
    default: rc = sn; rs = -cs; break;
  }
  *c = rc;
  *s = D < 0 ? -rs : rs;
}

static int64_t gcd64(int64_t a, int64_t b) {
  while (b) {
    int64_t r = a % b;
    a = b;
    b = r;
  }
  return a < 0 ? -a : a;
}

/* number of bits of x (Base.top_set_bit): 0 for x = 0 */
static int top_bit(int64_t x) {
  int k = 0;
  while (x > 0) {
    ++k;
    x >>= 1;
  }
  return k;
}

/* HvdH's t (hvdh.jl:75-92): smooth, coprime to n, at least n(1 + 1/128);
 * odd n: the smallest o * 2^k over SMOOTH_ODD (resample.jl:12) with
 * gcd(o, n) = 1; even n: the smallest odd 13-smooth m; -1 if t - n >= n */
static int is_smooth13(int64_t n);
static int64_t hvdh_t(int64_t n) {
  static const int64_t kSmoothOdd[7] = {1, 3, 5, 7, 9, 15, 21};
  const int64_t N = (int64_t)ceill((long double)n * (1.0L + 1.0L / 128));
  int64_t t = INT64_MAX;
  if (n % 2) {
    for (int i = 0; i < 7; ++i) {
      const int64_t o = kSmoothOdd[i];
      if (gcd64(o, n) != 1) continue;
      const int64_t c = o << top_bit((N + o - 1) / o - 1);
      if (c < t) t = c;
    }
  } else {
    for (int64_t m = N | 1; m <= 2 * n - 1; m += 2)
      if (is_smooth13(m) && gcd64(m, n) == 1) {
        t = m;
        break;
      }
  }
  return (t != INT64_MAX && t - n < n) ? t : -1;
}

static int64_t isqrt64(int64_t n) {
  if (n <= 0) return 0;
  int64_t r = (int64_t)sqrtl((long double)n);
  while (r * r > n) --r;
  while ((r + 1) * (r + 1) <= n) ++r;
  return r;
}

static const int kOdd[5] = {13, 11, 7, 5, 3}; /* ODD_RADICES, plan.jl:16 */

static int is_smooth13(int64_t n) {
  if (n < 1) return 0;
  while (n % 2 == 0) n /= 2;
  for (int i = 0; i < 5; ++i)
    while (n % kOdd[i] == 0) n /= kOdd[i];
  return n == 1;
}

/* build_stages (plan.jl:125-151) with _tower (plan.jl:105-117), allow_small:
 * leading radix-4 stages until the stride reaches W, a tower of 4s then 8s,
 * odd radices smallest first, then the twiddle-free final 2/4/8 */
static int build_chain(int64_t n, int W, int* rad) {
  if (!is_smooth13(n)) return -1;
  int a = 0;
  int64_t r = n;
  while (r % 2 == 0) {
    r /= 2;
    ++a;
  }
  int cnt[5] = {0, 0, 0, 0, 0};
  for (int i = 0; i < 5; ++i)
    while (r % kOdd[i] == 0) {
      r /= kOdd[i];
      ++cnt[i];
    }
  int nst = 0, f = 1;
  int tower[64], nt = 0;
  if (a == 0) {
    f = 1;
  } else if (a <= 3) {
    f = 1 << a;
  } else {
    int k = 0;
    int64_t p4 = 1;
    while (p4 < W && 2 * (k + 1) <= a - 2) {
      ++k;
      p4 *= 4;
    }
    int b = a - 2 * k;
    int n4 = (b % 3 == 0) ? 0 : (b % 3 == 1) ? 2 : 1;
    int all[64], na = 0;
    for (int i = 0; i < k + n4; ++i) all[na++] = 4;
    for (int i = 0; i < (b - 2 * n4) / 3; ++i) all[na++] = 8;
    f = all[na - 1];
    for (int i = 0; i < na - 1; ++i) tower[nt++] = all[i];
  }
  for (int i = 0; i < nt; ++i) rad[nst++] = tower[i];
  for (int i = 4; i >= 0; --i) /* reverse of ODD_RADICES: 3, 5, 7, 11, 13 */
    for (int c = 0; c < cnt[i]; ++c) rad[nst++] = kOdd[i];
  if (f > 1) rad[nst++] = f;
  return nst;
}

/* nextsmooth (resample.jl:12-21) */
static int64_t nextsmooth(int64_t N) {
  static const int64_t odd[7] = {1, 3, 5, 7, 9, 15, 21};
  int64_t best = INT64_MAX;
  for (int i = 0; i < 7; ++i) {
    int64_t o = odd[i];
    int64_t q = (N + o - 1) / o - 1;
    int k = 0;
    while (q) {
      q >>= 1;
      ++k;
    }
    if (k < 6) k = 6;
    if ((o << k) < best) best = o << k;
  }
  return best;
}

/* _gausslegendre (resample.jl:33-53) in FP64 */
static void gauss_legendre(int nq, double* z, double* wq) {
  for (int i = 1; i <= nq; ++i) {
    double x = cos(M_PI * (i - 0.25) / (nq + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 20; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= nq; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = nq * (x * p1 - p0) / (x * x - 1.0);
      double dx = p1 / dp;
      x -= dx;
      if (fabs(dx) < 4 * 2.220446049250313e-16) break;
    }
    z[i - 1] = x;
    wq[i - 1] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
}

/* _es (resample.jl:55) */
static double es_kernel(double z, double beta) {
  return fabs(z) >= 1.0 ? 0.0 : exp(beta * (sqrt(1.0 - z * z) - 1.0));
}

/* tforeach_chunks chunking (threads.jl:15-38), 0-based half-open ranges */
static void chunk_range(int64_t len, int nt, int c, int64_t align, int64_t* lo, int64_t* hi) {
  int64_t units = len / align;
  int64_t nchunks = nt < units ? nt : units;
  if (nchunks <= 1) {
    *lo = c == 0 ? 0 : len;
    *hi = len;
    return;
  }
  if (c >= nchunks) {
    *lo = *hi = len;
    return;
  }
  int64_t per = units / nchunks, rem = units % nchunks;
  int64_t l = 0;
  for (int i = 0; i < c; ++i) l += (per + (i < rem)) * align;
  int64_t u = per + (c < rem);
  *lo = l;
  *hi = (c == nchunks - 1) ? len : l + u * align;
}

/* bf4_core (kernels.jl:105-120), per type */
#define DEF_BF4(T, SUF)                                                                              \
  static void bf4_core##SUF(T ar, T ai, T br, T bi, T cr, T ci, T dr, T di, int D, T* yr, T* yi) { \
    T apcr = ar + cr, apci = ai + ci, amcr = ar - cr, amci = ai - ci;                              \
    T bpdr = br + dr, bpdi = bi + di, bmdr = br - dr, bmdi = bi - di;                              \
    yr[0] = apcr + bpdr;                                                                           \
    yi[0] = apci + bpdi;                                                                           \
    if (D == -1) {                                                                                 \
      yr[1] = amcr + bmdi;                                                                         \
      yi[1] = amci - bmdr;                                                                         \
      yr[3] = amcr - bmdi;                                                                         \
      yi[3] = amci + bmdr;                                                                         \
    } else {                                                                                       \
      yr[1] = amcr - bmdi;                                                                         \
      yi[1] = amci + bmdr;                                                                         \
      yr[3] = amcr + bmdi;                                                                         \
      yi[3] = amci - bmdr;                                                                         \
    }                                                                                              \
    yr[2] = apcr - bpdr;                                                                           \
    yi[2] = apci - bpdi;                                                                           \
  }

DEF_BF4(float, _f32)
DEF_BF4(double, _f64)
DEF_BF4(long double, _ld)

/* resampling tolerance 10*eps(T) (resample.jl:203) */
static long double resample_tol_f32(void) { return 10.0L * 1.1920928955078125e-07L; }
static long double resample_tol_f64(void) { return 10.0L * 2.220446049250313e-16L; }
static long double resample_tol_ld(void) { return 10.0L * 2.220446049250313e-16L; }

#define T float
#define SUF _f32
#define TW 16
#include "oracle_core.h"
#undef T
#undef SUF
#undef TW

#define T double
#define SUF _f64
#define TW 8
#include "oracle_core.h"
#undef T
#undef SUF
#undef TW

#define T long double
#define SUF _ld
#define TW 8
#include "oracle_core.h"
#undef T
#undef SUF
#undef TW

/* --------------------------------------------------- exported extras */

/* chain radices the reference builds for n on an AVX-512 host (F32: W=16,
 * F64: W=8); returns the stage count or -1 */
int oracle_chain_radices(int64_t n, int W, int* rad) { return build_chain(n, W, rad); }

int64_t oracle_nextsmooth(int64_t n) { return nextsmooth(n); }

int64_t oracle_hvdh_t(int64_t n) { return hvdh_t(n); }

void oracle_gauss_legendre(int nq, double* z, double* w) { gauss_legendre(nq, z, w); }

void oracle_chunk_range(int64_t len, int nt, int c, int64_t align, int64_t* lo, int64_t* hi) {
  chunk_range(len, nt, c, align, lo, hi);
}

void oracle_expi(int64_t e, int64_t L, int D, double* c, double* s) {
  long double cc, ss;
  expi_exact(e, L, D, &cc, &ss);
  *c = (double)cc;
  *s = (double)ss;
}

/* splitmix64 stream: draw i of seed s is mix(s + (i+1)*0x9e3779b97f4a7c15) */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

void oracle_splitmix64(uint64_t seed, uint64_t first, int64_t count, uint64_t* out) {
  for (int64_t i = 0; i < count; ++i) out[i] = mix64(seed + (first + (uint64_t)i + 1) * 0x9e3779b97f4a7c15ULL);
}
