// Host planner and C-ABI entry points (include/fftjh.h).
//
// A plan is a list of steps (kernel launches) over five buffer roles: the
// caller's x and y and up to four scratch buffers owned by the plan.
// Planning runs in two phases: layout() decides every launch on the host
// alone (it also serves fftjh_describe_request without a GPU), then
// materialize() generates the twiddle/correction tables on the GPU (FP64,
// shared through a process-wide cache like the reference's _STAGE_CACHE,
// plan.jl:153-167), allocates scratch and uploads the pass descriptors.
// Execution only launches kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fftjh.h"
#include "common.h"
#include "launch.h"
#include "ragged.h"
#include "stages_mat.h"

namespace fftjh {

// ---------------------------------------------------------------- errors --

static thread_local std::string g_err;

struct Fail {
  int code;
  std::string msg;
};

[[noreturn]] static void fail(int code, const std::string& msg) { throw Fail{code, msg}; }

static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(FFTJH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------- device buffers --

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  explicit DevBuf(size_t b) : bytes(b) {
    if (b) hip_check(hipMalloc(&p, b), "hipMalloc");
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};
using DevBufPtr = std::shared_ptr<DevBuf>;

// Immutable tables shared across plans: key -> weak reference, so a table
// lives while some plan uses it (the reference keeps its entries for the
// session; here device memory is returned when the last plan goes).
static std::mutex g_cache_mu;
static std::map<std::string, std::weak_ptr<DevBuf>> g_cache;

template <typename F>
static DevBufPtr cached(const std::string& key, size_t bytes, F&& fill) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    if (auto sp = it->second.lock()) return sp;
  }
  auto buf = std::make_shared<DevBuf>(bytes);
  fill(buf->p);
  hip_check(hipDeviceSynchronize(), "table generation");
  g_cache[key] = buf;
  return buf;
}

// ------------------------------------------------------- factorization --

static const int kOddRadices[] = {3, 5, 7, 11, 13};

// n = 2^a * prod p^c over {3,5,7,11,13} (plan.jl:20-34 uses the same prime set)
static bool factor13(int64_t n, int& a, int cnt[5]) {
  if (n < 1) return false;
  a = 0;
  while ((n & 1) == 0) {
    n >>= 1;
    ++a;
  }
  for (int i = 0; i < 5; ++i) {
    cnt[i] = 0;
    while (n % kOddRadices[i] == 0) {
      n /= kOddRadices[i];
      ++cnt[i];
    }
  }
  return n == 1;
}

// default radix chain for one pass: the power-of-two part in radix-16
// stages (remainder 2/4/8 first), then the odd primes ascending
static std::vector<int> default_radices(int64_t n) {
  int a, cnt[5];
  std::vector<int> r;
  if (!factor13(n, a, cnt)) return r;
  if (a % 4) r.push_back(1 << (a % 4));
  for (int i = 0; i < a / 4; ++i) r.push_back(16);
  for (int i = 0; i < 5; ++i)
    for (int k = 0; k < cnt[i]; ++k) r.push_back(kOddRadices[i]);
  return r;
}

// threads one column of a chain needs: every stage's butterflies covered,
// kElemsPerThread / R butterflies per thread
static int64_t chain_threads(int64_t n, const std::vector<int>& radices) {
  int64_t t = 1;
  for (int R : radices) {
    const int64_t bpt = kElemsPerThread / R;
    t = std::max<int64_t>(t, (n / R + bpt - 1) / bpt);
  }
  return t;
}

// The fewest stages over the VALU radices, composites 6, 9, 10, 12, 14, 15
// included (e.g. 100 = 10*10, 1000 = 10*10*10, 81 = 9*9 where the prime
// chain takes 3, 4 and 4 stages); ties go to the chain with fewer threads
// per column, then more power-of-two stages.  Power-of-two stages first
// (smallest first), the rest ascending.  Empty when n is not 13-smooth.
static std::vector<int> compact_radices(int64_t n) {
  static const int kValu[] = {16, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2};
  int a, cnt[5];
  if (n <= 1 || !factor13(n, a, cnt)) return {};
  std::vector<int> best, cur;
  std::function<void(int64_t, int)> rec = [&](int64_t m, int maxr) {
    if (!best.empty() && cur.size() >= best.size() && m > 1) return;
    if (m == 1) {
      std::vector<int> c = cur;
      std::sort(c.begin(), c.end(), [](int x, int y) {
        const bool px = (x & (x - 1)) == 0, py = (y & (y - 1)) == 0;
        if (px != py) return px;
        return x < y;
      });
      if (best.empty() || c.size() < best.size()) {
        best = c;
        return;
      }
      if (c.size() > best.size()) return;
      const int64_t tc = chain_threads(n, c), tb = chain_threads(n, best);
      auto pow2s = [](const std::vector<int>& v) {
        int k = 0;
        for (int R : v) k += (R & (R - 1)) == 0;
        return k;
      };
      if (tc < tb || (tc == tb && pow2s(c) > pow2s(best))) best = c;
      return;
    }
    for (int R : kValu) {
      if (R > maxr || m % R) continue;
      cur.push_back(R);
      rec(m / R, R);
      cur.pop_back();
    }
  };
  rec(n, 16);
  return best;
}

static int esize(int dtype) { return dtype == FFTJH_F32 ? 8 : 16; }

// MFMA chains: powers of two as radix-16 stages after one radix-2/4/8 stage,
// pairs of threes as radix-9 stages (both on MFMA), the rest on VALU
static std::vector<int> mat_radices(int64_t n) {
  int a, cnt[5];
  std::vector<int> r;
  if (!factor13(n, a, cnt)) return r;
  if (a % 4) r.push_back(1 << (a % 4));
  for (int i = 0; i < a / 4; ++i) r.push_back(16);
  for (int k = 0; k < cnt[0] / 2; ++k) r.push_back(9);
  if (cnt[0] % 2) r.push_back(3);
  for (int i = 1; i < 5; ++i)
    for (int k = 0; k < cnt[i]; ++k) r.push_back(kOddRadices[i]);
  return r;
}

// threads of an MFMA chain: a wave holds mat_tiles_per_wave(R) tiles of 16
// butterflies in a matrix stage; VALU stages as in pass_threads
static int pass_threads_mat(int n, int C, const std::vector<int>& radices) {
  int64_t t = 64;
  for (int R : radices) {
    const int64_t nb = static_cast<int64_t>(C) * (n / R);
    if (mat_radix(R)) {
      const int64_t tiles = (nb + 15) / 16;
      const int64_t waves = (tiles + mat_tiles_per_wave(R) - 1) / mat_tiles_per_wave(R);
      t = std::max<int64_t>(t, 64 * waves);
    } else {
      const int bpt = kElemsPerThread / R;
      t = std::max<int64_t>(t, (nb + bpt - 1) / bpt);
    }
  }
  t = (t + 63) / 64 * 64;
  return static_cast<int>(std::max<int64_t>(t, C));
}

constexpr int kMatThreads = kMaxThreads / 2;  // launch bound of the MFMA pass kernels

static bool pass_fits_mat(int n, int C, const std::vector<int>& radices, int* threads) {
  if (static_cast<int64_t>(lds_elems(C * n)) * 8 > kLdsBytes) return false;
  const int t = pass_threads_mat(n, C, radices);
  if (t > kMatThreads) return false;
  if (threads) *threads = t;
  return true;
}

// threads one workgroup needs for C transforms of length n with `radices`
static int pass_threads(int n, int C, const std::vector<int>& radices) {
  int t = 64;
  for (int R : radices) {
    const int bpt = kElemsPerThread / R;
    const int64_t nb = static_cast<int64_t>(C) * (n / R);
    t = std::max<int64_t>(t, (nb + bpt - 1) / bpt);
  }
  t = (t + 63) / 64 * 64;
  return std::max(t, C);
}

static bool pass_fits(int dtype, int n, int C, const std::vector<int>& radices, int* threads) {
  if (static_cast<int64_t>(lds_elems(C * n)) * esize(dtype) > kLdsBytes) return false;
  const int t = pass_threads(n, C, radices);
  if (t > max_threads(dtype)) return false;
  if (threads) *threads = t;
  return true;
}

// a row pass may drop to C = cap if a workgroup of that size still has a wave
static bool pass_fits_rows_ok(int dtype, int n, int C, int cap) {
  (void)dtype;
  (void)C;
  return static_cast<int64_t>(cap) * (n / 16) >= 64;
}

// elements per thread of the compile-time-chain kernel for length n (16 for
// the powers of two), 0 when n has none
static int chain_elems(int n) {
  int pr[8], pc = 0, e = 0;
  return fixed_chain(n, pr, &pc, &e) ? e : 0;
}

// the compile-time-chain kernel: E elements per thread, C*n/E threads
static bool pow2_fits(int dtype, int n, int C, int* threads) {
  if (static_cast<int64_t>(lds_elems(C * n)) * esize(dtype) > kLdsBytes) return false;
  const int E = chain_elems(n);
  if (!E) return false;
  const int t = C * (n / E);
  if (t > pow2_max_threads(dtype) || t % 64) return false;
  if (threads) *threads = t;
  return true;
}

// Rows per workgroup of a power-of-two row pass fused with the real edges;
// 0 when the length has no fused kernel
static int real_fused_C(int dtype, int64_t h, int64_t batch, int* threads) {
  static const bool off = [] {
    const char* e = std::getenv("FFTJH_REAL_FUSED");
    return (e && std::atoi(e) == 0) || std::getenv("FFTJH_NO_POW2") != nullptr;
  }();
  int pr[8], pc = 0;
  if (off || h > (1 << 14) || !pow2_chain(static_cast<int>(h), pr, &pc)) return 0;
  (void)batch;
  const int n = static_cast<int>(h);
  // the fewest rows per workgroup that fill a wave (as the row passes)
  for (int C = 1; C <= 64; C <<= 1) {
    int t;
    if (!pow2_fits(dtype, n, C, &t)) continue;
    if (threads) *threads = t;
    return C;
  }
  return 0;
}

static bool single_pass_ok(int dtype, int64_t n) {
  if (n == 1) return true;
  if (n > (1 << 16)) return false;
  auto r = default_radices(n);
  if (r.empty()) return false;
  return pass_fits(dtype, static_cast<int>(n), 1, r, nullptr);
}

// ---------------------------------------------------------------- steps --

enum Role { RX = 0, RY = 1, RS0 = 2, RS1 = 3, RS2 = 4, RS3 = 5, kRoles = 6 };

enum StepKind {
  ST_PASS,
  ST_R2C_TWIST,
  ST_C2R_PRE,
  ST_REAL_TO_COMPLEX,
  ST_COPY_PREFIX,
  ST_HERMITIAN_FILL,
  ST_TAKE_REAL,
  ST_RS_DECONV,
  ST_RS_INTERP,
  ST_STRIDED_COPY,
  ST_RAGGED,
  ST_LEAN,
  ST_HVDH_PERMUTE,
  ST_HVDH_SPREAD,
  ST_HVDH_SOLVE,
};

static const char* kStepNames[] = {"pass",           "r2c_twist", "c2r_pre",   "real_to_complex", "copy_prefix",
                                   "hermitian_fill", "take_real", "rs_deconv", "rs_interp",       "strided_copy",
                                   "ragged",         "lean",      "hvdh_permute", "hvdh_spread", "hvdh_solve"};

struct Step {
  int kind = ST_PASS;
  int src = RX, dst = RY;
  // roles when the caller's x and y are the same array (in-place call);
  // -1: the same as src / dst
  int alias_src = -1, alias_dst = -1;
  // ST_LEAN
  LeanDesc ld{};
  // ST_HVDH_SPREAD, ST_HVDH_SOLVE
  HvdhDesc hv{};
  int dir = -1;
  // ST_PASS
  PassDesc d{};
  bool pow2 = false;  // compile-time-chain kernel
  int mat = kMatNone;  // split-precision MFMA chain (radix 9/16/32 stages on MFMA)
  std::vector<int> radices;
  // other kernels
  int64_t n = 0, m = 0, h = 0, batch = 0, s1 = 0, s2 = 0, keep = 0;
  int w = 0;
  int pieces = 0, degree = 0;  // ST_RS_INTERP: piecewise-polynomial weights
  double beta = 0.0, pre = 1.0, scale = 1.0;
  // strided sides of rs_deconv (ga), rs_interp (gb), strided_copy (ga -> gb)
  int64_t Bi = 1;
  Geometry ga{1, 0, 0}, gb{1, 0, 0};
  // device resources
  DevBufPtr tw, desc, aux, mat_tab, ptw;
  std::shared_ptr<struct RaggedData> rg;  // ST_RAGGED: items of at most kRaggedTile points
  std::shared_ptr<struct RaggedData> rg2; // ST_RAGGED: longer items, on the large-item kernel
  std::shared_ptr<struct RaggedData> rg3; // ST_RAGGED: long smooth items as two tile-granular passes
};

struct Layout {
  fftjh_tuning tune{};  // search-space overrides (zero = planner's choice)
  int npass = 0;        // pass launches laid out so far (index into tune.pass)
  int mat = kMatNone;   // split-precision MFMA mode of every pass (FP32 C2C)
  unsigned route = 0;   // FFTJH_FLAG_ROUTE_* for non-smooth lengths (0 = planner)
  bool two_pass512 = true;  // the smaller 1-D lengths may take two passes of up to 512 points
  bool lean = false;        // a 1-D complex plan: its three-pass schedule may take the lean layout
  int dtype = 0, kind = FFTJH_KIND_C2C, sign = -1, rank = 1;
  int64_t dims[8] = {0};
  int64_t batch = 1;
  double scale = 1.0;
  int64_t in_elems = 0, out_elems = 0;
  std::vector<Step> steps;
  int64_t scratch_elems[4] = {0, 0, 0, 0};  // complex elements per scratch role
};

static void need_scratch(Layout& L, int role, int64_t elems) {
  if (role >= RS0) L.scratch_elems[role - RS0] = std::max(L.scratch_elems[role - RS0], elems);
}

// x / D = (x * m) >> s for 0 <= x < 2^31, 1 <= D < 2^31 (fdiv in passcommon.h)
static void fastdiv_params(int64_t D, uint32_t* m, int32_t* s) {
  int l = 0;
  while ((int64_t(1) << l) < D) ++l;
  *s = 31 + l;
  const unsigned __int128 p = static_cast<unsigned __int128>(1) << *s;
  *m = static_cast<uint32_t>((p + static_cast<unsigned __int128>(D) - 1) / static_cast<unsigned __int128>(D));
}

static uint32_t magic_for(int ns) {
  return ns > 1 ? static_cast<uint32_t>((uint64_t(1) << 32) / static_cast<uint64_t>(ns) + 1) : 0u;
}

// one launch of the pass kernel
// multi-pass twiddle of a pass (period 0 = none), see PassDesc
struct PassTw {
  int64_t period = 0, qdiv = 1, acoef = 0, bcoef = 0;
};

static void add_pass(Layout& L, int dir, int n, int64_t batch, int64_t Bi, Geometry in, Geometry out,
                     double scale, int src, int dst, PassTw tw = PassTw(), int64_t B2 = 1, int real_io = 0) {
  Step s;
  s.kind = ST_PASS;
  s.src = src;
  s.dst = dst;
  s.dir = dir;
  PassDesc& d = s.d;
  d.n = n;
  d.batch = batch;
  d.Bi = Bi;
  d.B2 = B2;
  d.tw_period = tw.period;
  d.tw_qdiv = tw.qdiv;
  d.tw_acoef = tw.acoef;
  d.tw_bcoef = tw.bcoef;
  const int64_t lim = int64_t(1) << 31;
  d.fast_batch = (batch < lim && Bi < lim && B2 < lim && tw.qdiv >= 1 && tw.qdiv < lim) ? 1 : 0;
  if (d.fast_batch) {
    fastdiv_params(Bi, &d.bi_m, &d.bi_s);
    fastdiv_params(B2, &d.b2_m, &d.b2_s);
    fastdiv_params(tw.qdiv, &d.qd_m, &d.qd_s);
  }
  if (tw.period) {
    const int64_t P = tw.period;
    d.tw_ar = tw.acoef % P;
    d.tw_br = tw.bcoef % P;
    d.tw_invp = 1.0 / static_cast<double>(P);
    int bits = 0;
    while ((int64_t(1) << bits) < P) ++bits;
    d.ptw_sh = (bits + 1) / 2;
    d.ptw_lo = ((P - 1) >> d.ptw_sh) + 1;  // entries of the high table
  }
  d.in = in;
  d.out = out;
  d.scale = scale;
  const fftjh_pass_tuning* pt = L.npass < 4 ? &L.tune.pass[L.npass] : nullptr;
  ++L.npass;
  s.mat = L.mat;
  s.radices = (n == 1) ? std::vector<int>{} : L.mat ? mat_radices(n) : default_radices(n);
  if (pt && pt->nradix > 0) {
    if (pt->nradix > kMaxStages) fail(FFTJH_ERR_ARGUMENT, "too many radices in tuning");
    std::vector<int> r(pt->radices, pt->radices + pt->nradix);
    int64_t prod = 1;
    for (int R : r) {
      const bool valu = R >= 2 && R <= 16;  // primes to 13, powers of two, composites 6..15
      if (!valu && !(L.mat && mat_radix(R)))
        fail(FFTJH_ERR_ARGUMENT, "tuning radix " + std::to_string(R) + " has no kernel");
      prod *= R;
    }
    if (prod != n) fail(FFTJH_ERR_ARGUMENT, "tuning radices do not multiply to the pass length");
    s.radices = r;
  }
  if (n > 1 && s.radices.empty()) fail(FFTJH_ERR_UNSUPPORTED, "pass length has a prime factor above 13");
  if (static_cast<int>(s.radices.size()) > kMaxStages) fail(FFTJH_ERR_UNSUPPORTED, "too many stages");
  // lengths with a compile-time chain (the powers of two, and the listed
  // others: FFTJH_FIXED_CHAINS=0 leaves those to the runtime chain) take the
  // fused kernel
  {
    int pr[8], pc = 0, pe = 0;
    static const bool no_pow2 = std::getenv("FFTJH_NO_POW2") != nullptr;
    static const bool fixed_env = [] {
      const char* e = std::getenv("FFTJH_FIXED_CHAINS");
      return !e || std::atoi(e) != 0;
    }();
    const bool p2 = pow2_chain(n, pr, &pc);
    // 81 = 9 x 9 (27 elements per thread) wins on strided loads (F32
    // 531,441 x 64 column passes 183 -> 130 us) and loses on contiguous rows
    // (81 x 200,000: 93 -> 104 us), so it takes only passes with strided loads
    const bool fixed = fixed_env && fixed_chain(n, pr, &pc, &pe) && (n != 81 || in.stride > 1);
    if (!no_pow2 && !L.mat && !(pt && pt->nradix > 0) && (p2 || fixed)) {
      s.pow2 = true;
      s.radices.assign(pr, pr + pc);
    }
  }
  int chainE = s.pow2 ? chain_elems(n) : 16;
  // transforms per workgroup: strided passes take adjacent columns for
  // contiguous HBM runs; row passes group short transforms to fill a workgroup
  const bool colmode = in.stride > 1 || out.stride > 1;
  const int row_cap = L.dtype == FFTJH_F32 ? 4 : 8;
  // the generic kernel's group size for a chain (0: none fits)
  auto generic_C = [&](const std::vector<int>& rad, int* tout) {
    for (int C = colmode ? 16 : 64; C >= 1; C >>= 1) {
      int t;
      if (!pass_fits(L.dtype, n, C, rad, &t)) continue;
      if (C > 1 && C > batch) continue;
      if (!colmode && t > 256 && C > 1) continue;
      if (!colmode && C > row_cap && t >= 64 * (C / row_cap) && pass_fits_rows_ok(L.dtype, n, C, row_cap)) continue;
      if (tout) *tout = t;
      return C;
    }
    return 0;
  };
  // composite radices: fewer stages, taken where the chain still fits as
  // many transforms per workgroup (FFTJH_COMPOSITE_RADICES=0 turns it off)
  auto apply_compact = [&] {
    static const bool comp_env = [] {
      const char* e = std::getenv("FFTJH_COMPOSITE_RADICES");
      return !e || std::atoi(e) != 0;
    }();
    if (comp_env && !s.pow2 && !L.mat && !(pt && pt->nradix > 0) && n > 1) {
      std::vector<int> rc = compact_radices(n);
      if (!rc.empty() && rc.size() < s.radices.size() && static_cast<int>(rc.size()) <= kMaxStages) {
        const int c_old = generic_C(s.radices, nullptr), c_new = generic_C(rc, nullptr);
        if (c_new > 0 && c_new >= c_old) s.radices = rc;
      }
    }
  };
  apply_compact();
  // a compile-time chain that no group size fits: the generic kernel, with
  // its own elements per thread and radix choice
  auto to_generic = [&] {
    s.pow2 = false;
    s.radices = default_radices(n);
    chainE = 16;
    apply_compact();
  };
  auto fits = [&](int C, int* t) {
    if (L.mat) return pass_fits_mat(n, C, s.radices, t);
    return s.pow2 ? pow2_fits(L.dtype, n, C, t) : pass_fits(L.dtype, n, C, s.radices, t);
  };
  int bestC = 0, bestT = 0;
  if (real_io) {
    // rows of a real transform's half-length spectrum, real edges fused
    bestC = real_fused_C(L.dtype, n, batch, &bestT);
    if (!bestC || !s.pow2) fail(FFTJH_ERR_UNSUPPORTED, "internal: no fused real-edge kernel for this length");
  } else if (pt && pt->C > 0) {
    int t;
    if ((pt->C & (pt->C - 1)) || pt->C > 64) fail(FFTJH_ERR_ARGUMENT, "tuning C must be a power of two <= 64");
    // (a chain that fits only other group sizes -- the 81 chain runs three
    // threads per transform and needs C = 64 -- gives way to the generic kernel)
    bool ok = fits(pt->C, &t);
    if (!ok && s.pow2 && !(pt->nradix > 0)) {
      to_generic();
      ok = fits(pt->C, &t);
    }
    if (!ok) fail(FFTJH_ERR_ARGUMENT, "tuning C=" + std::to_string(pt->C) + " does not fit one workgroup");
    bestC = pt->C;
    bestT = t;
  } else {
    // row passes: a wave covers C rows x 64/C consecutive elements, so keep
    // C <= 128 bytes / 8*esize ... i.e. at least 128-byte runs
    for (int attempt = 0; attempt < 2 && !bestC; ++attempt) {
    if (attempt == 1) {
      if (!s.pow2) break;
      to_generic();  // no group size fits the compile-time chain
    }
    if (!colmode && s.pow2) {
      // power-of-two row passes: the fewest rows per workgroup that fill a
      // wave, so each wave reads the longest contiguous runs (measured on
      // 256 MiB of rows: n = 64 C 64 -> 16 2.4 -> 3.5 TB/s, n = 1024 C 4 -> 1
      // 3.4 -> 4.4 TB/s)
      for (int C = 1; C <= 64 && !bestC; C <<= 1) {
        int t;
        if (fits(C, &t)) {
          bestC = C;
          bestT = t;
        }
      }
      if (bestC) break;
    }
    // (a fixed chain of E != 16 elements may need 32 columns to fill waves)
    for (int C = colmode ? (chainE != 16 ? 64 : 16) : 64; C >= 1; C >>= 1) {
      int t;
      if (!fits(C, &t)) continue;
      // more columns than transforms only to fill one wave (idle columns)
      if (C > 1 && C > batch && !(s.pow2 && (C / 2) * (n / chainE) < 64)) continue;
      if (!colmode && t > 256 && C > 1) continue;
      if (!colmode && C > row_cap && t >= 64 * (C / row_cap) && pass_fits_rows_ok(L.dtype, n, C, row_cap)) continue;
      bestC = C;
      bestT = t;
      break;
    }
    }
  }
  if (!bestC) fail(FFTJH_ERR_UNSUPPORTED, "pass does not fit one workgroup");
  // small strided passes (8 MiB and less) whose groups would not cover the
  // 256 CUs: fewer transforms per workgroup, more workgroups (latency bound).
  // Measured (us, off -> on): 147,456 x 4 grid 19.7 -> 16.0 (65,537 x 4 28.4
  // -> 24.9), 100,003 29.1 -> 22.1, 3^9 x 2 14.0 -> 11.7, 54,675 x 2 14.4 ->
  // 13.2, F64 10,007 x 3 even; single-pass contiguous rows lose (1000 x 100
  // 7.1 -> 7.7), so they keep their group size
  static const bool spread_env = [] {
    const char* e = std::getenv("FFTJH_SMALL_SPREAD");
    return !e || std::atoi(e) != 0;
  }();
  if (spread_env && colmode && !(pt && pt->C > 0) && !real_io && !s.pow2 &&
      n * batch * esize(L.dtype) <= (int64_t(8) << 20)) {
    while (bestC > 1 && (batch + bestC - 1) / bestC < 256) {
      int t;
      if (!fits(bestC / 2, &t)) break;
      bestC /= 2;
      bestT = t;
    }
  }
  d.C = bestC;
  d.logC = 0;
  while ((1 << d.logC) < bestC) ++d.logC;
  d.nthreads = bestT;
  const int64_t groups = (batch + bestC - 1) / bestC;
  if (groups > 0x7fffffff) fail(FFTJH_ERR_UNSUPPORTED, "batch too large for one launch");
  d.ngroups = static_cast<int>(groups);
  d.xcd_remap = colmode ? 1 : 0;
  // mixed passes (contiguous rows on one side, strided columns on the other)
  // stage the row side through LDS with row-major lanes
  // (measured slower than register I/O at the variant's register cost: off
  // unless FFTJH_ROWS_IO=1)
  static const bool rows_io = std::getenv("FFTJH_ROWS_IO") != nullptr;
  if (rows_io && s.pow2 && (n & (n - 1)) == 0 && d.C > (L.dtype == FFTJH_F32 ? 4 : 8)) {
    // the C transforms of a group must sit at a constant offset step
    int64_t cin = -1, cout = -1;
    if (Bi > 1 && Bi % d.C == 0) {
      cin = in.s_in;
      cout = out.s_in;
    } else if (Bi == 1 && B2 > 1 && B2 % d.C == 0) {
      cin = in.s_mid;
      cout = out.s_mid;
    } else if (Bi == 1 && B2 == 1) {
      cin = in.s_out;
      cout = out.s_out;
    }
    if (cin >= 0) {
      d.in_cdelta = cin;
      d.out_cdelta = cout;
      d.rows_in = (in.stride == 1 && out.stride > 1) ? 1 : 0;
      d.rows_out = (out.stride == 1 && in.stride > 1 && d.tw_period == 0) ? 1 : 0;
    }
  }
  if (real_io) {
    // row r of a group at base(g*C) + r*row stride on both sides
    d.real_io = real_io;
    d.real_h = n;
    int bits = 0;
    while ((int64_t(1) << bits) < 2 * static_cast<int64_t>(n)) ++bits;
    d.rtw_sh = (bits + 1) / 2;
    d.rtw_hi = ((2 * static_cast<int64_t>(n) - 1) >> d.rtw_sh) + 1;
    d.rows_in = d.rows_out = 0;
    d.in_cdelta = in.s_out;
    d.out_cdelta = out.s_out;
    d.xcd_remap = 0;
  }
  if (pt && pt->xcd_remap > 0) d.xcd_remap = pt->xcd_remap == 2 ? 1 : 0;
  {
    // generic chains: first stage fused into the load, last into the store.
    // Measured (fused / unfused time): rows of 1000+ points and strided
    // passes of 64+ gain (4000 x 16384 0.83, 6000 x 4096 0.81, 3^10 x 64
    // 0.89, 10^6 x 8 0.97); short rows and short strided passes lose a
    // little (100 x 2^20 1.06, 12 x 2^22 1.03, 49^3 x 256 1.03), so they keep
    // the unfused kernel.  FFTJH_GENERIC_FUSE=0 turns it off.
    static const bool fuse_env = [] {
      const char* e = std::getenv("FFTJH_GENERIC_FUSE");
      return !e || std::atoi(e) != 0;
    }();
    d.fuse_io = (fuse_env && !s.pow2 && !L.mat && (n >= 512 || (colmode && n >= 64))) ? 1 : 0;
  }
  int ns = 1, off = 0;
  d.nstages = static_cast<int>(s.radices.size());
  for (size_t i = 0; i < s.radices.size(); ++i) {
    const int R = s.radices[i];
    d.st[i].radix = R;
    d.st[i].ns = ns;
    d.st[i].tw_off = off;
    d.st[i].ns_magic = magic_for(ns);
    off += (R - 1) * ns;
    ns *= R;
  }
  L.steps.push_back(s);
}

static Geometry geo(int64_t stride, int64_t s_in, int64_t s_out, int64_t s_mid = 0) {
  return Geometry{stride, s_in, s_out, s_mid};
}

// Factors n = F[0]*F[1]*...*F[P-1] of a multi-pass plan, each a single pass:
// F[0] is the contiguous row pass (run last), F[P-1] the outermost column
// pass (run first).  The fewest passes whose factors fit max_factor, then
// the most balanced split (the role of _fourstep_split/_fs_pick,
// fourstep.jl:166-199: nearest divisor to sqrt(n) with buildable factors).
static void factor_search(int dtype, int64_t n, int parts, int64_t max_factor, std::vector<int64_t>& cur,
                          std::vector<int64_t>& best, double& best_score) {
  if (parts == 1) {
    if (n > max_factor || !single_pass_ok(dtype, n)) return;
    cur.push_back(n);
    double lo = 1e300, hi = 0;
    int non_pow2 = 0;
    for (int64_t f : cur) {
      lo = std::min(lo, static_cast<double>(f));
      hi = std::max(hi, static_cast<double>(f));
      non_pow2 += (f & (f - 1)) != 0;
    }
    const double score = std::log(hi / lo) + 0.01 * non_pow2;
    if (score < best_score - 1e-12) {
      best_score = score;
      best = cur;
    }
    cur.pop_back();
    return;
  }
  for (int64_t f = 2; f <= std::min<int64_t>(n, max_factor); ++f) {
    if (n % f || !single_pass_ok(dtype, f)) continue;
    cur.push_back(f);
    factor_search(dtype, n / f, parts - 1, max_factor, cur, best, best_score);
    cur.pop_back();
  }
}

static std::vector<int64_t> choose_factors(int dtype, int64_t n, int64_t max_factor) {
  for (int parts = 2; parts <= 4; ++parts) {
    std::vector<int64_t> cur, best;
    double score = 1e300;
    factor_search(dtype, n, parts, max_factor, cur, best, score);
    if (!best.empty()) return best;
  }
  return {};
}

// The P-part split with every column factor F[1..P-1] on the 81 = 9 x 9 or
// 100 = 10 x 10 compile-time chain, at least one of them 81, and the row
// factor F[0] a single pass of at most 512 points, most balanced first;
// empty when there is none.  Out of the Infinity Cache (FP32, 256-512 MiB
// per array) the 81 chain's strided passes run at about twice the rate of
// the runtime chain's (4*3^12 x 64: 0.51 ms per 81-point column pass against
// 1.01-1.09 ms per 162-point one), and every such split measured was
// faster than the balanced one (52 lengths x 2 batches, 1.01x - 1.57x); with
// power-of-two or 100-only columns the result was mixed (0.75x - 1.35x), and
// in cache (32 MiB) mixed too (0.88x - 1.21x).
static std::vector<int64_t> chain_column_factors(int dtype, int64_t n, int parts) {
  std::vector<int64_t> best, cur;
  double best_score = 1e300;
  std::function<void(int64_t, int)> rec = [&](int64_t m, int left) {
    if (left == 0) {
      if (m > 512 || !single_pass_ok(dtype, m)) return;
      if (std::find(cur.begin(), cur.end(), int64_t(81)) == cur.end()) return;
      double lo = static_cast<double>(m), hi = lo;
      for (int64_t f : cur) {
        lo = std::min(lo, static_cast<double>(f));
        hi = std::max(hi, static_cast<double>(f));
      }
      const double score = std::log(hi / lo);
      if (score < best_score - 1e-12) {
        best_score = score;
        best.assign(1, m);
        best.insert(best.end(), cur.begin(), cur.end());
      }
      return;
    }
    for (int64_t f : {int64_t(81), int64_t(100)}) {
      if (m % f) continue;
      cur.push_back(f);
      rec(m / f, left - 1);
      cur.pop_back();
    }
  };
  rec(n, parts - 1);
  return best;
}

// largest factor the default multi-pass planner uses (the tile of one pass)
static int64_t default_max_factor(int dtype) {
  static const int64_t env = [] {
    const char* e = std::getenv("FFTJH_MAX_FACTOR");
    return e ? std::atoll(e) : 0LL;
  }();
  if (env > 0) return env;
  return 256;
}

// ---- resampling parameters (resample.jl:12-21, 33-53, 202-256) ----

static int64_t nextsmooth(int64_t N) {
  static const int64_t odd[] = {1, 3, 5, 7, 9, 15, 21};
  int64_t best = INT64_MAX;
  for (int64_t o : odd) {
    int64_t q = (N + o - 1) / o - 1;
    int k = 0;
    while (q) {
      q >>= 1;
      ++k;
    }
    k = std::max(k, 6);
    best = std::min(best, o << k);
  }
  return best;
}

struct RsParams {
  int64_t m;
  int w;
  double beta;
};

static RsParams rs_params(int dtype, int64_t n) {
  RsParams p;
  p.m = nextsmooth(2 * n);
  const double eps = dtype == FFTJH_F32 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
  const double tol = 10.0 * eps;
  const double ratio = std::max(static_cast<double>(n) / p.m, 0.5);
  const int w0 = static_cast<int>(std::ceil(std::log(1.0 / tol) / (M_PI * std::sqrt(1.0 - ratio))));
  p.w = 4 * ((w0 + 3) / 4);
  p.beta = 0.97 * M_PI * p.w * (1.0 - static_cast<double>(n) / (2.0 * p.m));
  return p;
}

// Gauss-Legendre nodes/weights on [-1, 1] by Newton on the three-term
// recurrence (the reference's _gausslegendre, resample.jl:33-53)
static void gauss_legendre(int nq, std::vector<double>& z, std::vector<double>& wq) {
  z.assign(nq, 0.0);
  wq.assign(nq, 0.0);
  for (int i = 1; i <= nq; ++i) {
    double x = std::cos(M_PI * (i - 0.25) / (nq + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 20; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= nq; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = nq * (x * p1 - p0) / (x * x - 1.0);
      const double dx = p1 / dp;
      x -= dx;
      if (std::fabs(dx) < 4 * 2.220446049250313e-16) break;
    }
    z[i - 1] = x;
    wq[i - 1] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
}

// Interpolation weights as piecewise polynomials in the offset delta = r/n
// of an output from its grid point (resample.jl:59-170 chooses the same
// representation for long transforms): delta's range [0, 1) is cut into
// `pieces`; on piece p, tap i's weight psi((delta + h - 1 - i)/h) as a
// function of t = 2(delta*pieces - p) - 1 in [-1, 1] is interpolated at the
// degree+1 Chebyshev points (long double), kept as Chebyshev coefficients
// (evaluated by Clenshaw's recurrence on the device), the degree raised until
// the error on a dense grid meets the precision's target or stops halving.
// coef[((p*w + i)*(degree + 1)) + k].
struct RsWeights {
  int degree = 0;
  std::vector<long double> coef;
};

static RsWeights fit_rs_weights(int dtype, int w, double beta, int pieces) {
  const int h = w / 2;
  const long double target = dtype == FFTJH_F32 ? 2e-9L : 2e-16L;
  const int dmax = dtype == FFTJH_F32 ? 16 : 32;
  auto psi = [&](long double z) -> long double {
    const long double a = std::fabs(z);
    return a >= 1.0L ? 0.0L : std::exp(static_cast<long double>(beta) * (std::sqrt(1.0L - a * a) - 1.0L));
  };
  const long double pi = 3.141592653589793238462643383279502884L;
  RsWeights best;
  long double best_err = 1e30L;
  for (int d = 2; d <= dmax; ++d) {
    const int N = d + 1;
    std::vector<long double> coef(static_cast<size_t>(pieces) * w * N);
    long double err = 0;
    for (int p = 0; p < pieces; ++p)
      for (int i = 0; i < w; ++i) {
        auto f = [&](long double t) {
          const long double delta = (p + (t + 1.0L) / 2.0L) / pieces;
          return psi((delta + (h - 1 - i)) / h);
        };
        long double* a = &coef[(static_cast<size_t>(p) * w + i) * N];
        std::vector<long double> fj(N);
        for (int j = 0; j < N; ++j) fj[j] = f(std::cos(pi * (j + 0.5L) / N));
        for (int k = 0; k < N; ++k) {
          long double sum = 0;
          for (int j = 0; j < N; ++j) sum += fj[j] * std::cos(pi * k * (j + 0.5L) / N);
          a[k] = sum * 2.0L / N;
        }
        a[0] /= 2.0L;
        // error on a dense grid, with the coefficients rounded as stored
        for (int g = 0; g <= 4 * N + 4; ++g) {
          const long double t = -1.0L + 2.0L * g / (4 * N + 4);
          long double b1 = 0, b2 = 0;
          for (int k = d; k >= 1; --k) {
            const long double ak = dtype == FFTJH_F32 ? static_cast<long double>(static_cast<float>(a[k]))
                                                      : static_cast<long double>(static_cast<double>(a[k]));
            const long double b0 = ak + 2.0L * t * b1 - b2;
            b2 = b1;
            b1 = b0;
          }
          const long double a0 = dtype == FFTJH_F32 ? static_cast<long double>(static_cast<float>(a[0]))
                                                    : static_cast<long double>(static_cast<double>(a[0]));
          err = std::max(err, std::fabs(a0 + t * b1 - b2 - f(t)));
        }
      }
    if (err < best_err) {
      const bool halving = err <= best_err / 2;
      best.degree = d;
      best.coef = coef;
      best_err = err;
      if (err <= target || !halving) break;
    } else {
      break;
    }
  }
  return best;
}

static bool is_rows(const Geometry& g, int64_t n, int64_t Bi) { return Bi == 1 && g.stride == 1 && g.s_out == n; }

// Complex transform of `batch` length-n transforms laid out at geometry gin
// (read from src) and gout (written to dst), batch index split by Bi, using
// the scratch roles in `free` (front first).
// Column block of an N-d axis: transform b is column b % Bi of outer block
// b / Bi, element j at j*Bi (the layout of every axis but the first)
static bool is_colblock(const Geometry& g, int64_t n, int64_t Bi) {
  return Bi > 1 && g.stride == Bi && g.s_in == 1 && g.s_out == Bi * n && g.s_mid == 0;
}

// Strided axes of at least this length split into passes of 16 columns per
// workgroup (128-byte runs) instead of one pass whose tile holds only 1-2
// columns (8-16-byte runs).  Measured (2-D, one column axis): at 4096 the
// split is 10-30% faster (F64 4096 x 4096: 0.52 -> 0.35 ms), at 8192 2x;
// at 2048 it is slower.  FFTJH_COL_SPLIT=<length> moves the threshold.
static bool split_column(int dtype, int64_t n) {
  (void)dtype;
  static const int64_t env = [] {
    const char* e = std::getenv("FFTJH_COL_SPLIT");
    return e ? std::atoll(e) : -1LL;
  }();
  if (env >= 0) return env > 0 && n >= env;
  return n >= 4096;
}

// Composite route for a non-smooth length n with 13-smooth part s > 1 (the
// caller checks that one pass holds s): the route flag decides, else the
// measured rule.  The composite route pays one extra pass (the length-s
// step) and gains when the resampled length-p transforms fit one pass where
// the whole length's would not.  Measured on one MI355X (composite / full
// resampling time, F32 and F64): 48*1009 0.87 / 0.92, 100*1009 0.89 / 0.94,
// 1000*1009 0.77 / 1.07, 1024*1031 0.96 / 1.01, 256*4099 1.02 / 0.90; it
// loses with a small smooth part (2*10007 1.31, 16*10007 1.11), a length-p
// step of more than one pass (64*10007 1.09), a tiny p (32*17 1.95) or a
// step too long for 16 columns per workgroup (4096*17 1.78).
static bool use_composite(const Layout& L, int64_t n, int64_t s) {
  if (L.route & FFTJH_FLAG_ROUTE_COMPOSITE) return true;
  if (L.route & (FFTJH_FLAG_ROUTE_RESAMPLE | FFTJH_FLAG_ROUTE_HVDH)) return false;
  const int64_t p = n / s;
  if (s < 32 || s > 1024 || p < 512) return false;
  return single_pass_ok(L.dtype, rs_params(L.dtype, p).m) && !single_pass_ok(L.dtype, rs_params(L.dtype, n).m);
}

// ---- HvdH parameters (hvdh.jl:60-190) ----

static int64_t gcd64(int64_t a, int64_t b) {
  while (b) {
    const int64_t r = a % b;
    a = b;
    b = r;
  }
  return a;
}

// t coprime to odd n: the smallest o * 2^k >= ceil(n (1 + 1/128)) over the
// odd smooth o in {1, 3, 5, 7, 9, 15, 21} coprime to n (hvdh.jl:76-92);
// 0 when there is none or t - n >= n
static int64_t hvdh_grid(int64_t n) {
  if (n % 2 == 0) return 0;
  const int64_t N = static_cast<int64_t>(std::ceil(static_cast<long double>(n) * (1.0L + 1.0L / 128)));
  int64_t t = INT64_MAX;
  for (int64_t o : {1, 3, 5, 7, 9, 15, 21}) {
    if (gcd64(o, n) != 1) continue;
    int64_t c = o;
    while (c < N) c <<= 1;
    t = std::min(t, c);
  }
  return (t != INT64_MAX && t - n < n) ? t : 0;
}

// x^-1 mod m for gcd(x, m) = 1 (extended Euclid)
static int64_t inv_mod(int64_t x, int64_t m) {
  int64_t r0 = m, r1 = ((x % m) + m) % m, s0 = 0, s1 = 1;
  while (r1) {
    const int64_t q = r0 / r1, r2 = r0 - q * r1, s2 = s0 - q * s1;
    r0 = r1;
    r1 = r2;
    s0 = s1;
    s1 = s2;
  }
  return ((s0 % m) + m) % m;
}

// Halo rows per side (hvdh.jl:118-146): for probe unknowns l0 (16 spread over
// [0, n) and the two with Delta*l0 = n/2, n/2 + 1 mod n, where beta is near
// +-1/2 and the inverse decays slowest), the row of the inverse of the
// (2P+1)-window of the l-indexed system around l0, solved densely in long
// double, decays below tol within H columns; H' = ceil(rho H) + B rows of q.
// 0 when no window up to P = 8192 shows the decay.
static int64_t hvdh_halo(int64_t n, int64_t t, long double a, int B, long double tol) {
  const long double rho = static_cast<long double>(t) / n;
  const int64_t c = inv_mod(t - n, n);
  std::vector<int64_t> probes;
  const int64_t cnt = std::min<int64_t>(n, 16);
  for (int64_t k = 0; k < cnt; ++k)
    probes.push_back(cnt > 1 ? static_cast<int64_t>(std::llround(static_cast<long double>(k) * (n - 1) / (cnt - 1))) : 0);
  probes.push_back(static_cast<int64_t>((static_cast<__int128>(n / 2) * c) % n));
  probes.push_back(static_cast<int64_t>((static_cast<__int128>(n / 2 + 1) * c) % n));
  for (int64_t P = 64; P <= 8192; P *= 2) {
    const int64_t m = 2 * P + 1, w = B + 1;
    int64_t H = 0;
    for (int64_t l0 : probes) {
      // transpose of the window, banded (half width w): MT[j][i] = M[i][j],
      // M[i][j] = exp(-a (rho l_j - floor((2t l_i + n)/(2n)))^2); solve
      // MT r = e_P by Gaussian elimination without pivoting on the band
      std::vector<long double> A(static_cast<size_t>(m * (2 * w + 1)), 0.0L), r(static_cast<size_t>(m), 0.0L);
      auto at = [&](int64_t row, int64_t col) -> long double& { return A[row * (2 * w + 1) + (col - row + w)]; };
      for (int64_t i = 0; i < m; ++i) {
        const int64_t li = l0 - P + i;
        const int64_t num = 2 * t * li + n, den = 2 * n;
        const int64_t qi = num >= 0 ? num / den : -((-num + den - 1) / den);
        for (int64_t j = std::max<int64_t>(0, i - w); j <= std::min<int64_t>(m - 1, i + w); ++j) {
          const long double dd = rho * static_cast<long double>(l0 - P + j) - static_cast<long double>(qi);
          at(j, i) = std::exp(-a * dd * dd);
        }
      }
      r[P] = 1.0L;
      for (int64_t k = 0; k < m; ++k) {
        const long double piv = at(k, k);
        if (piv == 0.0L) return 0;
        for (int64_t i = k + 1; i <= std::min<int64_t>(m - 1, k + w); ++i) {
          const long double f = at(i, k) / piv;
          if (f == 0.0L) continue;
          for (int64_t j = k; j <= std::min<int64_t>(m - 1, k + w); ++j) at(i, j) -= f * at(k, j);
          r[i] -= f * r[k];
        }
      }
      for (int64_t k = m - 1; k >= 0; --k) {
        long double sacc = r[k];
        for (int64_t j = k + 1; j <= std::min<int64_t>(m - 1, k + w); ++j) sacc -= at(k, j) * r[j];
        r[k] = sacc / at(k, k);
      }
      long double mx = 0.0L;
      for (int64_t j = 0; j < m; ++j) mx = std::max(mx, std::fabs(r[j]));
      for (int64_t j = 0; j < m; ++j)
        if (std::fabs(r[j]) > tol * mx) H = std::max<int64_t>(H, j > P ? j - P : P - j);
    }
    if (H < P - 2 * B) return static_cast<int64_t>(std::ceil(rho * H)) + B;
  }
  return 0;
}

// The route's parameters for FP32 (hvdh.jl:154-190): alpha = 1, a = pi
// alpha^2, b = pi/alpha^2, tol = eps/4, h = ceil(alpha sqrt(ln(1/tol)/pi))
// taps per side, band B = floor(sqrt(ln(1/tol)/a) + 1/2), the probed halo.
// False when the length has no grid or the kernels' limits do not hold.
static bool hvdh_params(int dtype, int64_t n, int64_t batch, double scale, HvdhDesc& d) {
  if (dtype != FFTJH_F32 || n % 2 == 0 || n < 64 || n >= (int64_t(1) << 30)) return false;
  const int64_t t = hvdh_grid(n);
  if (!t) return false;
  const long double alpha = 1.0L, pi = 3.14159265358979323846264338327950288L;
  const long double a = pi * alpha * alpha, b = pi / (alpha * alpha);
  const long double tol = std::ldexp(1.0L, -25);
  const int h = static_cast<int>(std::ceil(alpha * std::sqrt(std::log(1 / tol) / pi)));
  const int B = static_cast<int>(std::floor(std::sqrt(std::log(1 / tol) / a) + 0.5L));
  if (B != 2 || h < 1 || h > 7 || 2 * (t - n) >= n) return false;
  const int64_t H = hvdh_halo(n, t, a, B, tol);
  if (H < 1 || H > hvdh_max_halo()) return false;
  d = HvdhDesc{};
  d.n = n;
  d.t = t;
  d.delta = t - n;
  d.cinv = inv_mod(d.delta, n);
  for (int o = 1 - h; o <= h; ++o) d.tap[o + h] = (((o - 1) * d.delta) % n + n) % n;
  d.batch = batch;
  d.chunks = (t + hvdh_chunk_rows() - 1) / hvdh_chunk_rows();
  d.h = h;
  d.B = B;
  d.H = static_cast<int32_t>(H);
  d.b = static_cast<double>(b);
  d.ialpha = static_cast<double>(1 / alpha);
  for (int o = 0; o <= h; ++o) d.kc[o] = static_cast<double>(std::exp(-b * o * o));
  d.a = static_cast<float>(a);
  d.nt = static_cast<float>(static_cast<long double>(n) / t);
  d.c1 = static_cast<float>(std::exp(-a));
  d.c2 = static_cast<float>(std::exp(-4 * a));
  d.scale = static_cast<float>(scale);
  d.o_stride = 1;
  d.o_bi = 1;
  d.o_out = n;
  d.o_in = 0;
  return true;
}

// The HvdH route for an odd non-smooth length: the route flag forces it (where
// its parameters hold), the other route flags exclude it.  By default the
// planner keeps resampling, measured faster on one MI355X at every length
// tried (F32 x 4, forward, us: 16,387 49 / 20, 65,537 59 / 34, 100,003 75 /
// 38, 300,007 100 / 57, 1,000,003 213 / 115): on the GPU the grid transform
// is cheap, and the route's two permutations and its solve (a sequential
// recurrence per chunk) cost more than resampling's doubled grid.  The same
// holds for even lengths split around HvdH (F32 x 4, forward, us: 2*65,537
// 96 / 50, 16*65,537 278 / 157, 4*1,000,003 1,091 / 520).
// FFTJH_HVDH=1 makes it the default where it applies.
static bool use_hvdh(const Layout& L, int64_t n) {
  if (L.route & FFTJH_FLAG_ROUTE_HVDH) return true;
  if (L.route & (FFTJH_FLAG_ROUTE_RESAMPLE | FFTJH_FLAG_ROUTE_COMPOSITE)) return false;
  static const int env = [] {
    const char* e = std::getenv("FFTJH_HVDH");
    return e ? std::atoi(e) : 0;
  }();
  (void)n;
  return env != 0;
}

// Three power-of-two passes run by the lean kernel inside the output array
// (FP32, factors 64..256, contiguous rows, default tuning).  With
// j = j1 + N1*j2 + N1*N2*j3 and k = k3 + N3*k2 + N2*N3*k1:
//   A  x -> y: columns j3 -> k3 of (j1, j2) at stride N1*N2, written as the
//      contiguous row k3 at N3*j2 + N2*N3*j1, times w_{N2 N3}^(j2*k3);
//   B  y -> y: j2 -> k2 of (j1, k3) at stride N3, in place, times
//      w_n^(j1*(k3 + N3*k2));
//   C  y -> y: j1 -> k1 of k' = k3 + N3*k2 at stride N2*N3, in place, to the
//      natural-order output k' + N2*N3*k1.
// The plan touches only x and y (two arrays where the rotating layout
// touches three: measured on one MI355X, a 128 MiB pass of 128-byte runs
// copies in 33 us within two buffers and 48-50 us across three).  When the
// caller passes x == y the same launches run x -> z -> z -> y through the
// scratch array z.  FFTJH_LEAN=0 keeps the generic three-pass layout.
static bool tune_is_default(const fftjh_tuning& t) {
  static const fftjh_tuning zero{};
  return std::memcmp(&t, &zero, sizeof t) == 0;
}

static bool lean3_ok(const Layout& L, const std::vector<int64_t>& F, int64_t n, int64_t batch, int src, int dst) {
  static const bool env = [] {
    const char* e = std::getenv("FFTJH_LEAN");
    return !e || std::atoi(e) != 0;
  }();
  // split-precision plans take the lean kernel's MFMA variant (radix-16
  // stages on MFMA; FFTJH_LEAN_MAT=0 keeps them on the generic MFMA chain)
  static const bool mat_env = [] {
    const char* e = std::getenv("FFTJH_LEAN_MAT");
    return !e || std::atoi(e) != 0;
  }();
  fftjh_tuning t = L.tune;
  t.flags &= ~(FFTJH_FLAG_MFMA_F16 | FFTJH_FLAG_MFMA_BF16 | FFTJH_FLAG_MFMA_BF16X3);
  if (!env || !L.lean || L.dtype != FFTJH_F32 || (L.mat && !mat_env) || !tune_is_default(t) || src != RX ||
      dst != RY)
    return false;
  if (F.size() != 3) return false;
  // square factors (100 = 10 x 10) on the square kernel: all three equal, C
  // columns per group dividing every pass's column count
  const int sq = lean_sq_cols(static_cast<int>(F[0]));
  if (sq && !L.mat && F[1] == F[0] && F[2] == F[0]) {
    static const bool sq_env = [] {
      const char* e = std::getenv("FFTJH_LEAN_SQ");
      return !e || std::atoi(e) != 0;
    }();
    return sq_env && F[0] % sq == 0 && (F[1] * F[2]) % sq == 0 && n < (int64_t(1) << 31) &&
           batch * (n / (F[0] * sq)) < (int64_t(1) << 31);
  }
  for (int64_t f : F)
    if (!lean_length(static_cast<int>(f))) return false;
  // 32-bit lane offsets (< n/16 + 16) and group counts (< 2^31)
  return n < (int64_t(1) << 31) && batch * (n / 1024) < (int64_t(1) << 31);
}

static void lean_twiddle(LeanDesc& d, int64_t P, int64_t a0, int64_t b) {
  int bits = 0;
  while ((int64_t(1) << bits) < P) ++bits;
  d.psh = bits;
  d.ptw_sh = (bits + 1) / 2;
  d.ptw_lo = ((P - 1) >> d.ptw_sh) + 1;
  d.tw_a0 = a0;
  d.tw_b = b;
  // a period that is not a power of two (the square kernel): reduced mod P
  d.pmod = (P & (P - 1)) ? P : 0;
  d.pinv = 1.0 / static_cast<double>(P);
}

static void add_lean3(Layout& L, const std::vector<int64_t>& F, int64_t n, int64_t batch, int dir, double scale,
                      int src, int dst, int z) {
  const int64_t N1 = F[0], N2 = F[1], N3 = F[2], N23 = N2 * N3;
  // columns per group: 16 (powers of two), or the square kernel's
  const int64_t CW = lean_sq_cols(static_cast<int>(N1)) && N2 == N1 && N3 == N1 ? lean_sq_cols(static_cast<int>(N1))
                                                                                  : 16;
  need_scratch(L, z, n * batch);  // the x == y variant
  auto mk = [&](int N, int rows, int64_t D0, int64_t D1, int s_, int d_, int as_, int ad_) {
    Step st;
    st.kind = ST_LEAN;
    st.dir = dir;
    st.src = s_;
    st.dst = d_;
    st.alias_src = as_;
    st.alias_dst = ad_;
    st.ld.n = N;
    st.ld.rows = rows;
    st.ld.D0 = static_cast<int32_t>(D0);
    st.ld.D1 = static_cast<int32_t>(D1);
    st.ld.groups = static_cast<int32_t>(batch * (n / (CW * N)));
    st.ld.scale = 1.0;
    st.ld.mat = L.mat;
    st.ld.cols = CW == 16 ? 0 : static_cast<int32_t>(CW);
    st.n = N;
    st.batch = batch * (n / N);
    return st;
  };
  Step a = mk(static_cast<int>(N3), 1, N1 / CW, N2, src, dst, src, z);
  a.ld.in_s0 = CW;
  a.ld.in_s1 = N1;
  a.ld.in_s2 = n;
  a.ld.istride = N1 * N2;
  a.ld.out_s0 = CW * N23;
  a.ld.out_s1 = N3;
  a.ld.out_s2 = n;
  a.ld.ostride = N23;
  lean_twiddle(a.ld, N23, 0, 1);
  Step b = mk(static_cast<int>(N2), 0, N3 / CW, N1, dst, dst, z, z);
  b.ld.in_s0 = b.ld.out_s0 = CW;
  b.ld.in_s1 = b.ld.out_s1 = N23;
  b.ld.in_s2 = b.ld.out_s2 = n;
  b.ld.istride = b.ld.ostride = N3;
  lean_twiddle(b.ld, n, 1, N3);
  Step c = mk(static_cast<int>(N1), 0, N23 / CW, batch, dst, dst, z, dst);
  c.ld.in_s0 = c.ld.out_s0 = CW;
  c.ld.in_s1 = c.ld.out_s1 = n;
  c.ld.istride = c.ld.ostride = N23;
  c.ld.scale = scale;
  // Streaming hints (the split-exchange VALU kernel only).  Pass A's loads
  // of x are non-temporal when x and y together (16*n*batch bytes) fill the
  // 256 MB Infinity Cache: the plan does not read x again, and streamed it
  // leaves the cache to y, which passes B and C work on in place.  Measured
  // on one MI355X, 2^24 forward + inverse through three arrays: 0.281 ->
  // 0.270 ms per step; streaming pass C's stores as well lost (0.275 ms:
  // the inverse reads y next).  FFTJH_LEAN_NT: -1 (default) that rule, else
  // bit 0 streams pass A's loads and bit 1 pass C's stores.
  static const int lean_nt = [] {
    const char* e = std::getenv("FFTJH_LEAN_NT");
    return e ? std::atoi(e) : -1;
  }();
  if (!L.mat && CW == 16) {
    const int nt = lean_nt >= 0 ? lean_nt : (16 * n * batch >= (int64_t(256) << 20) ? 1 : 0);
    if (nt & 1) a.ld.nt = 1;
    if (nt & 2) c.ld.nt = 2;
  }
  L.steps.push_back(a);
  L.steps.push_back(b);
  L.steps.push_back(c);
  L.npass += 3;
}

static void build_c2c(Layout& L, int64_t n, int64_t batch, int64_t Bi, Geometry gin, Geometry gout, int dir,
                      double scale, int src, int dst, std::vector<int> free) {
  const bool split_asked = L.tune.splits[0] > 0 && L.npass == 0;
  int a, cnt[5];
  const bool smooth = factor13(n, a, cnt);
  const bool colblock = is_colblock(gin, n, Bi) && is_colblock(gout, n, Bi);
  if (single_pass_ok(L.dtype, n) && !split_asked && !(colblock && smooth && split_column(L.dtype, n))) {
    add_pass(L, dir, static_cast<int>(n), batch, Bi, gin, gout, scale, src, dst);
    return;
  }
  const Geometry rows = geo(1, 0, n);
  if (smooth) {
    if (!colblock && (!is_rows(gin, n, Bi) || !is_rows(gout, n, Bi))) {
      // a long strided axis: gather into rows, transform, scatter back
      if (free.size() < 2) fail(FFTJH_ERR_UNSUPPORTED, "internal: no scratch for a strided long axis");
      const int r = free[0];
      need_scratch(L, r, n * batch);
      Step g;
      g.kind = ST_STRIDED_COPY;
      g.src = src;
      g.dst = r;
      g.n = n;
      g.batch = batch;
      g.Bi = Bi;
      g.ga = gin;
      // row b*n under the two-level index b = (b % Bi) + Bi*(b / Bi)
      const Geometry rows_bi = geo(1, n, Bi * n);
      g.gb = rows_bi;
      L.steps.push_back(g);
      build_c2c(L, n, batch, 1, rows, rows, dir, scale, r, r, std::vector<int>(free.begin() + 1, free.end()));
      Step s2 = g;
      s2.src = r;
      s2.dst = dst;
      s2.ga = rows_bi;
      s2.gb = gout;
      L.steps.push_back(s2);
      return;
    }
    std::vector<int64_t> F;
    if (L.tune.splits[0] > 0 && L.npass == 0) {
      int64_t prod = 1;
      for (int i = 0; i < 4 && L.tune.splits[i] > 0; ++i) {
        F.push_back(L.tune.splits[i]);
        prod *= L.tune.splits[i];
        if (!single_pass_ok(L.dtype, L.tune.splits[i]))
          fail(FFTJH_ERR_ARGUMENT, "tuning factor " + std::to_string(L.tune.splits[i]) + " is not a single pass");
      }
      if (prod != n) fail(FFTJH_ERR_ARGUMENT, "tuning factors do not multiply to " + std::to_string(n));
    } else {
      // two passes of up to 512 points beat three of up to 256 for the
      // smaller 1-D lengths (measured: 2^17 x 64 F32 0.094 -> 0.081 ms, 10^5 x 64
      // 0.096 -> 0.060 ms, F64 2^17 x 64 0.165 -> 0.130 ms); from 2^17 on (F32)
      // the 512-point passes lose (196,608 x 32 0.064 -> 0.066 ms, 2^18 x 32
      // 0.082 -> 0.094 ms).  FFTJH_MAX_FACTOR overrides both.
      static const bool mf_env = std::getenv("FFTJH_MAX_FACTOR") != nullptr;
      const int64_t two_pass_max = L.dtype == FFTJH_F32 ? (int64_t(1) << 17) : (int64_t(1) << 18);
      if (!mf_env && L.two_pass512 && !colblock && n <= two_pass_max) {
        std::vector<int64_t> cur;
        double score = 1e300;
        factor_search(L.dtype, n, 2, 512, cur, F, score);
      }
      // small footprints (the array inside the L2s, 8 MiB and less) above that
      // length: two passes where three run latency bound, the column pass a
      // power of two of 64-256 points (the power-of-two kernel, and a
      // resampled grid's fused deconvolution), the row pass up to 4096.
      // Measured on the 147,456-point grid x 4 of 65,537: three passes
      // (64 x 48 x 48) 24.2 us, 2304 x 64 19.6 us, 1152 x 128 20.1 us,
      // 576 x 256 20.6 us, the balanced 384 x 384 20.9 us; F64 2^19 x 1
      // 27.9 -> 23.1 us.  FP32 powers of two keep the three lean passes
      // (2^18 x 1 13.4 us against 16.9, 2^20 x 1 24.1 against 25.3).
      static const bool small_env = [] {
        const char* e = std::getenv("FFTJH_SMALL_TWO_PASS");
        return !(e && std::atoi(e) == 0);
      }();
      const bool lean_pow2 = L.dtype == FFTJH_F32 && (n & (n - 1)) == 0;
      const bool small = small_env && !lean_pow2 && n * batch * esize(L.dtype) <= (int64_t(8) << 20);
      if (!mf_env && F.empty() && L.two_pass512 && !colblock && small && n > two_pass_max) {
        for (int64_t c = 64; c <= 256 && F.empty(); c *= 2)
          if (n % c == 0 && n / c <= 4096 && single_pass_ok(L.dtype, n / c)) F = {n / c, c};
      }
      if (F.empty()) F = choose_factors(L.dtype, n, default_max_factor(L.dtype));
      // FP32 plans of 128 MiB and more per array: column passes on the 81
      // and 100 chains where the balanced split leaves one to the runtime
      // chain (FFTJH_CHAIN_COLUMNS=0: off)
      static const bool chain_cols = [] {
        const char* e = std::getenv("FFTJH_CHAIN_COLUMNS");
        return !e || std::atoi(e) != 0;
      }();
      bool generic_col = false;
      for (size_t i = 1; i < F.size(); ++i) {
        int pr[8], pc = 0, pe = 0;
        generic_col |= !fixed_chain(static_cast<int>(F[i]), pr, &pc, &pe);
      }
      const bool big = n * batch * esize(L.dtype) >= (int64_t(128) << 20);
      if (chain_cols && !mf_env && L.dtype == FFTJH_F32 && !L.mat && !colblock && big && generic_col) {
        std::vector<int64_t> alt = chain_column_factors(L.dtype, n, static_cast<int>(F.size()));
        if (!alt.empty()) F = alt;
      }
    }
    if (F.size() < 2) fail(FFTJH_ERR_UNSUPPORTED, "no multi-pass split of length " + std::to_string(n));
    if (free.empty()) fail(FFTJH_ERR_UNSUPPORTED, "internal: no scratch for a multi-pass plan");
    const int z = free[0];
    need_scratch(L, z, n * batch);
    // Pass d (P down to 1) transforms digit j_d of j = j1 + N1*(j2 + N2*(...)):
    // for each low index jl < S_d and processed outputs K_d < M_d it reads
    // jl + S_d*(j_d + N_d*K_d), writes jl + S_d*(K_d + M_d*k_d), and (d >= 2)
    // multiplies by w_{N_{d-1} M_{d-1}}^{j_{d-1}(K_d + M_d k_d)} -- the nested
    // four-step of fourstep.jl:1-17 with every level's twiddle fused into
    // the pass that produces it.  Output digits end as k = kP + NP*(...+ N2*k1).
    const int P = static_cast<int>(F.size());
    std::vector<int64_t> S(P + 1, 1), M(P + 2, 1);
    for (int d = 2; d <= P; ++d) S[d] = S[d - 1] * F[d - 2];
    for (int d = P - 1; d >= 0; --d) M[d] = M[d + 1] * F[d];  // M[d] = prod F[d..P-1] (0-based)
    // (measured on 2^24: the in-place middle pass runs from the Infinity
    // Cache, 70 -> 55 us; FFTJH_MID_INPLACE=0 restores the rotating layout)
    static const bool mid_inplace = [] {
      const char* e = std::getenv("FFTJH_MID_INPLACE");
      return !e || std::atoi(e) != 0;
    }();
    // rows (cs = 1) or a column block of cs = Bi columns: every offset of
    // the row schedule below scales by cs, and the column joins the low
    // digits, b % (cs*S_d) = column + cs*jl
    const int64_t cs = colblock ? Bi : 1;
    if (P == 3 && cs == 1 && lean3_ok(L, F, n, batch, src, dst)) {
      add_lean3(L, F, n, batch, dir, scale, src, dst, z);
      return;
    }
    if (P == 3 && mid_inplace && cs == 1) {
      // Three passes through one intermediate z: x -> z (digit j3 -> k3 in
      // place), z -> z (j2 -> k2 in place, twiddle w_n^{j1 (k3 + N3 k2)}),
      // z -> y (rows j1 at j1 + N1 (k2 + N2 k3), written to natural order
      // k3 + N3 k2 + N2 N3 k1).  The middle pass works inside one array.
      const int64_t N1 = F[0], N2 = F[1], N3 = F[2];
      PassTw t3;
      t3.period = N2 * N3;
      t3.qdiv = N1;
      t3.acoef = 1;
      t3.bcoef = 1;
      add_pass(L, dir, static_cast<int>(N3), batch * (n / N3), N1 * N2, geo(N1 * N2, 1, n, n), geo(N1 * N2, 1, n, n),
               1.0, src, z, t3, 1);
      PassTw t2;
      t2.period = n;
      t2.qdiv = 1;
      t2.acoef = 1;
      t2.bcoef = N3;
      add_pass(L, dir, static_cast<int>(N2), batch * (n / N2), N1, geo(N1, 1, n, N1 * N2), geo(N1, 1, n, N1 * N2), 1.0,
               z, z, t2, N3);
      add_pass(L, dir, static_cast<int>(N1), batch * (n / N1), N3, geo(1, N1 * N2, n, N1), geo(N2 * N3, 1, n, N3),
               scale, z, dst, PassTw(), N2);
      return;
    }
    int cur = src;
    for (int d = P; d >= 1; --d) {
      const int64_t Nd = F[d - 1];
      const int64_t Sd = S[d];
      const int64_t Md = M[d];  // prod of F[d..P-1] = N_{d+1}...N_P
      const int target = (d == 1) ? dst : ((d % 2 == 1) ? dst : z);
      // the first pass reads x with the same geometry it writes, so it may run in place
      const int out = (d == P && P % 2 == 1) ? dst : target;
      PassTw tw;
      if (d >= 2) {
        tw.period = F[d - 2] * M[d - 1];
        tw.qdiv = S[d - 1] * cs;
        tw.acoef = 1;
        tw.bcoef = Md;
      }
      add_pass(L, dir, static_cast<int>(Nd), batch * (n / Nd), Sd * cs, geo(Sd * cs, 1, n * cs, Sd * Nd * cs),
               geo(Sd * Md * cs, 1, n * cs, Sd * cs), d == 1 ? scale : 1.0, cur, out, tw, Md);
      cur = out;
    }
    return;
  }
  // not 13-smooth, n = p*s with s its 13-smooth part: one exact
  // Cooley-Tukey step around the resampled length p (composite.jl:1-101),
  //   X[k2 + s*k1] = sum_j1 w_p^(j1 k1) w_n^(j1 k2) sum_j2 w_s^(j2 k2) x[j1 + p*j2]:
  // a pass of p length-s transforms at stride p with the twiddle w_n^(j1 k2)
  // at its store, in place into dst, then the s length-p transforms over j1
  // resampled from dst (rows k2) back into dst at stride s (outputs k1)
  {
    int64_t sm = 1;
    for (int f : {2, 3, 5, 7, 11, 13})
      while ((n / sm) % f == 0) sm *= f;
    const int64_t p = n / sm;
    const bool ok = sm > 1 && is_rows(gin, n, Bi) && is_rows(gout, n, Bi) && single_pass_ok(L.dtype, sm);
    // the HvdH route on a length it cannot take whole (even n: no grid
    // coprime to it) takes the same step around length-p HvdH transforms,
    // whose output permutation writes at stride s: the reference's SplitPlan
    // (composite.jl:103-200, length-s combine after the length-p HvdH; here
    // the length-s step runs first, as above, so that the HvdH transforms
    // read contiguous rows)
    bool split_hv = false;
    if (ok && (L.route & FFTJH_FLAG_ROUTE_HVDH)) {
      HvdhDesc whole, part;
      split_hv = !hvdh_params(L.dtype, n, batch, scale, whole) && hvdh_params(L.dtype, p, batch * sm, scale, part);
    }
    if (ok && (split_hv || use_composite(L, n, sm))) {
      PassTw t;
      t.period = n;
      t.qdiv = 1;
      t.acoef = 0;
      t.bcoef = 1;
      add_pass(L, dir, static_cast<int>(sm), batch * p, p, geo(p, 1, n), geo(p, 1, n), 1.0, src, dst, t);
      build_c2c(L, p, batch * sm, sm, geo(1, p, n), geo(sm, 1, n), dir, scale, dst, dst, free);
      return;
    }
  }
  // the HvdH route (hvdh.jl:222-241): x permuted to Delta order, spread onto
  // grid rows of t, the t-point transform in place, the chunked banded
  // solves to the unknowns in l order, permuted into y
  const bool rows_in = gin.stride == 1 && gin.s_mid == 0 &&
                       (Bi == 1 ? gin.s_out == n : gin.s_in == n && gin.s_out == n * Bi);
  if (use_hvdh(L, n) && rows_in && gout.s_mid == 0 && free.size() >= 3) {
    HvdhDesc hd;
    if (hvdh_params(L.dtype, n, batch, scale, hd)) {
      const int g = free[0], u = free[1];
      need_scratch(L, g, hd.t * batch);
      need_scratch(L, u, n * batch);
      Step st;
      st.dir = dir;
      st.n = n;
      st.m = hd.t;
      st.batch = batch;
      st.hv = hd;
      Step pin = st;
      pin.kind = ST_HVDH_PERMUTE;
      pin.src = src;
      pin.dst = u;
      pin.scale = 1.0;
      L.steps.push_back(pin);
      Step sp = st;
      sp.kind = ST_HVDH_SPREAD;
      sp.src = u;
      sp.dst = g;
      L.steps.push_back(sp);
      build_c2c(L, hd.t, batch, 1, geo(1, 0, hd.t), geo(1, 0, hd.t), dir, 1.0, g, g,
                std::vector<int>(free.begin() + 2, free.end()));
      Step so = st;
      so.kind = ST_HVDH_SOLVE;
      so.src = g;
      so.dst = u;
      L.steps.push_back(so);
      Step pout = st;
      pout.kind = ST_HVDH_PERMUTE;
      pout.src = u;
      pout.dst = dst;
      pout.scale = scale;
      pout.hv.o_stride = gout.stride;
      pout.hv.o_bi = Bi;
      pout.hv.o_in = gout.s_in;
      pout.hv.o_out = gout.s_out;
      L.steps.push_back(pout);
      return;
    }
    if (L.route & FFTJH_FLAG_ROUTE_HVDH)
      fail(FFTJH_ERR_UNSUPPORTED, "the HvdH route needs FP32, contiguous rows and an odd length with a coprime "
                                  "smooth grid, or a length whose odd part after its 13-smooth part has one "
                                  "(length " + std::to_string(n) + ")");
  }
  // otherwise resample onto a smooth m (resample.jl:258-276)
  if (free.size() < 2) fail(FFTJH_ERR_UNSUPPORTED, "internal: no scratch for resampling");
  const RsParams rp = rs_params(L.dtype, n);
  const int g = free[0];
  need_scratch(L, g, rp.m * batch);
  {
    // FP32 grids of three lean factors: the lean layout from x into the grid
    // (pass A reads x through the deconvolution, passes B and C in place on
    // the grid), then the interpolation.  FFTJH_RS_LEAN=0: the pass kernel.
    static const bool rs_lean = [] {
      const char* e = std::getenv("FFTJH_RS_LEAN");
      return !e || std::atoi(e) != 0;
    }();
    fftjh_tuning t = L.tune;
    t.flags &= ~(FFTJH_FLAG_MFMA_F16 | FFTJH_FLAG_MFMA_BF16 | FFTJH_FLAG_MFMA_BF16X3);
    const std::vector<int64_t> F = rp.m > 16384 ? choose_factors(L.dtype, rp.m, default_max_factor(L.dtype))
                                                : std::vector<int64_t>{};
    bool lean_ok = rs_lean && L.dtype == FFTJH_F32 && !L.mat && tune_is_default(t) && is_rows(gin, n, Bi) &&
                   gin.stride == 1 && rp.m < (int64_t(1) << 31) && F.size() == 3 &&
                   batch * (rp.m / 1024) < (int64_t(1) << 31);
    for (int64_t f : F) lean_ok = lean_ok && lean_length(static_cast<int>(f));
    if (lean_ok) {
      const size_t first = L.steps.size();
      add_lean3(L, F, rp.m, batch, dir, 1.0, src, g, g);
      Step& a = L.steps[first];
      a.ld.rs = 1;
      a.ld.rs_n = n;
      a.ld.rs_m = rp.m;
      a.ld.nt = 0;
      a.n = n;  // the correction table's length and parameters
      a.m = rp.m;
      a.w = rp.w;
      a.beta = rp.beta;
      a.scale = scale;
      Step ip;
      ip.kind = ST_RS_INTERP;
      ip.src = g;
      ip.dst = dst;
      ip.dir = dir;
      ip.n = n;
      ip.m = rp.m;
      ip.batch = batch;
      ip.w = rp.w;
      ip.beta = rp.beta;
      ip.Bi = Bi;
      ip.gb = gout;
      L.steps.push_back(ip);
      return;
    }
  }
  {
    // the deconvolution at the load of the m-point transform's first pass,
    // when that is a power-of-two pass over the grid rows and x is rows
    // (FFTJH_RS_FUSE=0: its own launch)
    static const bool rs_fuse = [] {
      const char* e = std::getenv("FFTJH_RS_FUSE");
      return !e || std::atoi(e) != 0;
    }();
    // (the default split first, then passes of at most 256 points)
    for (int attempt = 0; attempt < 2 && rs_fuse && is_rows(gin, n, Bi) && gin.stride == 1 &&
                          rp.m < (int64_t(1) << 31);
         ++attempt) {
      Layout trial = L;
      trial.two_pass512 = attempt == 0;
      const size_t first = trial.steps.size();
      build_c2c(trial, rp.m, batch, 1, geo(1, 0, rp.m), geo(1, 0, rp.m), dir, 1.0, g, g,
                std::vector<int>(free.begin() + 1, free.end()));
      Step& f = trial.steps[first];
      if (f.kind == ST_PASS && f.pow2 && (f.d.n & (f.d.n - 1)) == 0 && !f.mat && f.src == g && !f.d.rows_in &&
          !f.d.real_io &&
          f.d.in.s_out == rp.m && f.d.batch == batch * (rp.m / f.d.n)) {
        trial.two_pass512 = L.two_pass512;
        f.src = src;
        f.d.rs_io = 1;
        f.d.rs_n = n;
        f.d.rs_m = rp.m;
        f.n = n;
        f.m = rp.m;
        f.w = rp.w;
        f.beta = rp.beta;
        f.scale = scale;
        Step ip;
        ip.kind = ST_RS_INTERP;
        ip.src = g;
        ip.dst = dst;
        ip.dir = dir;
        ip.n = n;
        ip.m = rp.m;
        ip.batch = batch;
        ip.w = rp.w;
        ip.beta = rp.beta;
        ip.Bi = Bi;
        ip.gb = gout;
        trial.steps.push_back(ip);
        L = std::move(trial);
        return;
      }
    }
  }
  Step dc;
  dc.kind = ST_RS_DECONV;
  dc.src = src;
  dc.dst = g;
  dc.dir = dir;
  dc.n = n;
  dc.m = rp.m;
  dc.batch = batch;
  dc.w = rp.w;
  dc.beta = rp.beta;
  dc.scale = scale;
  dc.Bi = Bi;
  dc.ga = gin;
  L.steps.push_back(dc);
  std::vector<int> rest(free.begin() + 1, free.end());
  build_c2c(L, rp.m, batch, 1, geo(1, 0, rp.m), geo(1, 0, rp.m), dir, 1.0, g, g, rest);
  Step ip;
  ip.kind = ST_RS_INTERP;
  ip.src = g;
  ip.dst = dst;
  ip.dir = dir;
  ip.n = n;
  ip.m = rp.m;
  ip.batch = batch;
  ip.w = rp.w;
  ip.beta = rp.beta;
  ip.Bi = Bi;
  ip.gb = gout;
  L.steps.push_back(ip);
}

static void build_c2c_rows(Layout& L, int64_t n, int64_t batch, int dir, double scale, int src, int dst,
                           std::vector<int> free) {
  build_c2c(L, n, batch, 1, geo(1, 0, n), geo(1, 0, n), dir, scale, src, dst, std::move(free));
}

static void check_dtype(int dtype) {
  if (dtype != FFTJH_F32 && dtype != FFTJH_F64)
    fail(FFTJH_ERR_ARGUMENT, "dtype must be FFTJH_F32 or FFTJH_F64");
}

// split-precision mode from plan flags (at most one FFTJH_FLAG_MFMA_* bit; FP32 only)
static const unsigned kRouteFlags = FFTJH_FLAG_ROUTE_COMPOSITE | FFTJH_FLAG_ROUTE_RESAMPLE | FFTJH_FLAG_ROUTE_HVDH;

static int mat_mode_of(unsigned flags, int dtype) {
  const unsigned known = FFTJH_FLAG_MFMA_F16 | FFTJH_FLAG_MFMA_BF16 | FFTJH_FLAG_MFMA_BF16X3;
  if (flags & ~(known | kRouteFlags)) fail(FFTJH_ERR_ARGUMENT, "unknown plan flags");
  {
    const unsigned r = flags & kRouteFlags;
    if (r & (r - 1)) fail(FFTJH_ERR_ARGUMENT, "at most one route flag");
  }
  flags &= known;
  if (!flags) return kMatNone;
  if (flags & (flags - 1)) fail(FFTJH_ERR_ARGUMENT, "at most one split-precision flag");
  if (dtype != FFTJH_F32) fail(FFTJH_ERR_ARGUMENT, "split-precision MFMA stages are for FP32 plans");
  return flags == FFTJH_FLAG_MFMA_F16 ? kMatF16 : flags == FFTJH_FLAG_MFMA_BF16 ? kMatBF16 : kMatBF16x3;
}

static Layout layout_c2c(int dtype, int rank, const int64_t* dims, int64_t batch, int sign, double scale,
                         const fftjh_tuning* tune = nullptr, unsigned flags = 0) {
  check_dtype(dtype);
  if (rank < 1 || rank > 8) fail(FFTJH_ERR_ARGUMENT, "rank must be in 1..8");
  if (!dims) fail(FFTJH_ERR_ARGUMENT, "dims is NULL");
  if (batch < 1) fail(FFTJH_ERR_ARGUMENT, "batch must be positive (got " + std::to_string(batch) + ")");
  if (sign != -1 && sign != 1) fail(FFTJH_ERR_ARGUMENT, "sign must be -1 or +1");
  int64_t len = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] < 1)
      fail(FFTJH_ERR_ARGUMENT, rank == 1 ? "transform length must be positive (got " + std::to_string(dims[i]) + ")"
                                         : "dimensions must be positive");
    len *= dims[i];
  }
  Layout L;
  if (tune) L.tune = *tune;
  L.mat = mat_mode_of(flags | (tune ? tune->flags : 0u), dtype);
  L.route = (flags | (tune ? tune->flags : 0u)) & kRouteFlags;
  L.dtype = dtype;
  L.kind = FFTJH_KIND_C2C;
  L.sign = sign;
  L.rank = rank;
  for (int i = 0; i < rank; ++i) L.dims[i] = dims[i];
  L.batch = batch;
  L.scale = scale;
  L.in_elems = L.out_elems = len * batch;
  if (rank == 1) {
    L.lean = true;
    build_c2c_rows(L, dims[0], batch, sign, scale, RX, RY, {RS0, RS1, RS2, RS3});
    return L;
  }
  // N-d: one strided axis pass per dim of length > 1 (ndplan.jl:57-118,
  // _nd_core! :270-327), x -> y then in place on y; scale on the last pass
  std::vector<int> axes;
  for (int i = 0; i < rank; ++i)
    if (dims[i] > 1) axes.push_back(i);
  if (axes.empty()) {
    add_pass(L, sign, 1, batch, 1, geo(1, 0, 1), geo(1, 0, 1), scale, RX, RY);
    return L;
  }
  int src = RX;
  for (size_t ai = 0; ai < axes.size(); ++ai) {
    const int ax = axes[ai];
    int64_t Bi = 1, Bo = batch;
    for (int i = 0; i < ax; ++i) Bi *= dims[i];
    for (int i = ax + 1; i < rank; ++i) Bo *= dims[i];
    const int64_t nd = dims[ax];
    const double sc = (ai + 1 == axes.size()) ? scale : 1.0;
    build_c2c(L, nd, Bi * Bo, Bi, geo(Bi, 1, Bi * nd), geo(Bi, 1, Bi * nd), sign, sc, src, RY,
              {RS0, RS1, RS2, RS3});
    src = RY;
  }
  return L;
}

// The edges of an even-length real transform fused into the last (R2C,
// real_io 3) or first (C2R, real_io 4) pass of its multi-pass half-length
// transform of h points, `rows` rows: that pass must be a power-of-two pass
// whose transform b = T + M*row (M = h/n per row, a multiple of 32) sits at
// T + row*h on the edge side with element stride M, and 32 of them must fit
// one workgroup.  The edge side's row stride becomes h + 1.  On success the
// step is rewritten; FFTJH_REAL_PAIRED=0 keeps the edges as launches of
// their own.
static bool pair_real_pass(Layout& L, Step& s, int64_t h, int64_t rows, int real_io, double scale) {
  static const bool off = [] {
    const char* e = std::getenv("FFTJH_REAL_PAIRED");
    return e && std::atoi(e) == 0;
  }();
  if (off || s.kind != ST_PASS || !s.pow2 || s.mat || (s.d.n & (s.d.n - 1))) return false;
  PassDesc& d = s.d;
  if (h % d.n) return false;
  const int64_t M = h / d.n;
  if (M % 32 || M > (int64_t(1) << 30) || d.batch != M * rows || (real_io == 3 && d.tw_period != 0)) return false;
  Geometry& g = real_io == 3 ? d.out : d.in;
  const bool rowmap = g.stride == M && g.s_out == h &&
                      ((d.Bi == 1 && d.B2 == M && g.s_mid == 1) ||
                       (d.Bi * d.B2 == M && g.s_in == 1 && (d.B2 == 1 || g.s_mid == d.Bi)));
  if (!rowmap) return false;
  int t;
  if (!pow2_fits(L.dtype, d.n, 32, &t)) return false;
  g.s_out = h + 1;
  d.C = 32;
  d.logC = 5;
  d.nthreads = t;
  d.ngroups = static_cast<int>(d.batch / 32);
  d.rows_in = d.rows_out = 0;
  d.real_io = real_io;
  d.pair_M = static_cast<int32_t>(M);
  d.real_h = h;
  d.scale = scale;
  int bits = 0;
  while ((int64_t(1) << bits) < 2 * h) ++bits;
  d.rtw_sh = (bits + 1) / 2;
  d.rtw_hi = ((2 * h - 1) >> d.rtw_sh) + 1;  // entries of the high table
  return true;
}

// Real-to-complex rows: `batch` rows of n reals at src (row stride n) ->
// rows of n/2+1 complex at dst (row stride n/2+1), through the scratch roles
// in `free` (front first; at least four).  Even n: the half-length complex
// transform of the reals read as n/2 complex points (real.jl:88-131); odd n:
// the full complex transform of the widened reals, prefix kept.
static void add_r2c_rows(Layout& L, int64_t n, int64_t batch, int src, int dst, const std::vector<int>& free) {
  if (n % 2 == 0 && !L.mat && real_fused_C(L.dtype, n / 2, batch, nullptr)) {
    // one launch: the reals read as n/2 complex points, the half-length
    // transform, the twist at the store
    const int64_t h = n / 2;
    add_pass(L, -1, static_cast<int>(h), batch, 1, geo(1, 0, h), geo(1, 0, h + 1), 1.0, src, dst, PassTw(), 1, 1);
  } else if (n % 2 == 0) {
    const int64_t h = n / 2;
    // the twist at the store of the half-length transform's last pass (with
    // the default split, else with passes of at most 256 points, which the
    // paired kernel takes in either precision)
    for (int attempt = 0; attempt < 2; ++attempt) {
      Layout trial = L;
      trial.two_pass512 = attempt == 0;
      build_c2c_rows(trial, h, batch, -1, 1.0, src, dst, free);
      if (pair_real_pass(trial, trial.steps.back(), h, batch, 3, 1.0)) {
        trial.two_pass512 = L.two_pass512;
        L = std::move(trial);
        return;
      }
    }
    need_scratch(L, free[0], h * batch);
    build_c2c_rows(L, h, batch, -1, 1.0, src, free[0], std::vector<int>(free.begin() + 1, free.end()));
    Step t;
    t.kind = ST_R2C_TWIST;
    t.src = free[0];
    t.dst = dst;
    t.dir = -1;
    t.n = n;
    t.h = h;
    t.batch = batch;
    t.s1 = h;
    t.s2 = h + 1;
    L.steps.push_back(t);
  } else {
    need_scratch(L, free[0], n * batch);
    need_scratch(L, free[1], n * batch);
    Step e;
    e.kind = ST_REAL_TO_COMPLEX;
    e.src = src;
    e.dst = free[0];
    e.n = n;
    e.batch = batch;
    L.steps.push_back(e);
    build_c2c_rows(L, n, batch, -1, 1.0, free[0], free[1], std::vector<int>(free.begin() + 2, free.end()));
    Step c;
    c.kind = ST_COPY_PREFIX;
    c.src = free[1];
    c.dst = dst;
    c.n = n;
    c.keep = n / 2 + 1;
    c.batch = batch;
    L.steps.push_back(c);
  }
}

// Complex-to-real rows: `batch` rows of n/2+1 complex at src -> rows of n
// reals at dst, scaled by 1/n when normalized (irfft), else 1 (brfft).  src
// is read once, by the first step, so a scratch src may also be listed in
// `free` behind the roles that step writes.
static void add_c2r_rows(Layout& L, int64_t n, int64_t batch, bool normalized, int src, int dst,
                         const std::vector<int>& free) {
  if (n % 2 == 0 && !L.mat && real_fused_C(L.dtype, n / 2, batch, nullptr)) {
    // one launch: the pre-pass at the load, the inverse half-length
    // transform, the reals stored as n/2 complex points; the scale carries
    // the pre-factor (1/2 * 1/h = 1/n for irfft, 1 for brfft)
    const int64_t h = n / 2;
    add_pass(L, 1, static_cast<int>(h), batch, 1, geo(1, 0, h + 1), geo(1, 0, h),
             normalized ? 1.0 / static_cast<double>(n) : 1.0, src, dst, PassTw(), 1, 2);
  } else if (n % 2 == 0) {
    const int64_t h = n / 2;
    // the pre-pass at the load of the half-length transform's first pass
    // (its store carries the pre-factor, the last pass the 1/h); default
    // split first, then passes of at most 256 points
    for (int attempt = 0; attempt < 2; ++attempt) {
      Layout trial = L;
      trial.two_pass512 = attempt == 0;
      const size_t first = trial.steps.size();
      build_c2c_rows(trial, h, batch, 1, normalized ? 1.0 / static_cast<double>(h) : 1.0, src, dst, free);
      if (trial.steps.size() > first + 1 &&
          pair_real_pass(trial, trial.steps[first], h, batch, 4, normalized ? 0.5 : 1.0)) {
        trial.two_pass512 = L.two_pass512;
        L = std::move(trial);
        return;
      }
    }
    need_scratch(L, free[0], h * batch);
    Step p;
    p.kind = ST_C2R_PRE;
    p.src = src;
    p.dst = free[0];
    p.dir = 1;
    p.n = n;
    p.h = h;
    p.batch = batch;
    p.s1 = h + 1;
    p.s2 = h;
    p.pre = normalized ? 0.5 : 1.0;
    L.steps.push_back(p);
    // inner inverse of length h: scale 1/h (irfft) or 1 (brfft), real.jl:64-65
    build_c2c_rows(L, h, batch, 1, normalized ? 1.0 / static_cast<double>(h) : 1.0, free[0], dst,
                   std::vector<int>(free.begin() + 1, free.end()));
  } else {
    need_scratch(L, free[0], n * batch);
    need_scratch(L, free[1], n * batch);
    Step f;
    f.kind = ST_HERMITIAN_FILL;
    f.src = src;
    f.dst = free[0];
    f.n = n;
    f.batch = batch;
    L.steps.push_back(f);
    build_c2c_rows(L, n, batch, 1, normalized ? 1.0 / static_cast<double>(n) : 1.0, free[0], free[1],
                   std::vector<int>(free.begin() + 2, free.end()));
    Step t;
    t.kind = ST_TAKE_REAL;
    t.src = free[1];
    t.dst = dst;
    t.n = n;
    t.batch = batch;
    L.steps.push_back(t);
  }
}

// Strided complex passes along dims[1..rank-1] of the column-major array
// cdims (dims[0] is the batch within each pass), src -> dst then in place on
// dst (ndplan.jl:57-118), `scale` on the last pass.  Lays out nothing when
// every one of those dims is 1.
static void add_axis_passes(Layout& L, int rank, const int64_t* cdims, int64_t batch, int sign, double scale, int src,
                            int dst, const std::vector<int>& free) {
  std::vector<int> axes;
  for (int i = 1; i < rank; ++i)
    if (cdims[i] > 1) axes.push_back(i);
  for (size_t ai = 0; ai < axes.size(); ++ai) {
    const int ax = axes[ai];
    int64_t Bi = 1, Bo = batch;
    for (int i = 0; i < ax; ++i) Bi *= cdims[i];
    for (int i = ax + 1; i < rank; ++i) Bo *= cdims[i];
    const int64_t nd = cdims[ax];
    const double sc = (ai + 1 == axes.size()) ? scale : 1.0;
    build_c2c(L, nd, Bi * Bo, Bi, geo(Bi, 1, Bi * nd), geo(Bi, 1, Bi * nd), sign, sc, ai == 0 ? src : dst, dst, free);
  }
}

static void check_real_request(int dtype, int rank, const int64_t* dims, int64_t batch) {
  check_dtype(dtype);
  if (rank < 1 || rank > 8) fail(FFTJH_ERR_ARGUMENT, "rank must be in 1..8");
  if (!dims) fail(FFTJH_ERR_ARGUMENT, "dims is NULL");
  for (int i = 0; i < rank; ++i)
    if (dims[i] < 1)
      fail(FFTJH_ERR_ARGUMENT, rank == 1 ? "transform length must be positive (got " + std::to_string(dims[i]) + ")"
                                         : "dimensions must be positive");
  if (batch < 1) fail(FFTJH_ERR_ARGUMENT, "batch must be positive");
}

// R2C over all dims of the column-major real array dims (dims[0] halved):
// real rows along dims[0] into y, then the complex passes of the other dims
// in place on y (ndplan.jl:355-418).  rank 1 is the 1-D plan (real.jl:41-52).
static Layout layout_r2c(int dtype, int rank, const int64_t* dims, int64_t batch) {
  check_real_request(dtype, rank, dims, batch);
  Layout L;
  L.dtype = dtype;
  L.kind = FFTJH_KIND_R2C;
  L.sign = -1;
  L.rank = rank;
  int64_t rest = 1;
  for (int i = 0; i < rank; ++i) L.dims[i] = dims[i];
  for (int i = 1; i < rank; ++i) rest *= dims[i];
  L.batch = batch;
  const int64_t n1 = dims[0], h1p = n1 / 2 + 1;
  L.in_elems = n1 * rest * batch;
  L.out_elems = h1p * rest * batch;
  add_r2c_rows(L, n1, rest * batch, RX, RY, {RS0, RS1, RS2, RS3});
  int64_t cdims[8];
  for (int i = 0; i < rank; ++i) cdims[i] = i == 0 ? h1p : dims[i];
  add_axis_passes(L, rank, cdims, batch, -1, 1.0, RY, RY, {RS0, RS1, RS2, RS3});
  return L;
}

// C2R over all dims (ndplan.jl:435-521): the inverse complex passes of
// dims[1..] from x into scratch (x is left intact), then the real rows along
// dims[0] into y.  normalized scales by 1/prod(dims) (irfft), else 1 (brfft).
static Layout layout_c2r(int dtype, int rank, const int64_t* dims, int64_t batch, int normalized) {
  check_real_request(dtype, rank, dims, batch);
  Layout L;
  L.dtype = dtype;
  L.kind = FFTJH_KIND_C2R;
  L.sign = 1;
  L.rank = rank;
  int64_t rest = 1;
  for (int i = 0; i < rank; ++i) L.dims[i] = dims[i];
  for (int i = 1; i < rank; ++i) rest *= dims[i];
  L.batch = batch;
  const int64_t n1 = dims[0], h1p = n1 / 2 + 1;
  L.scale = normalized ? 1.0 / (static_cast<double>(n1) * static_cast<double>(rest)) : 1.0;
  L.in_elems = h1p * rest * batch;
  L.out_elems = n1 * rest * batch;
  int64_t cdims[8];
  for (int i = 0; i < rank; ++i) cdims[i] = i == 0 ? h1p : dims[i];
  // irfft: 1/prod(dims[1..]) on the last complex pass, 1/n1 on the rows
  const double rest_scale = normalized ? 1.0 / static_cast<double>(rest) : 1.0;
  bool passes = false;
  for (int i = 1; i < rank; ++i) passes |= dims[i] > 1;
  if (passes) {
    need_scratch(L, RS0, h1p * rest * batch);
    add_axis_passes(L, rank, cdims, batch, 1, rest_scale, RX, RS0, {RS1, RS2, RS3});
    add_c2r_rows(L, n1, rest * batch, normalized != 0, RS0, RY, {RS1, RS2, RS3, RS0});
  } else {
    add_c2r_rows(L, n1, rest * batch, normalized != 0, RX, RY, {RS0, RS1, RS2, RS3});
  }
  return L;
}

// ----------------------------------------------------------- ragged --

// Host side of one ragged launch.  Tables are per plan (the reference keeps
// no ragged state at all; its baseline is a loop of 1-D plans).
struct RaggedData {
  int dir = -1;
  int groups = 0;
  int radix_pref = 16;     // 16: radix-16 towers; 8: radix-8 towers
  int tile_max = kRaggedTile;
  int big = 0;                      // 1: the large-item kernel (one workgroup per CU)
  int threads = kRaggedThreads;     // its workgroup size
  std::map<int64_t, int64_t> corr_of;  // resampled length -> offset in the correction pool
  std::vector<RaggedItem> items;
  std::vector<RaggedSched> scheds;
  std::map<int, int> sched_of;  // tile length -> schedule
  std::vector<int32_t> wg_ptr, wg_items;
  std::vector<int64_t> wg_ws;
  int64_t ws_total = 0;
  int64_t tw_total = 0, corr_total = 0;
  struct StageJob { int64_t off; int ns, R; };
  struct PowJob { int64_t off, count, mult, L; };
  struct CorrJob { int64_t off, n, m; int w; double beta, scale; };
  std::vector<StageJob> stage_jobs;
  std::vector<PowJob> pow_jobs;
  std::vector<CorrJob> corr_jobs;
  std::vector<double> cost;     // per item
  int64_t counts[4] = {0, 0, 0, 0};
  std::vector<RaggedTile> tiles_a, tiles_b;  // tile-granular passes: phase A and phase B tiles
  // tile-granular passes: FP64 two-level tables of w_n per length (hi, lo offsets in double2)
  std::map<int64_t, std::pair<int64_t, int64_t>> t64_of;
  int64_t tw64_total = 0;
  std::vector<PowJob> pow64_jobs;
  DevBufPtr d_items, d_wgptr, d_wgitems, d_scheds, d_tw, d_corr, d_ws, d_wgws, d_tiles_a, d_tiles_b, d_tw64;
};

static std::vector<int> radices_pref(int64_t n, int pref) {
  int a, cnt[5];
  std::vector<int> r;
  if (!factor13(n, a, cnt)) return r;
  const int big = pref == 8 ? 3 : 4;
  if (a % big) r.push_back(1 << (a % big));
  for (int i = 0; i < a / big; ++i) r.push_back(1 << big);
  for (int i = 0; i < 5; ++i)
    for (int k = 0; k < cnt[i]; ++k) r.push_back(kOddRadices[i]);
  return r;
}

static int ragged_tile_C(const RaggedData& R, int64_t L, const std::vector<int>& rad);

// The stage chain of a tile length.  The tile-granular passes take the
// shortest chain with the composite radices 6..15 (140 = 14 x 10 instead of
// 4 x 5 x 7: fewer LDS round trips per tile) where it fits as many columns
// per tile as the prime chain; FFTJH_RAGGED_COMPACT=0 keeps the prime chains.
static std::vector<int> ragged_radices(const RaggedData& R, int64_t L) {
  static const bool compact_env = [] {
    const char* e = std::getenv("FFTJH_RAGGED_COMPACT");
    return !e || std::atoi(e) != 0;
  }();
  std::vector<int> rad = radices_pref(L, R.radix_pref);
  if (!compact_env || R.big != 2 || rad.size() < 3 || L > R.tile_max) return rad;
  std::vector<int> rc = compact_radices(L);
  if (rc.empty() || rc.size() >= rad.size()) return rad;
  int cp = 0, cc = 0;
  for (int C = 64; C >= 1 && !cp; C >>= 1)
    if (C * L <= R.tile_max && pass_threads(static_cast<int>(L), C, rad) <= R.threads) cp = C;
  for (int C = 64; C >= 1 && !cc; C >>= 1)
    if (C * L <= R.tile_max && pass_threads(static_cast<int>(L), C, rc) <= R.threads) cc = C;
  return cc >= cp ? rc : rad;
}

// largest power-of-two number of length-L columns one tile holds
static int ragged_tile_C(const RaggedData& R, int64_t L, const std::vector<int>& rad) {
  for (int C = 64; C >= 1; C >>= 1) {
    if (C * L > R.tile_max) continue;
    if (pass_threads(static_cast<int>(L), C, rad) > R.threads) continue;
    return C;
  }
  return 0;
}

static bool ragged_direct_ok(const RaggedData& R, int64_t L) {
  if (L == 1) return true;  // no stages: a scaled copy
  if (L > R.tile_max) return false;
  auto rad = ragged_radices(R, L);
  if (rad.empty() || static_cast<int>(rad.size()) > kMaxStages) return false;
  return ragged_tile_C(R, L, rad) >= 1;
}

static int ragged_sched(RaggedData& R, int64_t L) {
  auto it = R.sched_of.find(static_cast<int>(L));
  if (it != R.sched_of.end()) return it->second;
  RaggedSched sc{};
  sc.n = static_cast<int>(L);
  auto rad = ragged_radices(R, L);
  sc.nstages = static_cast<int>(rad.size());
  int64_t ns = 1;
  for (size_t i = 0; i < rad.size(); ++i) {
    sc.st[i].radix = rad[i];
    sc.st[i].ns = static_cast<int>(ns);
    sc.st[i].tw_off = static_cast<int>(R.tw_total);
    sc.st[i].ns_magic = magic_for(static_cast<int>(ns));
    R.stage_jobs.push_back({R.tw_total, static_cast<int>(ns), rad[i]});
    R.tw_total += (rad[i] - 1) * ns;
    ns *= rad[i];
  }
  if (R.tw_total >= (int64_t(1) << 31)) fail(FFTJH_ERR_UNSUPPORTED, "ragged twiddle pool too large");
  R.scheds.push_back(sc);
  const int id = static_cast<int>(R.scheds.size()) - 1;
  R.sched_of[static_cast<int>(L)] = id;
  return id;
}

// two-level split of L with both factors direct, nearest sqrt(L)
static void ragged_two_level(RaggedData& R, RaggedItem& it, int64_t L) {
  int64_t best_n2 = 0;
  double best = 1e300;
  for (int64_t q = 2; q * q <= L * 64 && q < L; ++q) {
    if (L % q || !ragged_direct_ok(R, q) || !ragged_direct_ok(R, L / q)) continue;
    const double d = std::fabs(std::log(static_cast<double>(q) / std::sqrt(static_cast<double>(L))));
    if (d < best) {
      best = d;
      best_n2 = q;
    }
  }
  if (!best_n2) fail(FFTJH_ERR_UNSUPPORTED, "ragged length " + std::to_string(L) + " has no two-level split");
  it.n2 = static_cast<int>(best_n2);
  it.n1 = static_cast<int>(L / best_n2);
  it.sA = ragged_sched(R, it.n2);
  it.sB = ragged_sched(R, it.n1);
  auto radA = ragged_radices(R, it.n2), radB = ragged_radices(R, it.n1);
  it.cA = ragged_tile_C(R, it.n2, radA);
  it.cB = ragged_tile_C(R, it.n1, radB);
  it.logcA = 0;
  while ((1 << it.logcA) < it.cA) ++it.logcA;
  it.logcB = 0;
  while ((1 << it.logcB) < it.cB) ++it.logcB;
  int lb = 0;
  while ((int64_t(1) << (2 * lb)) < L) ++lb;
  it.lo_bits = lb;
  const int64_t K = int64_t(1) << lb, nh = (L + K - 1) / K;
  it.tw_lo = R.tw_total;
  R.pow_jobs.push_back({R.tw_total, K, 1, L});
  R.tw_total += K;
  it.tw_hi = R.tw_total;
  R.pow_jobs.push_back({R.tw_total, nh, K, L});
  R.tw_total += nh;
  if (R.big == 2) {
    // the same split in FP64 for the tile passes' twiddle start and steps
    auto t = R.t64_of.find(L);
    if (t == R.t64_of.end()) {
      const int64_t lo = R.tw64_total;
      R.pow64_jobs.push_back({lo, K, 1, L});
      R.pow64_jobs.push_back({lo + K, nh, K, L});
      R.tw64_total += K + nh;
      t = R.t64_of.emplace(L, std::make_pair(lo + K, lo)).first;
    }
    it.t64_hi = t->second.first;
    it.t64_lo = t->second.second;
  }
}

static double ragged_cost(const RaggedItem& it) {
  const double n = it.n;
  switch (it.kind) {
    case kRaggedDirect: return n * std::log2(std::max(n, 2.0)) + 2000.0;
    case kRaggedTwoLevel: return 2.6 * n * std::log2(n) + 6000.0;
    case kRaggedRsDirect: return it.m * std::log2(static_cast<double>(it.m)) + 3.0 * n * it.w + 3000.0;
    default: return 2.6 * it.m * std::log2(static_cast<double>(it.m)) + 3.0 * n * it.w + 8000.0;
  }
}

static Layout layout_ragged(int dtype, int64_t count, const int64_t* sizes, const int64_t* offsets, int sign,
                            int normalized, const fftjh_ragged_tuning* tune) {
  check_dtype(dtype);
  if (count < 1) fail(FFTJH_ERR_ARGUMENT, "ragged batch needs at least one transform");
  if (count > (int64_t(1) << 30)) fail(FFTJH_ERR_UNSUPPORTED, "too many transforms");
  if (!sizes) fail(FFTJH_ERR_ARGUMENT, "sizes is NULL");
  if (sign != -1 && sign != 1) fail(FFTJH_ERR_ARGUMENT, "sign must be -1 or +1");
  auto R = std::make_shared<RaggedData>();
  R->dir = sign;
  if (tune && tune->radix_pref) {
    if (tune->radix_pref != 8 && tune->radix_pref != 16) fail(FFTJH_ERR_ARGUMENT, "radix_pref must be 8 or 16");
    R->radix_pref = tune->radix_pref;
  }
  if (tune && tune->tile_max) {
    if (tune->tile_max < 256 || tune->tile_max > kRaggedTile)
      fail(FFTJH_ERR_ARGUMENT, "tile_max must be in [256, " + std::to_string(kRaggedTile) + "]");
    R->tile_max = tune->tile_max;
  }
  // items longer than the small kernel's tile go to the large-item kernel
  // (FFTJH_RAGGED_BIG=0: every item on the small one, as two-level tiles)
  static const bool big_env = [] {
    const char* e = std::getenv("FFTJH_RAGGED_BIG");
    return !e || std::atoi(e) != 0;
  }();
  auto R2 = std::make_shared<RaggedData>();
  R2->dir = sign;
  R2->radix_pref = R->radix_pref;
  R2->big = 1;
  R2->threads = ragged_big_threads(dtype);
  R2->tile_max = ragged_big_tile(dtype);
  // smooth items longer than pass_min run as two tile-granular passes on the
  // small kernel's tiles (phase A: every column tile of every such item in
  // one launch, phase B every row tile in a second); the others stay on the
  // item-granular kernels.  pass_min: the tuning's, else FFTJH_RAGGED_PASS_MIN,
  // else the default; negative: never.
  static const int pass_env = [] {
    const char* e = std::getenv("FFTJH_RAGGED_PASS_MIN");
    return e ? std::atoi(e) : 0;
  }();
  int pass_min = kRaggedPassMin;
  if (pass_env) pass_min = pass_env;
  if (tune && tune->pass_min) pass_min = tune->pass_min;
  auto R3 = std::make_shared<RaggedData>();
  R3->dir = sign;
  R3->radix_pref = R->radix_pref;
  R3->big = 2;
  const int64_t nmax = dtype == FFTJH_F32 ? 65536 : 32768;
  int64_t extent = 0, off = 0;
  std::vector<double> cost_small, cost_big;
  R->items.reserve(count);
  for (int64_t i = 0; i < count; ++i) {
    const int64_t n = sizes[i];
    if (n < 1) fail(FFTJH_ERR_ARGUMENT, "transform length must be positive (got " + std::to_string(n) + ")");
    if (n > nmax) fail(FFTJH_ERR_UNSUPPORTED, "ragged lengths above " + std::to_string(nmax) + " need a plan of their own");
    const int64_t o = offsets ? offsets[i] : off;
    if (o < 0) fail(FFTJH_ERR_ARGUMENT, "offsets must be non-negative");
    off = o + n;
    extent = std::max(extent, o + n);
    int a, cnt[5];
    const bool smooth = factor13(n, a, cnt);
    const RsParams rp = smooth ? RsParams{n, 0, 0.0} : rs_params(dtype, n);
    const int64_t e = smooth ? n : rp.m;  // the length one tile must hold
    const bool as_pass = smooth && pass_min >= 0 && n > pass_min && n > R3->tile_max;
    RaggedData& C = as_pass ? *R3 : (big_env && e > R->tile_max) ? *R2 : *R;
    RaggedItem it{};
    it.off = o;
    it.n = static_cast<int>(n);
    it.scale = (sign > 0 && normalized) ? 1.0 / static_cast<double>(n) : 1.0;
    if (smooth) {
      if (ragged_direct_ok(C, n)) {
        it.kind = kRaggedDirect;
        it.sA = ragged_sched(C, n);
      } else {
        it.kind = kRaggedTwoLevel;
        ragged_two_level(C, it, n);
        if (as_pass) {
          const int32_t idx = static_cast<int32_t>(C.items.size());
          it.woff = (C.ws_total + 15) & ~int64_t(15);  // each region starts on a 128-byte line
          C.ws_total = it.woff + n;
          for (int c0 = 0; c0 < it.n1; c0 += it.cA) C.tiles_a.push_back({idx, c0});
          for (int r0 = 0; r0 < it.n2; r0 += it.cB) C.tiles_b.push_back({idx, r0});
        }
      }
    } else {
      it.m = static_cast<int>(rp.m);
      it.w = rp.w;
      it.beta = rp.beta;
      auto c = C.corr_of.find(n);
      if (c == C.corr_of.end()) {
        C.corr_jobs.push_back({C.corr_total, n, rp.m, rp.w, rp.beta, it.scale});
        c = C.corr_of.emplace(n, C.corr_total).first;
        C.corr_total += n;
      }
      it.corr_off = c->second;
      if (ragged_direct_ok(C, rp.m)) {
        it.kind = kRaggedRsDirect;
        it.sA = ragged_sched(C, rp.m);
      } else {
        it.kind = kRaggedRsTwo;
        ragged_two_level(C, it, rp.m);
      }
    }
    C.counts[it.kind]++;
    C.items.push_back(it);
    // predicted per-item cost: the caller's (a learned model) or the analytic one
    C.cost.push_back((tune && tune->costs) ? tune->costs[i] : ragged_cost(it));
  }
  // workgroups: resident ones (set at materialize time), or the tuning's
  // (the small-item kernel's)
  R->groups = (tune && tune->groups > 0) ? tune->groups : 0;
  Layout L;
  L.dtype = dtype;
  L.kind = FFTJH_KIND_RAGGED;
  L.sign = sign;
  L.rank = 1;
  L.dims[0] = count;
  L.batch = 1;
  L.scale = 1.0;  // per-transform scales live in the items
  L.in_elems = L.out_elems = extent;
  Step st;
  st.kind = ST_RAGGED;
  st.src = RX;
  st.dst = RY;
  st.dir = sign;
  st.rg = R;
  if (!R2->items.empty()) st.rg2 = R2;
  if (!R3->items.empty()) st.rg3 = R3;
  L.steps.push_back(st);
  return L;
}

// XCD-local tile order for the tile-granular passes.  Workgroups are
// dispatched to the 8 XCDs round-robin by index, and each XCD has its own
// L2; packed items start anywhere, so a 16-column run of one tile shares its
// first and last 128-byte lines with the neighbouring tiles.  Laying each
// item's tiles out at indices x, x + 8, x + 16, ... (bin x, items assigned to
// bins longest first onto the least-filled) keeps them on one XCD, close in
// time, so the shared lines are fetched once.  Empty slots (unequal bins)
// hold item -1.
static std::vector<RaggedTile> xcd_tile_order(const std::vector<RaggedTile>& tiles) {
  constexpr int kX = 8;
  std::vector<std::pair<size_t, size_t>> runs;  // [begin, end) per item
  for (size_t i = 0; i < tiles.size();) {
    size_t j = i;
    while (j < tiles.size() && tiles[j].item == tiles[i].item) ++j;
    runs.push_back({i, j});
    i = j;
  }
  std::stable_sort(runs.begin(), runs.end(),
                   [](const auto& a, const auto& b) { return a.second - a.first > b.second - b.first; });
  std::vector<std::vector<RaggedTile>> bins(kX);
  std::vector<size_t> fill(kX, 0);
  for (const auto& r : runs) {
    const int x = static_cast<int>(std::min_element(fill.begin(), fill.end()) - fill.begin());
    for (size_t t = r.first; t < r.second; ++t) bins[x].push_back(tiles[t]);
    fill[x] += r.second - r.first;
  }
  const size_t slots = *std::max_element(fill.begin(), fill.end());
  std::vector<RaggedTile> out(slots * kX, RaggedTile{-1, 0});
  for (int x = 0; x < kX; ++x)
    for (size_t s = 0; s < bins[x].size(); ++s) out[s * kX + x] = bins[x][s];
  return out;
}

// longest-predicted-first onto the least-loaded workgroup (deterministic)
static void ragged_assign(RaggedData& R) {
  const int G = R.groups;
  const int64_t count = static_cast<int64_t>(R.items.size());
  std::vector<int64_t> order(count);
  for (int64_t i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return R.cost[a] > R.cost[b]; });
  std::vector<std::vector<int32_t>> lists(G);
  std::vector<int64_t> need(G, 0);
  using Slot = std::pair<double, int>;
  std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
  for (int g = 0; g < G; ++g) heap.push({0.0, g});
  for (int64_t i : order) {
    auto [load, g] = heap.top();
    heap.pop();
    lists[g].push_back(static_cast<int32_t>(i));
    const RaggedItem& it = R.items[i];
    const int64_t ws = it.kind == kRaggedTwoLevel ? it.n : it.kind == kRaggedRsTwo ? 2 * int64_t(it.m) : 0;
    need[g] = std::max(need[g], ws);
    heap.push({load + R.cost[i], g});
  }
  R.wg_ptr.assign(G + 1, 0);
  R.wg_items.clear();
  R.wg_ws.assign(G, 0);
  R.ws_total = 0;
  for (int g = 0; g < G; ++g) {
    R.wg_ptr[g] = static_cast<int32_t>(R.wg_items.size());
    R.wg_items.insert(R.wg_items.end(), lists[g].begin(), lists[g].end());
    R.wg_ws[g] = R.ws_total;
    R.ws_total += need[g];
  }
  R.wg_ptr[G] = static_cast<int32_t>(R.wg_items.size());
}

static std::string describe(const Layout& L) {
  static const char* roles[] = {"x", "y", "s0", "s1", "s2", "s3"};
  std::ostringstream o;
  o << (L.kind == FFTJH_KIND_C2C ? "c2c" : L.kind == FFTJH_KIND_R2C ? "r2c" : L.kind == FFTJH_KIND_C2R ? "c2r" : "ragged")
    << (L.dtype == FFTJH_F32 ? " f32" : " f64") << " dims=";
  for (int i = 0; i < L.rank; ++i) o << (i ? "x" : "") << L.dims[i];
  o << " batch=" << L.batch << " sign=" << L.sign << " launches=" << L.steps.size() << "\n";
  for (const Step& s : L.steps) {
    if (s.kind == ST_LEAN) {
      // a pass line (the planners' features read pass lines)
      const LeanDesc& d = s.ld;
      const int sqr = d.cols ? static_cast<int>(std::lround(std::sqrt(static_cast<double>(d.n)))) : 0;
      o << "  pass " << roles[s.src] << "->" << roles[s.dst] << " n=" << d.n << " radices="
        << (sqr ? sqr : 16) << "," << (sqr ? sqr : d.n / 16)
        << " kernel=lean" << (d.mat == kMatF16 ? " kernel=mfma_f16" : d.mat == kMatBF16 ? " kernel=mfma_bf16"
                              : d.mat == kMatBF16x3 ? " kernel=mfma_bf16x3" : "")
        << (d.rin ? " rows_in" : "") << (d.rows ? " rows_out" : "") << " batch=" << s.batch << " C="
        << (d.cols ? d.cols : d.real >= 3 ? 32 : 16) << " threads=" << (d.cols ? d.cols * sqr : d.real >= 3 ? 2 * d.n : d.n)
        << " groups=" << d.groups << " in_stride=" << (d.rin ? 1 : d.istride) << " out_stride=" << (d.rows ? 1 : d.ostride);
      if (d.psh) o << " twiddle=w_" << (d.pmod ? d.pmod : int64_t(1) << d.psh);
      if (d.scale != 1.0) o << " scale=" << d.scale;
      if (d.nt) o << " stream=" << ((d.nt & 1) ? "in" : "") << ((d.nt & 2) ? "out" : "");
      if (d.rs) o << " deconv n=" << d.rs_n << " m=" << d.rs_m << " w=" << s.w;
      if (d.real) o << (d.real == 1 ? " real=r2c" : d.real == 2 ? " real=c2r" : d.real == 3 ? " real=r2c-paired" : " real=c2r-paired");
      const int as = s.alias_src >= 0 ? s.alias_src : s.src, ad = s.alias_dst >= 0 ? s.alias_dst : s.dst;
      if (as != s.src || ad != s.dst) o << " aliased=" << roles[as] << "->" << roles[ad];
      o << "\n";
      continue;
    }
    o << "  " << kStepNames[s.kind] << " " << roles[s.src] << "->" << roles[s.dst];
    if (s.kind == ST_PASS) {
      o << " n=" << s.d.n << " radices=";
      for (size_t i = 0; i < s.radices.size(); ++i) o << (i ? "," : "") << s.radices[i];
      if (s.pow2) o << ((s.d.n & (s.d.n - 1)) == 0 ? " kernel=pow2" : " kernel=chain");
      if (s.d.fuse_io && s.radices.size() >= 2) o << " kernel=fused";
      if (s.mat) o << (s.mat == kMatF16 ? " kernel=mfma_f16" : s.mat == kMatBF16 ? " kernel=mfma_bf16" : " kernel=mfma_bf16x3");
      if (s.d.real_io) {
        static const char* kReal[] = {"", " real=r2c", " real=c2r", " real=r2c-paired", " real=c2r-paired"};
        o << kReal[s.d.real_io];
      }
      if (s.d.rows_in) o << " rows_in";
      if (s.d.rows_out) o << " rows_out";
      o << " batch=" << s.d.batch << " C=" << s.d.C << " threads=" << s.d.nthreads << " groups=" << s.d.ngroups
        << " in_stride=" << s.d.in.stride << " out_stride=" << s.d.out.stride;
      if (s.d.tw_period) o << " twiddle=w_" << s.d.tw_period;
      if (s.d.scale != 1.0) o << " scale=" << s.d.scale;
      if (s.d.rs_io) o << " deconv n=" << s.d.rs_n << " m=" << s.d.rs_m << " w=" << s.w;
      if (s.d.nt_load || s.d.nt_store) o << " stream=" << (s.d.nt_load ? "in" : "") << (s.d.nt_store ? "out" : "");
    } else if (s.kind == ST_RAGGED) {
      const RaggedData& R = *s.rg;
      const RaggedData* B = s.rg2.get();
      const RaggedData* Pp = s.rg3.get();
      auto cnt = [&](int k) { return R.counts[k] + (B ? B->counts[k] : 0) + (Pp ? Pp->counts[k] : 0); };
      o << " transforms=" << R.items.size() + (B ? B->items.size() : 0) + (Pp ? Pp->items.size() : 0)
        << " direct=" << cnt(0) << " two_level=" << cnt(1) << " resample=" << cnt(2)
        << " resample_two_level=" << cnt(3)
        << " tile_lengths=" << R.scheds.size() + (B ? B->scheds.size() : 0) + (Pp ? Pp->scheds.size() : 0)
        << " radix_pref=" << R.radix_pref << " tile_max=" << R.tile_max;
      if (R.groups) o << " workgroups=" << R.groups;
      if (B)
        o << " large_transforms=" << B->items.size() << " large_tile=" << B->tile_max << " large_direct=" << B->counts[0]
          << " large_two_level=" << B->counts[1] + B->counts[3];
      if (Pp) {
        auto live = [](const std::vector<RaggedTile>& v) {
          return std::count_if(v.begin(), v.end(), [](const RaggedTile& t) { return t.item >= 0; });
        };
        o << " pass_transforms=" << Pp->items.size() << " pass_tiles=" << live(Pp->tiles_a) << "+"
          << live(Pp->tiles_b);
      }
    } else if (s.kind == ST_HVDH_PERMUTE || s.kind == ST_HVDH_SPREAD || s.kind == ST_HVDH_SOLVE) {
      o << " n=" << s.hv.n << " t=" << s.hv.t << " taps=" << 2 * s.hv.h << " band=" << s.hv.B << " halo=" << s.hv.H
        << " batch=" << s.hv.batch;
      if (s.kind == ST_HVDH_SOLVE) o << " chunks=" << s.hv.chunks * s.hv.batch << " rows=" << hvdh_chunk_rows();
      if (s.kind == ST_HVDH_PERMUTE && s.hv.o_stride != 1) o << " out_stride=" << s.hv.o_stride;
    } else if (s.kind == ST_RS_DECONV || s.kind == ST_RS_INTERP) {
      o << " n=" << s.n << " m=" << s.m << " w=" << s.w << " beta=" << s.beta;
    } else {
      o << " n=" << s.n << " batch=" << s.batch;
    }
    o << "\n";
  }
  return o.str();
}

// ------------------------------------------------------------ the plan --

}  // namespace fftjh

struct fftjh_plan_s {
  fftjh::Layout L;
  fftjh::DevBufPtr scratch[4];
  int device = 0;
  // ragged plans with long items: a second stream on which the short-item
  // launches run beside the long items' tile passes, forked from and joined
  // back to the caller's stream through two events
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  mutable std::mutex side_mu;  // one fork/join sequence at a time (calls from several threads)
  fftjh_plan_s() = default;
  fftjh_plan_s(const fftjh_plan_s&) = delete;
  fftjh_plan_s& operator=(const fftjh_plan_s&) = delete;
  ~fftjh_plan_s() {
    if (!side) return;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return;
    if (cur != device) (void)hipSetDevice(device);
    (void)hipStreamSynchronize(side);
    (void)hipEventDestroy(fork);
    (void)hipEventDestroy(join);
    (void)hipStreamDestroy(side);
    if (cur != device) (void)hipSetDevice(cur);
  }
};

namespace fftjh {

static std::string key_of(const char* tag, std::initializer_list<double> vals) {
  std::ostringstream o;
  o.precision(17);
  o << tag;
  for (double v : vals) o << ":" << v;
  return o.str();
}

static void materialize(fftjh_plan_s& P) {
  Layout& L = P.L;
  hip_check(hipGetDevice(&P.device), "hipGetDevice");
  const int dt = L.dtype, es = esize(dt);
  for (Step& s : L.steps) {
    const double dev = P.device;
    // corr_j = scale / (w * Phi(j'/m)) of a resampled length (deconvolution
    // step, or the first pass that does it at its load)
    auto rs_corr = [&](const Step& st) {
      const int64_t n = st.n, m = st.m;
      const int w = st.w;
      const double beta = st.beta, scale = st.scale;
      return cached(key_of("rs_corr", {dev, double(dt), double(n), double(m), double(w), beta, scale}), n * (es / 2),
                    [&](void* p) {
                      std::vector<double> zq, wq;
                      gauss_legendre(2 * w + 20, zq, wq);
                      DevBuf nodes(zq.size() * 2 * sizeof(double));
                      hip_check(hipMemcpy(nodes.p, zq.data(), zq.size() * sizeof(double), hipMemcpyHostToDevice),
                                "node upload");
                      double* wqd = static_cast<double*>(nodes.p) + zq.size();
                      hip_check(hipMemcpy(wqd, wq.data(), wq.size() * sizeof(double), hipMemcpyHostToDevice),
                                "node upload");
                      hip_check(launch_rs_corr(dt, p, n, m, w, beta, static_cast<const double*>(nodes.p), wqd,
                                               static_cast<int>(zq.size()), scale, 0),
                                "corr kernel");
                      hip_check(hipDeviceSynchronize(), "corr kernel");
                    });
    };
    if (s.kind == ST_PASS) {
      // stage twiddle table, shared by every pass with the same chain
      std::ostringstream k;
      k << "stage:" << dev << ":" << dt << ":" << s.dir;
      for (int r : s.radices) k << ":" << r;
      int64_t total = 0, ns = 1;
      for (int R : s.radices) {
        total += (R - 1) * ns;
        ns *= R;
      }
      const std::vector<int> rad = s.radices;
      const int dir = s.dir;
      s.tw = cached(k.str(), std::max<int64_t>(total, 1) * es, [&](void* p) {
        int64_t off = 0, nsl = 1;
        for (int R : rad) {
          hip_check(launch_tw_stage(dt, static_cast<char*>(p) + off * es, static_cast<int>(nsl), R, dir, 0),
                    "twiddle kernel");
          off += (R - 1) * nsl;
          nsl *= R;
        }
      });
      // persistent grid: as many workgroups as are resident at once
      // persistent grid (as many workgroups as are resident at once) for the
      // power-of-two kernel; one workgroup per group for the generic one
      static const int persist_env = [] {
        const char* e = std::getenv("FFTJH_PERSIST");
        return e ? std::atoi(e) : -1;
      }();
      // (small generic tiles leave room for several workgroups per CU and
      // gain from it; a 1-workgroup-per-CU generic tile loses)
      const bool small_tile = static_cast<int64_t>(lds_elems(s.d.C * s.d.n)) * es <= 48 * 1024;
      const bool persist = persist_env >= 0 ? persist_env != 0 : (s.pow2 || small_tile);
      const int resident = !persist  ? 0
                           : s.mat   ? pass_mat_resident_groups(s.mat, s.dir, s.d)
                           : s.pow2  ? pass_pow2_resident_groups(dt, s.dir, s.d)
                                     : pass_resident_groups(dt, s.dir, s.d);
      if (s.d.tw_period) {
        // multi-pass twiddle tables (FP64): w_P^(i*2^sh) for i < ptw_lo, then w_P^i for i < 2^sh
        const int64_t P = s.d.tw_period, hi = s.d.ptw_lo, lo = int64_t(1) << s.d.ptw_sh;
        std::ostringstream pk;
        pk << "ptw:" << dev << ":" << dir << ":" << P;  // FP64 for every plan precision
        s.ptw = cached(pk.str(), (hi + lo) * 16, [&](void* p) {
          hip_check(launch_tw_pow(FFTJH_F64, p, hi, lo, P, dir, 0), "twiddle kernel");
          hip_check(launch_tw_pow(FFTJH_F64, static_cast<char*>(p) + hi * 16, lo, 1, P, dir, 0), "twiddle kernel");
        });
        s.d.ptw = s.ptw->p;
      }
      if (s.mat) {
        // split DFT-matrix operands of this mode and direction (FP64-generated)
        std::ostringstream mk;
        mk << "mat:" << dev << ":" << s.mat << ":" << s.dir;
        const int mode = s.mat;
        s.mat_tab = cached(mk.str(), mat_table_bytes(mode),
                           [&](void* p) { hip_check(launch_mat_table(mode, dir, p, 0), "operand table kernel"); });
      }
      if (s.d.rs_io) {
        s.aux = rs_corr(s);
        s.d.rs_corr = s.aux->p;
      }
      if (s.d.real_io) {
        // w_{2h}^e = exp(dir*2*pi*i*e/(2h)) (real.jl:26-39) from an FP64
        // two-level table (w^(i*2^sh) for i < rtw_hi, then w^i for i < 2^sh)
        const int64_t P = 2 * s.d.real_h, hi = s.d.rtw_hi, lo = int64_t(1) << s.d.rtw_sh;
        std::ostringstream rk;
        rk << "ptw:" << dev << ":" << dir << ":" << P;
        s.aux = cached(rk.str(), (hi + lo) * 16, [&](void* p) {
          hip_check(launch_tw_pow(FFTJH_F64, p, hi, lo, P, dir, 0), "twiddle kernel");
          hip_check(launch_tw_pow(FFTJH_F64, static_cast<char*>(p) + hi * 16, lo, 1, P, dir, 0), "twiddle kernel");
        });
        s.d.rtw = s.aux->p;
      }
      s.d.grid = resident > 0 ? std::min(resident, s.d.ngroups) : s.d.ngroups;
      s.desc = std::make_shared<DevBuf>(sizeof(PassDesc));
      hip_check(hipMemcpy(s.desc->p, &s.d, sizeof(PassDesc), hipMemcpyHostToDevice), "descriptor upload");
    } else if (s.kind == ST_RAGGED) {
     for (RaggedData* Rp : {s.rg.get(), s.rg2.get()}) {
      if (!Rp || Rp->items.empty()) continue;
      RaggedData& R = *Rp;
      if (R.groups <= 0) {
        const int res = ragged_resident_groups(dt, R.dir, R.big);
        R.groups = res > 0 ? res : 1024;
      }
      R.groups = static_cast<int>(std::min<int64_t>(R.groups, static_cast<int64_t>(R.items.size())));
      ragged_assign(R);
      auto upload = [&](const void* src, size_t bytes) {
        auto b = std::make_shared<DevBuf>(std::max<size_t>(bytes, 16));
        if (bytes) hip_check(hipMemcpy(b->p, src, bytes, hipMemcpyHostToDevice), "ragged table upload");
        return b;
      };
      R.d_items = upload(R.items.data(), R.items.size() * sizeof(RaggedItem));
      R.d_wgptr = upload(R.wg_ptr.data(), R.wg_ptr.size() * sizeof(int32_t));
      R.d_wgitems = upload(R.wg_items.data(), R.wg_items.size() * sizeof(int32_t));
      R.d_wgws = upload(R.wg_ws.data(), R.wg_ws.size() * sizeof(int64_t));
      R.d_scheds = upload(R.scheds.data(), R.scheds.size() * sizeof(RaggedSched));
      R.d_tw = std::make_shared<DevBuf>(std::max<int64_t>(R.tw_total, 1) * es);
      for (const auto& j : R.stage_jobs)
        hip_check(launch_tw_stage(dt, static_cast<char*>(R.d_tw->p) + j.off * es, j.ns, j.R, R.dir, 0), "twiddles");
      for (const auto& j : R.pow_jobs)
        hip_check(launch_tw_pow(dt, static_cast<char*>(R.d_tw->p) + j.off * es, j.count, j.mult, j.L, R.dir, 0),
                  "twiddles");
      R.d_corr = std::make_shared<DevBuf>(std::max<int64_t>(R.corr_total, 1) * (es / 2));
      if (!R.corr_jobs.empty()) {
        int wmax = 0;
        for (const auto& j : R.corr_jobs) wmax = std::max(wmax, j.w);
        std::map<int, std::shared_ptr<DevBuf>> nodes;
        for (const auto& j : R.corr_jobs) {
          auto& nb = nodes[j.w];
          const int nq = 2 * j.w + 20;
          if (!nb) {
            std::vector<double> zq, wq;
            gauss_legendre(nq, zq, wq);
            nb = std::make_shared<DevBuf>(2 * nq * sizeof(double));
            hip_check(hipMemcpy(nb->p, zq.data(), nq * sizeof(double), hipMemcpyHostToDevice), "nodes");
            hip_check(hipMemcpy(static_cast<double*>(nb->p) + nq, wq.data(), nq * sizeof(double),
                                hipMemcpyHostToDevice),
                      "nodes");
          }
          hip_check(launch_rs_corr(dt, static_cast<char*>(R.d_corr->p) + j.off * (es / 2), j.n, j.m, j.w, j.beta,
                                   static_cast<const double*>(nb->p), static_cast<const double*>(nb->p) + nq, nq,
                                   j.scale, 0),
                    "corr kernel");
        }
        hip_check(hipDeviceSynchronize(), "corr kernel");
      }
      hip_check(hipDeviceSynchronize(), "ragged tables");
      R.d_ws = std::make_shared<DevBuf>(std::max<int64_t>(R.ws_total, 1) * es);
     }
     if (s.rg3) {
      RaggedData& R = *s.rg3;
      auto upload = [&](const void* src, size_t bytes) {
        auto b = std::make_shared<DevBuf>(std::max<size_t>(bytes, 16));
        if (bytes) hip_check(hipMemcpy(b->p, src, bytes, hipMemcpyHostToDevice), "ragged table upload");
        return b;
      };
      static const bool xcd_env = [] {
        const char* e = std::getenv("FFTJH_RAGGED_XCD");
        return !e || std::atoi(e) != 0;
      }();
      if (xcd_env) {
        R.tiles_a = xcd_tile_order(R.tiles_a);
        R.tiles_b = xcd_tile_order(R.tiles_b);
      }
      if (R.tiles_a.size() > (size_t(1) << 31) - 1 || R.tiles_b.size() > (size_t(1) << 31) - 1)
        fail(FFTJH_ERR_UNSUPPORTED, "too many ragged pass tiles");
      R.d_items = upload(R.items.data(), R.items.size() * sizeof(RaggedItem));
      R.d_scheds = upload(R.scheds.data(), R.scheds.size() * sizeof(RaggedSched));
      R.d_tiles_a = upload(R.tiles_a.data(), R.tiles_a.size() * sizeof(RaggedTile));
      R.d_tiles_b = upload(R.tiles_b.data(), R.tiles_b.size() * sizeof(RaggedTile));
      R.d_tw = std::make_shared<DevBuf>(std::max<int64_t>(R.tw_total, 1) * es);
      for (const auto& j : R.stage_jobs)
        hip_check(launch_tw_stage(dt, static_cast<char*>(R.d_tw->p) + j.off * es, j.ns, j.R, R.dir, 0), "twiddles");
      for (const auto& j : R.pow_jobs)
        hip_check(launch_tw_pow(dt, static_cast<char*>(R.d_tw->p) + j.off * es, j.count, j.mult, j.L, R.dir, 0),
                  "twiddles");
      R.d_tw64 = std::make_shared<DevBuf>(std::max<int64_t>(R.tw64_total, 1) * 16);
      for (const auto& j : R.pow64_jobs)
        hip_check(launch_tw_pow(FFTJH_F64, static_cast<char*>(R.d_tw64->p) + j.off * 16, j.count, j.mult, j.L, R.dir,
                                0),
                  "twiddles");
      hip_check(hipDeviceSynchronize(), "ragged tables");
      R.d_ws = std::make_shared<DevBuf>(std::max<int64_t>(R.ws_total, 1) * es);
      static const bool fork_env = [] {
        const char* e = std::getenv("FFTJH_RAGGED_FORK");
        return !e || std::atoi(e) != 0;
      }();
      if (fork_env && !P.side) {
        hip_check(hipStreamCreateWithFlags(&P.side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        hip_check(hipEventCreateWithFlags(&P.fork, hipEventDisableTiming), "hipEventCreateWithFlags");
        hip_check(hipEventCreateWithFlags(&P.join, hipEventDisableTiming), "hipEventCreateWithFlags");
      }
     }
    } else if (s.kind == ST_R2C_TWIST || s.kind == ST_C2R_PRE) {
      // w_k = exp(dir*2*pi*i*k/n), k = 0..h (real.jl:26-39)
      const int64_t cnt = s.h + 1, n = s.n;
      const int dir = s.dir;
      s.aux = cached(key_of("real", {dev, double(dt), double(dir), double(n)}), cnt * es,
                     [&](void* p) { hip_check(launch_tw_pow(dt, p, cnt, 1, n, dir, 0), "twiddle kernel"); });
    } else if (s.kind == ST_RS_INTERP) {
      const int pieces = dt == FFTJH_F32 ? 64 : 128;
      const RsWeights wf = fit_rs_weights(dt, s.w, s.beta, pieces);
      s.pieces = pieces;
      s.degree = wf.degree;
      const size_t cnt = wf.coef.size();
      s.aux = cached(key_of("rs_wpoly", {dev, double(dt), double(s.w), s.beta, double(pieces)}), cnt * (es / 2),
                     [&](void* p) {
                       if (dt == FFTJH_F32) {
                         std::vector<float> c(wf.coef.begin(), wf.coef.end());
                         hip_check(hipMemcpy(p, c.data(), cnt * sizeof(float), hipMemcpyHostToDevice), "weights upload");
                       } else {
                         std::vector<double> c(wf.coef.begin(), wf.coef.end());
                         hip_check(hipMemcpy(p, c.data(), cnt * sizeof(double), hipMemcpyHostToDevice), "weights upload");
                       }
                     });
    } else if (s.kind == ST_RS_DECONV) {
      s.aux = rs_corr(s);
    } else if (s.kind == ST_LEAN) {
      // stage table w_n^i (i < n) and the multi-pass twiddle's FP64 tables
      const int N = s.ld.n, dir = s.dir;
      if (s.ld.rs) {
        s.aux = rs_corr(s);
        s.ld.rs_corr = s.aux->p;
      }
      if (s.ld.real >= 3) {
        // w_{2h}^e (real.jl:26-39) from an FP64 two-level table, the same
        // table as the power-of-two kernel's paired passes
        const int64_t P = 2 * s.ld.real_h, hi = s.ld.rtw_hi, lo = int64_t(1) << s.ld.rtw_sh;
        std::ostringstream rk;
        rk << "ptw:" << dev << ":" << dir << ":" << P;
        s.aux = cached(rk.str(), (hi + lo) * 16, [&](void* p) {
          hip_check(launch_tw_pow(FFTJH_F64, p, hi, lo, P, dir, 0), "twiddle kernel");
          hip_check(launch_tw_pow(FFTJH_F64, static_cast<char*>(p) + hi * 16, lo, 1, P, dir, 0), "twiddle kernel");
        });
        s.ld.rtw = s.aux->p;
      } else if (s.ld.real) {
        // w_k = exp(dir*2*pi*i*k/(2N)), k = 0..N (real.jl:26-39)
        const int64_t n2 = 2 * int64_t(N), cnt = N + 1;
        s.aux = cached(key_of("real", {dev, double(dt), double(dir), double(n2)}), cnt * es,
                       [&](void* p) { hip_check(launch_tw_pow(dt, p, cnt, 1, n2, dir, 0), "twiddle kernel"); });
        s.ld.rtw = s.aux->p;
      }
      s.tw = cached(key_of("lean", {dev, double(dir), double(N)}), int64_t(N) * 8,
                    [&](void* p) { hip_check(launch_tw_pow(FFTJH_F32, p, N, 1, N, dir, 0), "twiddle kernel"); });
      if (s.ld.psh) {
        const int64_t P = s.ld.pmod ? s.ld.pmod : int64_t(1) << s.ld.psh, hi = s.ld.ptw_lo,
                      lo = int64_t(1) << s.ld.ptw_sh;
        std::ostringstream pk;
        pk << "ptw:" << dev << ":" << dir << ":" << P;
        s.ptw = cached(pk.str(), (hi + lo) * 16, [&](void* p) {
          hip_check(launch_tw_pow(FFTJH_F64, p, hi, lo, P, dir, 0), "twiddle kernel");
          hip_check(launch_tw_pow(FFTJH_F64, static_cast<char*>(p) + hi * 16, lo, 1, P, dir, 0), "twiddle kernel");
        });
        s.ld.ptw = s.ptw->p;
      }
      if (s.ld.mat) {
        // split DFT-matrix operands of this mode and direction (FP64-generated)
        std::ostringstream mk;
        mk << "mat:" << dev << ":" << s.ld.mat << ":" << dir;
        const int mode = s.ld.mat;
        s.mat_tab = cached(mk.str(), mat_table_bytes(mode),
                           [&](void* p) { hip_check(launch_mat_table(mode, dir, p, 0), "operand table kernel"); });
        s.ld.matp = s.mat_tab->p;
      }
    }
  }
  for (int i = 0; i < 4; ++i)
    if (L.scratch_elems[i]) P.scratch[i] = std::make_shared<DevBuf>(L.scratch_elems[i] * es);
}

// ------------------------------------------- fixed per-device scratch --
//
// fftjh_execute_static runs a plan with its scratch (the intermediates of
// multi-pass plans, the ragged per-workgroup workspace) carved out of this
// array instead of the buffers the plan allocated, so the call itself
// allocates nothing.  Every GPU the library runs on holds its own copy from
// module load to process end.  In host code `scratch` names a host-side
// stand-in, never the device memory: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
static __device__ __align__(256) unsigned char scratch[FFTJH_STATIC_SCRATCH_BYTES];

static std::mutex g_static_mu;
static std::vector<char*> g_static_ptr;  // device address of `scratch`, per device

// the current device's copy of `scratch` (looked up on its first use there)
static char* static_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_static_mu);
  if (static_cast<int>(g_static_ptr.size()) <= device) g_static_ptr.resize(device + 1, nullptr);
  if (!g_static_ptr[device]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) fail(FFTJH_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_static_ptr[device] = static_cast<char*>(p);
  }
  return g_static_ptr[device];
}

static int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

// Byte offsets of the scratch roles and of the ragged workspace inside the
// fixed array (each 256-byte aligned); returns the bytes the plan needs.
static int64_t static_layout(const fftjh_plan_s& P, int64_t role_off[4], int64_t* ws_off) {
  const int64_t es = esize(P.L.dtype);
  int64_t at = 0;
  for (int i = 0; i < 4; ++i) {
    role_off[i] = at;
    at += align256(P.L.scratch_elems[i] * es);
  }
  *ws_off = at;
  for (const Step& s : P.L.steps)
    if (s.kind == ST_RAGGED)
      at += align256(std::max({s.rg->ws_total, s.rg2 ? s.rg2->ws_total : int64_t(0),
                                s.rg3 ? s.rg3->ws_total : int64_t(0)}) *
                     es);
  return at;
}

static hipError_t launch_step(const fftjh_plan_s& P, const Step& s, void* const* ptr, hipStream_t st,
                              void* ws = nullptr) {
  const int dt = P.L.dtype;
  switch (s.kind) {
    case ST_PASS:
      if (s.mat)
        return launch_pass_mat(s.mat, s.dir, s.d, static_cast<const PassDesc*>(s.desc->p), ptr[s.src], ptr[s.dst],
                               s.tw->p, s.mat_tab->p, st);
      if (s.pow2)
        return launch_pass_pow2(dt, s.dir, s.d, static_cast<const PassDesc*>(s.desc->p), ptr[s.src], ptr[s.dst],
                                s.tw->p, st);
      return launch_pass(dt, s.dir, s.d, static_cast<const PassDesc*>(s.desc->p), ptr[s.src], ptr[s.dst], s.tw->p,
                         st);
    case ST_R2C_TWIST:
      return launch_r2c_twist(dt, ptr[s.src], ptr[s.dst], s.aux->p, s.h, s.batch, s.s1, s.s2, st);
    case ST_C2R_PRE:
      return launch_c2r_pre(dt, ptr[s.src], ptr[s.dst], s.aux->p, s.h, s.batch, s.s1, s.s2, s.pre, st);
    case ST_REAL_TO_COMPLEX:
      return launch_real_to_complex(dt, ptr[s.src], ptr[s.dst], s.n, s.batch, st);
    case ST_COPY_PREFIX:
      return launch_copy_prefix(dt, ptr[s.src], ptr[s.dst], s.n, s.keep, s.batch, st);
    case ST_HERMITIAN_FILL:
      return launch_hermitian_fill(dt, ptr[s.src], ptr[s.dst], s.n, s.batch, st);
    case ST_TAKE_REAL:
      return launch_take_real(dt, ptr[s.src], ptr[s.dst], s.n * s.batch, st);
    case ST_RS_DECONV:
      return launch_rs_deconv(dt, ptr[s.src], ptr[s.dst], s.aux->p, s.n, s.m, s.batch, s.Bi, s.ga, st);
    case ST_RS_INTERP:
      return launch_rs_interp(dt, ptr[s.src], ptr[s.dst], s.aux->p, s.pieces, s.degree, s.n, s.m, s.w, s.batch, s.Bi,
                              s.gb, st);
    case ST_LEAN:
      return launch_lean(s.dir, s.ld, ptr[s.src], ptr[s.dst], s.tw->p, st);
    case ST_HVDH_PERMUTE:
      return launch_hvdh_permute(s.hv, ptr[s.src], ptr[s.dst], static_cast<float>(s.scale), st);
    case ST_HVDH_SPREAD:
      return launch_hvdh_spread(s.hv, ptr[s.src], ptr[s.dst], st);
    case ST_HVDH_SOLVE:
      return launch_hvdh_solve(s.hv, ptr[s.src], ptr[s.dst], st);
    case ST_STRIDED_COPY:
      return launch_strided_copy(dt, ptr[s.src], ptr[s.dst], s.n, s.batch, s.Bi, s.ga, s.gb, st);
    case ST_RAGGED: {
      // the small-item launch, then the large-item one (same stream: they
      // may share a caller's workspace).  With the plan's own workspaces
      // (disjoint) and long items, the two run on the plan's side stream
      // beside the long items' passes.
      const bool small = (s.rg && !s.rg->items.empty()) || (s.rg2 && !s.rg2->items.empty());
      const bool fork = !ws && P.side && s.rg3 && small;
      hipStream_t sst = st;
      std::unique_lock<std::mutex> lock(P.side_mu, std::defer_lock);
      if (fork) {
        lock.lock();
        hipError_t e = hipEventRecord(P.fork, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(P.side, P.fork, 0);
        if (e != hipSuccess) return e;
        sst = P.side;
      }
      for (const RaggedData* Rp : {s.rg.get(), s.rg2.get()}) {
        if (!Rp || Rp->items.empty()) continue;
        const RaggedData& R = *Rp;
        RaggedLaunch rl;
        rl.big = R.big;
        rl.groups = R.groups;
        rl.items = R.d_items->p;
        rl.wg_ptr = R.d_wgptr->p;
        rl.wg_items = R.d_wgitems->p;
        rl.scheds = R.d_scheds->p;
        rl.tw = R.d_tw->p;
        rl.corr = R.d_corr->p;
        rl.x = ptr[s.src];
        rl.y = ptr[s.dst];
        rl.ws = ws ? ws : R.d_ws->p;
        rl.wg_ws = R.d_wgws->p;
        const hipError_t e = launch_ragged(dt, s.dir, rl, sst);
        if (e != hipSuccess) return e;
      }
      if (s.rg3) {
        // then the long items' column tiles, then their row tiles
        const RaggedData& R = *s.rg3;
        for (int ph = 0; ph < 2; ++ph) {
          RaggedPassLaunch pl;
          pl.phase = ph;
          pl.tiles = static_cast<int>(ph ? R.tiles_b.size() : R.tiles_a.size());
          pl.items = R.d_items->p;
          pl.tile_list = ph ? R.d_tiles_b->p : R.d_tiles_a->p;
          pl.scheds = R.d_scheds->p;
          pl.tw = R.d_tw->p;
          pl.tw64 = R.d_tw64->p;
          pl.x = ptr[s.src];
          pl.y = ptr[s.dst];
          pl.ws = ws ? ws : R.d_ws->p;
          const hipError_t e = launch_ragged_pass(dt, s.dir, pl, st);
          if (e != hipSuccess) return e;
        }
      }
      if (fork) {
        const hipError_t e = hipEventRecord(P.join, P.side);
        if (e != hipSuccess) return e;
        return hipStreamWaitEvent(st, P.join, 0);
      }
      return hipSuccess;
    }
  }
  return hipErrorInvalidValue;
}

// static_base: the device's fixed scratch array (fftjh_execute_static), or
// NULL for the plan's own scratch buffers
static void run(const fftjh_plan_s& P, void* y, const void* x, hipStream_t st, hipEvent_t* ev,
                char* static_base = nullptr) {
  void* ptr[kRoles] = {const_cast<void*>(x), y, nullptr, nullptr, nullptr, nullptr};
  void* ws = nullptr;
  if (static_base) {
    int64_t role_off[4], ws_off = 0;
    static_layout(P, role_off, &ws_off);
    for (int i = 0; i < 4; ++i) ptr[RS0 + i] = P.L.scratch_elems[i] ? static_base + role_off[i] : nullptr;
    ws = static_base + ws_off;
  } else {
    for (int i = 0; i < 4; ++i) ptr[RS0 + i] = P.scratch[i] ? P.scratch[i]->p : nullptr;
  }
  // an in-place call (x == y) takes the steps' alias roles where they have them
  const bool alias = x == static_cast<const void*>(y);
  for (size_t i = 0; i < P.L.steps.size(); ++i) {
    const Step& s = P.L.steps[i];
    if (ev) hip_check(hipEventRecord(ev[2 * i], st), "hipEventRecord");
    if (s.kind == ST_LEAN) {
      const int sr = alias && s.alias_src >= 0 ? s.alias_src : s.src;
      const int dr = alias && s.alias_dst >= 0 ? s.alias_dst : s.dst;
      hip_check(launch_lean(s.dir, s.ld, ptr[sr], ptr[dr], s.tw->p, st), kStepNames[s.kind]);
    } else {
      hip_check(launch_step(P, s, ptr, st, ws), kStepNames[s.kind]);
    }
    if (ev) hip_check(hipEventRecord(ev[2 * i + 1], st), "hipEventRecord");
  }
}

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return FFTJH_OK;
  } catch (const Fail& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return FFTJH_ERR_HIP;
  }
}

// Streaming hints (the power-of-two kernel honours them): a pass that reads
// the plan's input x loads it non-temporally, and the pass that writes the
// plan's output y last stores it non-temporally -- neither is touched again
// by the plan, and streamed they leave the 256 MB Infinity Cache to the
// intermediates.  Only for footprints (x, y and scratch) of 384 MiB and up:
// 2^24 forward + inverse 0.395 -> 0.347 ms per step, 4096 x 16384 0.249 ->
// 0.239 ms; a plan that fits the cache loses with them (its input, re-read
// by the next call, no longer stays there: 1,000,003 x 4 0.114 -> 0.133 ms,
// 2^22 x 4 C2R 0.095 -> 0.107 ms, 256^3 C2C 0.166 -> 0.172 ms).
// FFTJH_NT: -1 (default) that rule, else bit 0 loads and bit 1 stores.
static void stream_hints(Layout& L) {
  static const int nt_env = [] {
    const char* e = std::getenv("FFTJH_NT");
    return e ? std::atoi(e) : -1;
  }();
  int nt = nt_env;
  if (nt < 0) {
    const int64_t es = esize(L.dtype);
    const int64_t in_es = L.kind == FFTJH_KIND_R2C ? es / 2 : es;
    const int64_t out_es = L.kind == FFTJH_KIND_C2R ? es / 2 : es;
    int64_t bytes = L.in_elems * in_es + L.out_elems * out_es;
    for (int i = 0; i < 4; ++i) bytes += L.scratch_elems[i] * es;
    nt = bytes >= (int64_t(384) << 20) ? 3 : 0;
  }
  int last_y = -1;
  for (size_t i = 0; i < L.steps.size(); ++i)
    if (L.steps[i].dst == RY) last_y = static_cast<int>(i);
  for (size_t i = 0; i < L.steps.size(); ++i) {
    Step& s = L.steps[i];
    if (s.kind == ST_LEAN && s.ld.rin && (nt & 1) && s.src == RX) s.ld.nt = 1;
    // (the power-of-two kernel's register-I/O instantiations honour them)
    if (s.kind != ST_PASS || !s.pow2 || (s.d.n & (s.d.n - 1)) || s.d.real_io || s.d.rs_io || s.d.rows_in ||
        s.d.rows_out)
      continue;
    if ((nt & 1) && s.src == RX) s.d.nt_load = 1;
    if ((nt & 2) && static_cast<int>(i) == last_y && s.src != RY) s.d.nt_store = 1;
  }
}

// FP32 power-of-two row passes of 64, 128 or 256 points over contiguous
// rows in and out (groups of 16 rows, no multi-pass twiddle, no fused real
// or resampling edge, the planner's own schedule) run on the lean kernel
// with both edges transposed through LDS: thread t moves element t of each
// of the group's 16 rows, so every wave instruction reads or writes whole
// 512-byte runs.  FFTJH_LEAN_ROWS=0 keeps the power-of-two kernel.
static void lean_rows(Layout& L) {
  static const bool env = [] {
    const char* e = std::getenv("FFTJH_LEAN_ROWS");
    return !e || std::atoi(e) != 0;
  }();
  fftjh_tuning t = L.tune;
  t.flags &= ~(FFTJH_FLAG_MFMA_F16 | FFTJH_FLAG_MFMA_BF16 | FFTJH_FLAG_MFMA_BF16X3);
  if (!env || L.dtype != FFTJH_F32 || L.mat || !tune_is_default(t)) return;
  static const bool cols_env = [] {
    const char* e = std::getenv("FFTJH_LEAN_COLS");
    return !e || std::atoi(e) != 0;
  }();
  for (Step& s : L.steps) {
    if (s.kind != ST_PASS || !s.pow2 || s.mat) continue;
    const PassDesc& d = s.d;
    // rows of 2N reals with the real edge fused (R2C twist at the store, C2R
    // pre-pass at the load): rows in and out through LDS, the edge there
    static const bool real_env = [] {
      const char* e = std::getenv("FFTJH_LEAN_REAL");
      return !e || std::atoi(e) != 0;
    }();
    if (real_env && lean_length(d.n) && (d.real_io == 1 || d.real_io == 2) && !d.rs_io && !d.tw_period &&
        d.Bi == 1 && d.B2 == 1 && d.in.stride == 1 && d.out.stride == 1 && d.batch % 16 == 0 &&
        d.batch / 16 < (int64_t(1) << 31) && !d.nt_load && !d.nt_store &&
        d.in.s_out >= d.n + (d.real_io == 2 ? 1 : 0) && d.out.s_out >= d.n + (d.real_io == 1 ? 1 : 0)) {
      Step st;
      st.kind = ST_LEAN;
      st.dir = s.dir;
      st.src = s.src;
      st.dst = s.dst;
      st.alias_src = s.alias_src;
      st.alias_dst = s.alias_dst;
      st.n = d.n;
      st.batch = d.batch;
      LeanDesc& ld = st.ld;
      ld.n = d.n;
      ld.rows = 1;
      ld.rin = 1;
      ld.real = d.real_io;
      ld.groups = static_cast<int32_t>(d.batch / 16);
      ld.D0 = ld.groups;
      ld.D1 = 1;
      ld.in_s0 = 16 * d.in.s_out;
      ld.istride = d.in.s_out;
      ld.out_s0 = 16 * d.out.s_out;
      ld.ostride = d.out.s_out;
      ld.scale = d.scale;
      s = st;
      continue;
    }
    // the paired real edges of a multi-pass half-length transform (R2C last
    // pass, rows in; C2R first pass, columns in), 64 or 128 points: two
    // 16-column sub-groups (a column block and its mirror) per workgroup
    static const bool pair_env = [] {
      const char* e = std::getenv("FFTJH_LEAN_PAIRED");
      return !e || std::atoi(e) != 0;
    }();
    if (pair_env && (d.n == 64 || d.n == 128) && (d.real_io == 3 || d.real_io == 4) && !d.rs_io &&
        d.pair_M % 32 == 0 && d.Bi * d.B2 == d.pair_M && d.batch % d.pair_M == 0 && d.batch / 32 < (int64_t(1) << 31) &&
        d.Bi < (int64_t(1) << 31) && d.B2 < (int64_t(1) << 31) && !d.nt_load && !d.nt_store &&
        (d.real_io == 4 || (d.in.stride == 1 && !d.tw_period)) &&
        (!d.tw_period || ((d.tw_period & (d.tw_period - 1)) == 0 && d.tw_period <= (int64_t(1) << 32) &&
                          d.tw_qdiv >= 1 && d.tw_qdiv < (int64_t(1) << 31)))) {
      Step st;
      st.kind = ST_LEAN;
      st.dir = s.dir;
      st.src = s.src;
      st.dst = s.dst;
      st.alias_src = s.alias_src;
      st.alias_dst = s.alias_dst;
      st.n = d.n;
      st.batch = d.batch;
      LeanDesc& ld = st.ld;
      ld.n = d.n;
      ld.real = d.real_io;
      ld.groups = static_cast<int32_t>(d.batch / 32);
      ld.D0 = ld.groups;
      ld.D1 = 1;
      ld.pm = d.pair_M;
      ld.pbi = d.Bi;
      ld.pb2 = d.B2;
      ld.in_s0 = d.in.s_in;
      ld.in_s1 = d.in.s_mid;
      ld.in_s2 = d.in.s_out;
      ld.istride = d.in.stride;
      ld.out_s0 = d.out.s_in;
      ld.out_s1 = d.out.s_mid;
      ld.out_s2 = d.out.s_out;
      ld.ostride = d.out.stride;
      ld.real_h = d.real_h;
      ld.rtw_sh = d.rtw_sh;
      ld.rtw_hi = d.rtw_hi;
      if (d.tw_period) {
        lean_twiddle(ld, d.tw_period, d.tw_acoef % d.tw_period, d.tw_bcoef % d.tw_period);
        ld.pqdiv = d.tw_qdiv;
      }
      ld.scale = d.scale;
      s = st;
      continue;
    }
    if (!lean_length(d.n) || d.real_io || d.rs_io) continue;
    // a generic layout's column passes with a power-of-two multi-pass
    // twiddle (columns b % Bi contiguous, element stride and the middle and
    // outer indices as the geometry says, the same on both sides):
    //   qdiv a multiple of 16, B2 = 1: the twiddle's multiplier q = column /
    //   qdiv is a group digit (d1), e = d1*bcoef*k;
    //   qdiv = 1: the multiplier is the column, e = col*(acoef*bm + bcoef*k),
    //   bm the group digit d1 (TWQ form)
    if (cols_env && d.tw_period && (d.tw_period & (d.tw_period - 1)) == 0 && d.tw_period <= (int64_t(1) << 32) &&
        d.Bi % 16 == 0 && d.in.s_in == 1 && d.in.stride == d.out.stride && d.in.s_in == d.out.s_in &&
        d.in.s_mid == d.out.s_mid && d.in.s_out == d.out.s_out && !d.nt_load && !d.nt_store) {
      const bool digit = d.B2 == 1 && d.tw_qdiv % 16 == 0 && d.Bi % d.tw_qdiv == 0;
      const bool colq = d.tw_qdiv == 1;
      const int64_t outer = d.batch / (d.Bi * d.B2);
      if ((digit || colq) && outer * d.Bi * d.B2 == d.batch && d.batch / 16 < (int64_t(1) << 31)) {
        Step st;
        st.kind = ST_LEAN;
        st.dir = s.dir;
        st.src = s.src;
        st.dst = s.dst;
        st.alias_src = s.alias_src;
        st.alias_dst = s.alias_dst;
        st.n = d.n;
        st.batch = d.batch;
        LeanDesc& ld = st.ld;
        ld.n = d.n;
        ld.groups = static_cast<int32_t>(d.batch / 16);
        ld.in_s0 = ld.out_s0 = 16;
        ld.istride = ld.ostride = d.in.stride;
        ld.in_s2 = ld.out_s2 = d.in.s_out;
        if (digit) {
          ld.D0 = static_cast<int32_t>(d.tw_qdiv / 16);
          ld.D1 = static_cast<int32_t>(d.Bi / d.tw_qdiv);
          ld.in_s1 = ld.out_s1 = d.tw_qdiv;
          lean_twiddle(ld, d.tw_period, 0, d.tw_bcoef % d.tw_period);
        } else {
          ld.D0 = static_cast<int32_t>(d.Bi / 16);
          ld.D1 = static_cast<int32_t>(d.B2);
          ld.in_s1 = ld.out_s1 = d.in.s_mid;
          lean_twiddle(ld, d.tw_period, d.tw_acoef % d.tw_period, d.tw_bcoef % d.tw_period);
          ld.twq = 1;
        }
        ld.scale = d.scale;
        s = st;
        continue;
      }
    }
    // a generic layout's last pass: rows in (16 consecutive transforms at
    // row stride s_in), columns out (consecutive transforms adjacent), no
    // twiddle: the rows-in transposition with the column store
    if (cols_env && !d.tw_period && d.Bi % 16 == 0 && d.in.stride == 1 && d.in.s_in >= d.n && d.out.s_in == 1 &&
        d.out.stride >= d.Bi && !d.nt_load && !d.nt_store && d.batch % (d.Bi * d.B2) == 0 &&
        d.batch / 16 < (int64_t(1) << 31)) {
      Step st;
      st.kind = ST_LEAN;
      st.dir = s.dir;
      st.src = s.src;
      st.dst = s.dst;
      st.alias_src = s.alias_src;
      st.alias_dst = s.alias_dst;
      st.n = d.n;
      st.batch = d.batch;
      LeanDesc& ld = st.ld;
      ld.n = d.n;
      ld.rin = 1;
      ld.groups = static_cast<int32_t>(d.batch / 16);
      ld.D0 = static_cast<int32_t>(d.Bi / 16);
      ld.D1 = static_cast<int32_t>(d.B2);
      ld.in_s0 = 16 * d.in.s_in;
      ld.in_s1 = d.in.s_mid;
      ld.in_s2 = d.in.s_out;
      ld.istride = d.in.s_in;
      ld.out_s0 = 16;
      ld.out_s1 = d.out.s_mid;
      ld.out_s2 = d.out.s_out;
      ld.ostride = d.out.stride;
      ld.scale = d.scale;
      s = st;
      continue;
    }
    if (d.tw_period || d.B2 != 1) continue;
    // column blocks of an N-d axis (column b % Bi of block b / Bi, element j
    // at j*Bi), 16 adjacent columns per group: the lean column kernel
    auto colblock = [&](const Geometry& g) {
      return g.stride == d.Bi && g.s_in == 1 && g.s_mid == 0 && g.s_out == d.Bi * d.n;
    };
    // (a block of columns that is not a multiple of 16 ends in a partial group)
    const int64_t D0 = (d.Bi + 15) / 16;
    if (cols_env && d.Bi >= 16 && d.Bi < (int64_t(1) << 31) && colblock(d.in) && colblock(d.out) &&
        D0 * (d.batch / d.Bi) < (int64_t(1) << 31)) {
      Step st;
      st.kind = ST_LEAN;
      st.dir = s.dir;
      st.src = s.src;
      st.dst = s.dst;
      st.alias_src = s.alias_src;
      st.alias_dst = s.alias_dst;
      st.n = d.n;
      st.batch = d.batch;
      LeanDesc& ld = st.ld;
      ld.n = d.n;
      ld.D0 = static_cast<int32_t>(D0);
      ld.D1 = static_cast<int32_t>(d.batch / d.Bi);
      ld.groups = static_cast<int32_t>(D0 * (d.batch / d.Bi));
      ld.ncol = d.Bi % 16 ? static_cast<int32_t>(d.Bi) : 0;
      ld.in_s0 = ld.out_s0 = 16;
      ld.in_s1 = ld.out_s1 = d.Bi * d.n;
      ld.istride = ld.ostride = d.Bi;
      ld.scale = d.scale;
      s = st;
      continue;
    }
    if (d.Bi != 1 || d.batch % 16 || d.batch / 16 >= (int64_t(1) << 31)) continue;
    if (d.in.stride != 1 || d.out.stride != 1 || d.in.s_out < d.n || d.out.s_out < d.n) continue;
    Step st;
    st.kind = ST_LEAN;
    st.dir = s.dir;
    st.src = s.src;
    st.dst = s.dst;
    st.alias_src = s.alias_src;
    st.alias_dst = s.alias_dst;
    st.n = d.n;
    st.batch = d.batch;
    LeanDesc& ld = st.ld;
    ld.n = d.n;
    ld.rows = 1;
    ld.rin = 1;
    ld.groups = static_cast<int32_t>(d.batch / 16);
    ld.D0 = ld.groups;
    ld.D1 = 1;
    ld.in_s0 = 16 * d.in.s_out;
    ld.istride = d.in.s_out;
    ld.out_s0 = 16 * d.out.s_out;
    ld.ostride = d.out.s_out;
    ld.scale = d.scale;
    s = st;
  }
}

// Every lean launch takes the XCD-local group order (neighbouring column
// groups on one XCD).  Measured on one MI355X, forward, ms (off / only the
// launches with partial column groups / every launch): 256^3 R2C 0.1076 /
// 0.0844 / 0.0855, C2R 0.1069 / 0.0927 / 0.0879 (its 129-column half
// spectrum: every 128-byte run straddles two lines), 256^3 C2C 0.1257 /
// 0.1259 / 0.1250, 2^24 0.1353 / 0.1366 / 0.1341; 2^24 forward + inverse
// 0.2676 / 0.2670 / 0.2660.  FFTJH_LEAN_XCD: 1 (default) every launch, 2
// only partial groups, 0 none.
static void lean_xcd(Layout& L) {
  static const int env = [] {
    const char* e = std::getenv("FFTJH_LEAN_XCD");
    return e ? std::atoi(e) : 1;
  }();
  for (Step& s : L.steps)
    if (s.kind == ST_LEAN) s.ld.xcd = env == 1 || (env == 2 && s.ld.ncol) ? 1 : 0;
}

static int make_plan(fftjh_plan* out, Layout&& L) {
  if (!out) fail(FFTJH_ERR_ARGUMENT, "plan output pointer is NULL");
  auto P = std::make_unique<fftjh_plan_s>();
  P->L = std::move(L);
  lean_rows(P->L);
  lean_xcd(P->L);
  stream_hints(P->L);
  materialize(*P);
  *out = P.release();
  return 0;
}

}  // namespace fftjh

using namespace fftjh;

extern "C" {

int fftjh_plan_c2c(fftjh_plan* plan, int dtype, int rank, const int64_t* dims, int64_t batch, int sign,
                   double scale, unsigned flags) {
  return guarded([&] {
    if (plan) *plan = nullptr;
    make_plan(plan, layout_c2c(dtype, rank, dims, batch, sign, scale, nullptr, flags));
  });
}

int fftjh_plan_c2c_tuned(fftjh_plan* plan, int dtype, int rank, const int64_t* dims, int64_t batch, int sign,
                         double scale, const fftjh_tuning* tune) {
  return guarded([&] {
    if (plan) *plan = nullptr;
    make_plan(plan, layout_c2c(dtype, rank, dims, batch, sign, scale, tune));
  });
}

int fftjh_plan_ragged(fftjh_plan* plan, int dtype, int64_t count, const int64_t* sizes, const int64_t* offsets,
                      int sign, int normalized, unsigned flags) {
  return guarded([&] {
    if (plan) *plan = nullptr;
    if (flags) fail(FFTJH_ERR_ARGUMENT, "flags must be 0");
    make_plan(plan, layout_ragged(dtype, count, sizes, offsets, sign, normalized, nullptr));
  });
}

int fftjh_plan_ragged_tuned(fftjh_plan* plan, int dtype, int64_t count, const int64_t* sizes,
                            const int64_t* offsets, int sign, int normalized, const fftjh_ragged_tuning* tune) {
  return guarded([&] {
    if (plan) *plan = nullptr;
    make_plan(plan, layout_ragged(dtype, count, sizes, offsets, sign, normalized, tune));
  });
}

int fftjh_plan_r2c(fftjh_plan* plan, int dtype, int64_t n, int64_t batch, unsigned flags) {
  return guarded([&] {
    if (plan) *plan = nullptr;
    if (flags) fail(FFTJH_ERR_ARGUMENT, "flags must be 0");
    make_plan(plan, layout_r2c(dtype, 1, &n, batch));
  });
}

int fftjh_plan_c2r(fftjh_plan* plan, int dtype, int64_t n, int64_t batch, int normalized, unsigned flags) {
  return guarded([&] {
    if (plan) *plan = nullptr;
    if (flags) fail(FFTJH_ERR_ARGUMENT, "flags must be 0");
    make_plan(plan, layout_c2r(dtype, 1, &n, batch, normalized));
  });
}

int fftjh_plan_r2c_nd(fftjh_plan* plan, int dtype, int rank, const int64_t* dims, int64_t batch, unsigned flags) {
  return guarded([&] {
    if (plan) *plan = nullptr;
    if (flags) fail(FFTJH_ERR_ARGUMENT, "flags must be 0");
    make_plan(plan, layout_r2c(dtype, rank, dims, batch));
  });
}

int fftjh_plan_c2r_nd(fftjh_plan* plan, int dtype, int rank, const int64_t* dims, int64_t batch, int normalized,
                      unsigned flags) {
  return guarded([&] {
    if (plan) *plan = nullptr;
    if (flags) fail(FFTJH_ERR_ARGUMENT, "flags must be 0");
    make_plan(plan, layout_c2r(dtype, rank, dims, batch, normalized));
  });
}

int fftjh_execute(fftjh_plan plan, void* y, const void* x, fftjh_stream stream) {
  return guarded([&] {
    if (!plan) fail(FFTJH_ERR_ARGUMENT, "plan is NULL");
    if (!x || !y) fail(FFTJH_ERR_ARGUMENT, "x and y must be device pointers");
    run(*plan, y, x, static_cast<hipStream_t>(stream), nullptr);
  });
}

int fftjh_execute_static(fftjh_plan plan, void* y, const void* x, fftjh_stream stream) {
  return guarded([&] {
    if (!plan) fail(FFTJH_ERR_ARGUMENT, "plan is NULL");
    if (!x || !y) fail(FFTJH_ERR_ARGUMENT, "x and y must be device pointers");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != plan->device)
      fail(FFTJH_ERR_ARGUMENT, "plan was built on device " + std::to_string(plan->device) +
                                   ", the current device is " + std::to_string(dev));
    int64_t role_off[4], ws_off = 0;
    const int64_t need = static_layout(*plan, role_off, &ws_off);
    if (need > FFTJH_STATIC_SCRATCH_BYTES)
      fail(FFTJH_ERR_UNSUPPORTED, "plan needs " + std::to_string(need) +
                                      " bytes of scratch; the fixed scratch of fftjh_execute_static holds " +
                                      std::to_string(static_cast<int64_t>(FFTJH_STATIC_SCRATCH_BYTES)) +
                                      " (use fftjh_execute)");
    char* base = need > 0 ? static_scratch(dev) : nullptr;
    run(*plan, y, x, static_cast<hipStream_t>(stream), nullptr, base);
  });
}

int fftjh_execute_checked(fftjh_plan plan, void* y, int64_t y_elems, const void* x, int64_t x_elems,
                          fftjh_stream stream) {
  return guarded([&] {
    if (!plan) fail(FFTJH_ERR_ARGUMENT, "plan is NULL");
    const Layout& L = plan->L;
    if (x_elems != L.in_elems)
      fail(FFTJH_ERR_DIMENSION, "plan input length " + std::to_string(L.in_elems) + ", input length " +
                                    std::to_string(x_elems));
    if (y_elems != L.out_elems)
      fail(FFTJH_ERR_DIMENSION, "plan output length " + std::to_string(L.out_elems) + ", output length " +
                                    std::to_string(y_elems));
    if (!x || !y) fail(FFTJH_ERR_ARGUMENT, "x and y must be device pointers");
    run(*plan, y, x, static_cast<hipStream_t>(stream), nullptr);
  });
}

int fftjh_plan_info(fftjh_plan plan, int* kind, int* dtype, int* sign, int* rank, int64_t* dims, int64_t* batch,
                    double* scale, int64_t* in_elems, int64_t* out_elems) {
  return guarded([&] {
    if (!plan) fail(FFTJH_ERR_ARGUMENT, "plan is NULL");
    const Layout& L = plan->L;
    if (kind) *kind = L.kind;
    if (dtype) *dtype = L.dtype;
    if (sign) *sign = L.sign;
    if (rank) *rank = L.rank;
    if (dims)
      for (int i = 0; i < 8; ++i) dims[i] = i < L.rank ? L.dims[i] : 0;
    if (batch) *batch = L.batch;
    if (scale) *scale = L.scale;
    if (in_elems) *in_elems = L.in_elems;
    if (out_elems) *out_elems = L.out_elems;
  });
}

static void copy_out(const std::string& s, char* buf, int64_t len) {
  if (!buf || len <= 0) return;
  const size_t n = std::min<size_t>(s.size(), static_cast<size_t>(len - 1));
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
}

int fftjh_plan_describe(fftjh_plan plan, char* buf, int64_t len) {
  return guarded([&] {
    if (!plan) fail(FFTJH_ERR_ARGUMENT, "plan is NULL");
    copy_out(describe(plan->L), buf, len);
  });
}

int fftjh_describe_request(int kind, int dtype, int rank, const int64_t* dims, int64_t batch, int sign, char* buf,
                           int64_t len) {
  return guarded([&] {
    Layout L;
    if (kind == FFTJH_KIND_C2C) L = layout_c2c(dtype, rank, dims, batch, sign, 1.0);
    else if (kind == FFTJH_KIND_R2C)
      L = layout_r2c(dtype, rank, dims, batch);
    else if (kind == FFTJH_KIND_C2R)
      L = layout_c2r(dtype, rank, dims, batch, 1);
    else fail(FFTJH_ERR_ARGUMENT, "unknown kind");
    lean_rows(L);
    stream_hints(L);
    copy_out(describe(L), buf, len);
  });
}

int fftjh_describe_tuned(int dtype, int rank, const int64_t* dims, int64_t batch, int sign,
                         const fftjh_tuning* tune, char* buf, int64_t len) {
  return guarded([&] {
    Layout L = layout_c2c(dtype, rank, dims, batch, sign, 1.0, tune);
    lean_rows(L);
    stream_hints(L);
    copy_out(describe(L), buf, len);
  });
}

int fftjh_plan_destroy(fftjh_plan plan) {
  return guarded([&] { delete plan; });
}

const char* fftjh_last_error(void) { return g_err.c_str(); }

int fftjh_version(void) { return 100; }

int fftjh_profile_launches(fftjh_plan plan, void* y, const void* x, fftjh_stream stream, int iters, double* ms,
                           int* nlaunch) {
  return guarded([&] {
    if (!plan) fail(FFTJH_ERR_ARGUMENT, "plan is NULL");
    if (iters < 1) fail(FFTJH_ERR_ARGUMENT, "iters must be positive");
    const size_t ns = plan->L.steps.size();
    const int cap = nlaunch ? *nlaunch : 0;
    if (nlaunch) *nlaunch = static_cast<int>(ns);
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::vector<hipEvent_t> ev(2 * ns * iters);
    for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    std::vector<double> acc(ns, 0.0);
    try {
      for (int it = 0; it < iters; ++it) run(*plan, y, x, st, ev.data() + 2 * ns * it);
      hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
      for (int it = 0; it < iters; ++it)
        for (size_t i = 0; i < ns; ++i) {
          float t = 0.f;
          hip_check(hipEventElapsedTime(&t, ev[2 * ns * it + 2 * i], ev[2 * ns * it + 2 * i + 1]), "elapsed");
          acc[i] += t;
        }
    } catch (...) {
      for (auto& e : ev) (void)hipEventDestroy(e);
      throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (size_t i = 0; i < ns && static_cast<int>(i) < cap; ++i) ms[i] = acc[i] / iters;
  });
}

}  // extern "C"
