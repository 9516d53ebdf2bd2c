// Ragged batch: thousands of independent complex transforms of uneven
// lengths, packed back to back with an offset table, in ONE persistent
// launch.  Each workgroup walks its own list of whole transforms (assigned
// on the host, longest predicted cost first, to the least-loaded workgroup),
// so every transform's arithmetic is the same whichever workgroup runs it.
//
// Per transform, by kind:
//   DIRECT    n fits one LDS tile: load, all stages in LDS, store.
//   TWO_LEVEL n = n1*n2 (four-step inside one workgroup): column tiles
//             (length n2, twiddle w_n^{j1*k2}) into the workgroup's
//             workspace, then row tiles (length n1) to natural order.
//   RS_DIRECT non-smooth n, resampled onto m that fits one tile: the
//             deconvolution lands in LDS, the m-point transform runs there,
//             the w-tap interpolation reads it (resample.jl:258-276).
//   RS_TWO    the same with m two-level through the workspace.
// Two launches of this kernel share the batch: items whose tile length is
// at most 4096 points on 256-thread workgroups (4 per CU), longer ones on
// 1024-thread workgroups with a 16384-point tile (FP64: 512 threads, 8192),
// one per CU -- so items up to 16384 points run in one tile and the longer
// ones take 64-column tiles (512-byte runs) in their two phases.  Measured
// on the 10,000-transform benchmark batch: 1.26 -> 1.16 ms, HBM traffic
// 4.04 -> 2.86 GB.
// The reference has no ragged API: its baseline is a loop of 1-D plans.
#include "butterflies.h"
#include "common.h"
#include "launch.h"
#include "pass_kernel.h"
#include "passcommon.h"
#include "ragged.h"
#include "stages.h"

namespace fftjh {

namespace {

template <typename T, int DIR, int kRT>
__device__ void tile_stages(LdsPtr<T> lds, const RaggedSched& sc, int C, int logC, TwPtr<T> tw) {
  const int tid = static_cast<int>(threadIdx.x);
  const int c = tid & (C - 1), jt = tid >> logC, jstep = kRT >> logC;
  for (int s = 0; s < sc.nstages; ++s) {
    const StageDesc& st = sc.st[s];
    switch (st.radix) {
      case 2: run_stage<T, DIR, 2>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 3: run_stage<T, DIR, 3>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 4: run_stage<T, DIR, 4>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 5: run_stage<T, DIR, 5>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 7: run_stage<T, DIR, 7>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 8: run_stage<T, DIR, 8>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 11: run_stage<T, DIR, 11>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 13: run_stage<T, DIR, 13>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      default: run_stage<T, DIR, 16>(sc.n, C, st, lds, tw, c, jt, jstep); break;
    }
  }
}

// Tile load: this thread's elements u < kTile/kRT of a tile, element u
// at src[base + u*step], present while i0 + u*js < lim (all when `ok`),
// into LDS slot tid + u*kRT.  Every load is issued before the first LDS
// store (one load in flight per thread left the two-level items latency
// bound); out of line, so its registers are its own.
template <typename T, int kRT, int kTile, typename P>
__device__ __noinline__ void tile_load(P src_, const int64_t base, const int64_t step_, const int i0, const int js_,
                                       const int lim_, const bool ok, LdsPtr<T> lds_) {
  constexpr int kU = kTile / kRT;
  const int tid = static_cast<int>(threadIdx.x);
  // (src, step, js, lim and lds are uniform: scalar registers)
  const P src = uni_ptr(src_);
  const int64_t step = static_cast<int64_t>(
      (static_cast<uint64_t>(static_cast<unsigned>(uni(static_cast<int>(static_cast<uint64_t>(step_) >> 32)))) << 32) |
      static_cast<unsigned>(uni(static_cast<int>(step_ & 0xffffffff))));
  const int js = uni(js_), lim = uni(lim_);
  const LdsPtr<T> lds = uni_lds<T>(lds_) + lpad(tid);
  // rows present: u < cnt
  const int cnt = i0 < lim ? (lim - i0 + js - 1) / js : 0;
  const int got = ok ? cnt : 0;
  V2<T> v[kU];
  P p = src + base;
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    v[u] = mk<T>(T(0), T(0));
    if (u < got) v[u] = ldv<T>(p, 0);
    p += step;
  }
  // lpad(tid + u*kRT) = lpad(tid) + u*(kRT + kRT/16)
#pragma unroll
  for (int u = 0; u < kU; ++u)
    if (u < cnt) stv(lds, u * (kRT + kRT / 16), v[u]);
}

// Phase-A epilogue of a column tile: element f = tid + u*kRT of the tile
// (column c0 + (f & (C-1)), row k2 = f >> logC) times w_n^{col*k2} into
// mid[col + n1*k2].  All twiddle-table loads are issued before the first
// use (a load and its wait per element left the loop latency bound).
template <typename T, int kRT, int kTile>
__device__ __noinline__ void tile_store_tw(LdsPtr<T> lds_, GPtr<T> mid_, TwPtr<T> thi_, TwPtr<T> tlo_,
                                           const int lo_bits_, const int n1_, const int total_, const int C_,
                                           const int logC_, const int c0_) {
  constexpr int kU = kTile / kRT;
  const int tid = static_cast<int>(threadIdx.x);
  const int lo_bits = uni(lo_bits_), n1 = uni(n1_), total = uni(total_), C = uni(C_), logC = uni(logC_);
  const int c0 = uni(c0_);
  const unsigned lo_mask = (1u << lo_bits) - 1u;
  using TwBytes = const __attribute__((address_space(1))) char*;
  const TwBytes thi = reinterpret_cast<TwBytes>(uni_ptr(thi_));
  const TwBytes tlo = reinterpret_cast<TwBytes>(uni_ptr(tlo_));
  const LdsPtr<T> lds = uni_lds<T>(lds_) + lpad(tid);
  const int col = c0 + (tid & (C - 1));
  const int k20 = tid >> logC, ks = kRT >> logC;
  // rows of this thread present in the tile (the column itself must exist)
  const int cnt = (col < n1 && tid < total) ? (total - tid + kRT - 1) / kRT : 0;
  const unsigned e0 = static_cast<unsigned>(col) * static_cast<unsigned>(k20);
  const unsigned de = static_cast<unsigned>(col) * static_cast<unsigned>(ks);
  NV2<T> a[kU], b[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    if (u < cnt) {
      const unsigned e = e0 + static_cast<unsigned>(u) * de;
      a[u] = *reinterpret_cast<TwPtr<T>>(thi + (e >> lo_bits) * static_cast<unsigned>(sizeof(NV2<T>)));
      b[u] = *reinterpret_cast<TwPtr<T>>(tlo + (e & lo_mask) * static_cast<unsigned>(sizeof(NV2<T>)));
    }
  }
  GPtr<T> mp = uni_ptr(mid_) + (col + static_cast<int64_t>(n1) * k20);
  const int64_t mstep = static_cast<int64_t>(n1) * ks;
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    if (u < cnt) {
      const V2<T> w = cmul<T>(mk<T>(a[u].x, a[u].y), mk<T>(b[u].x, b[u].y));
      stv(mp, 0, cmul<T>(ldv<T>(lds, u * (kRT + kRT / 16)), w));
    }
    mp += mstep;
  }
}

// length-n transform src -> dst (both contiguous, global) as two tile phases
// through `mid` (global, n elements); dst may equal src
template <typename T, int DIR, int kRT, int kTile>
__device__ void two_level(LdsPtr<T> lds, const RaggedItem& it, const RaggedSched* scheds, TwPtr<T> tw,
                          const V2<T>* src_, V2<T>* mid_, V2<T>* dst_, T scale) {
  const TwPtr<T> src = as_tw<T>(src_);
  const GPtr<T> mid = as_g<T>(mid_);
  const GPtr<T> dst = as_g<T>(dst_);
  const int tid = static_cast<int>(threadIdx.x);
  const int n1 = it.n1, n2 = it.n2;
  const RaggedSched& sa = scheds[it.sA];
  const RaggedSched& sb = scheds[it.sB];
  TwPtr<T> thi = tw + it.tw_hi;
  TwPtr<T> tlo = tw + it.tw_lo;
  // phase A: columns j1 (stride n1) of length n2, twiddle w_n^{j1*k2}, mid[j1 + n1*k2]
  {
    const int C = it.cA, logC = it.logcA;
    const int ca = tid & (C - 1), ia = tid >> logC, ja = kRT >> logC;
    for (int c0 = 0; c0 < n1; c0 += C) {
      tile_load<T, kRT, kTile>(src, c0 + ca + static_cast<int64_t>(n1) * ia, static_cast<int64_t>(n1) * ja, ia,
                               ja, n2, c0 + ca < n1, lds);
      __syncthreads();
      tile_stages<T, DIR, kRT>(lds, sa, C, logC, tw);
      tile_store_tw<T, kRT, kTile>(lds, mid, thi, tlo, it.lo_bits, n1, C * n2, C, logC, c0);
      __syncthreads();
    }
  }
  __threadfence_block();
  __syncthreads();
  // phase B: rows k2 of mid (length n1), output dst[k2 + n2*k1]
  {
    const int C = it.cB, logC = it.logcB;
    const int cb = tid & (C - 1), ib = tid >> logC, jb = kRT >> logC;
    for (int r0 = 0; r0 < n2; r0 += C) {
      tile_load<T, kRT, kTile>(mid, static_cast<int64_t>(r0 + cb) * n1 + ib, jb, ib, jb, n1, r0 + cb < n2, lds);
      __syncthreads();
      tile_stages<T, DIR, kRT>(lds, sb, C, logC, tw);
      for (int f = tid; f < C * n1; f += kRT) {
        const int c = f & (C - 1), k1 = f >> logC;
        const int row = r0 + c;
        if (row < n2) stv(dst, row + static_cast<int64_t>(n2) * k1, cscale<T>(ldv<T>(lds, lpad(k1 * C + c)), scale));
      }
      __syncthreads();
    }
  }
  __threadfence_block();
  __syncthreads();
}

__device__ __forceinline__ double es_weight(double z, double beta) {
  return fabs(z) >= 1.0 ? 0.0 : exp(beta * (sqrt(1.0 - z * z) - 1.0));
}
// the same weight in single precision, for FP32 plans (relative error a few
// ulp of FP32, below the plans' tolerance)
__device__ __forceinline__ float es_weight(float z, float beta) {
  return fabsf(z) >= 1.0f ? 0.0f : expf(beta * (sqrtf(1.0f - z * z) - 1.0f));
}

// y[k] = sum_{i<w} psi((delta + h - 1 - i)/h) * G[(q - h + 1 + i) mod m],  k*m = q*n + r
template <typename T, bool FROM_LDS, int kRT>
__device__ void interp(const RaggedItem& it, const V2<T>* G, V2<T>* y) {
  const int n = it.n, m = it.m, w = it.w, h = it.w / 2;
  const T beta = static_cast<T>(it.beta), ih = T(1) / static_cast<T>(h);
  for (int k = static_cast<int>(threadIdx.x); k < n; k += kRT) {
    const int64_t km = static_cast<int64_t>(k) * m;
    const int64_t q = km / n, r = km - q * n;
    const T delta = static_cast<T>(static_cast<double>(r) / n);
    int l = static_cast<int>(q) - h + 1;
    if (l < 0) l += m;
    V2<T> acc = mk<T>(T(0), T(0));
    for (int i = 0; i < w; ++i) {
      const T wt = es_weight((delta + static_cast<T>(h - 1 - i)) * ih, beta);
      int li = l + i;
      if (li >= m) li -= m;
      const V2<T> g = FROM_LDS ? G[lpad(li)] : G[li];
      acc.x = fma(wt, g.x, acc.x);
      acc.y = fma(wt, g.y, acc.y);
    }
    y[k] = acc;
  }
}

}  // namespace

template <typename T, int DIR, int kRT, int kTile>
__global__ void __launch_bounds__(kRT)
k_ragged(const RaggedItem* __restrict__ items, const int32_t* __restrict__ wg_ptr,
         const int32_t* __restrict__ wg_items, const RaggedSched* __restrict__ scheds, const V2<T>* tw,
         const T* __restrict__ corr, const V2<T>* x, V2<T>* y, V2<T>* ws, const int64_t* __restrict__ wg_ws) {
  extern __shared__ __align__(16) unsigned char smem[];
  V2<T>* lds = reinterpret_cast<V2<T>*>(smem);
  const LdsPtr<T> l3 = as_lds<T>(lds);
  const TwPtr<T> t1 = as_tw<T>(tw);
  const int tid = static_cast<int>(threadIdx.x);
  const int wg = static_cast<int>(blockIdx.x);
  // this workgroup's workspace: two-level items use n elements, two-level
  // resampling items 2m (grid, then the phase-A intermediate)
  V2<T>* wsA = ws + wg_ws[wg];
  for (int p = wg_ptr[wg]; p < wg_ptr[wg + 1]; ++p) {
    const RaggedItem& it = items[wg_items[p]];
    const V2<T>* xi = x + it.off;
    V2<T>* yi = y + it.off;
    const T scale = static_cast<T>(it.scale);
    const int n = it.n;
    if (it.kind == kRaggedDirect) {
      const RaggedSched& sc = scheds[it.sA];
      tile_load<T, kRT, kTile>(as_tw<T>(xi), tid, kRT, tid, kRT, n, true, l3);
      __syncthreads();
      tile_stages<T, DIR, kRT>(l3, sc, 1, 0, t1);
      for (int k = tid; k < n; k += kRT) yi[k] = cscale<T>(lds[lpad(k)], scale);
      __syncthreads();
    } else if (it.kind == kRaggedTwoLevel) {
      two_level<T, DIR, kRT, kTile>(l3, it, scheds, t1, xi, wsA, yi, scale);
    } else {
      // resampling: corr carries the scale; the m-point transform is unscaled
      const int m = it.m, hp = (n + 1) / 2, hi0 = m - n + hp;
      const T* cr = corr + it.corr_off;
      if (it.kind == kRaggedRsDirect) {
        for (int j = tid; j < m; j += kRT) {
          const int src = j < hp ? j : (j >= hi0 ? j - (m - n) : -1);
          lds[lpad(j)] = src >= 0 ? cscale<T>(xi[src], cr[src]) : mk<T>(T(0), T(0));
        }
        __syncthreads();
        tile_stages<T, DIR, kRT>(l3, scheds[it.sA], 1, 0, t1);
        interp<T, true, kRT>(it, lds, yi);
        __syncthreads();
      } else {
        V2<T>* grid = wsA;
        V2<T>* mid = wsA + m;
        for (int j = tid; j < m; j += kRT) {
          const int src = j < hp ? j : (j >= hi0 ? j - (m - n) : -1);
          grid[j] = src >= 0 ? cscale<T>(xi[src], cr[src]) : mk<T>(T(0), T(0));
        }
        __threadfence_block();
        __syncthreads();
        two_level<T, DIR, kRT, kTile>(l3, it, scheds, t1, grid, mid, grid, T(1));
        interp<T, false, kRT>(it, grid, yi);
        __threadfence_block();
        __syncthreads();
      }
    }
  }
}

// Tile-granular passes over the long two-level items: one workgroup per
// tile, every tile of every item in one launch per phase, so that small
// workgroups (several per CU) share the load at tile granularity instead of
// one workgroup per CU walking whole items.  Phase A: the column block
// c0..c0+C-1 of an item (columns j1 at stride n1, length n2) from x, twiddle
// w_n^{j1*k2}, into the item's workspace region mid[j1 + n1*k2].  Phase B:
// the row block r0..r0+C-1 of mid (rows k2, length n1) to y[k2 + n2*k1].
// The first stage of a tile's chain runs on the registers its HBM loads
// fill and the last one stores from registers (the pass kernel's fused
// edges; the phase-A twiddle from an FP64 start and steps), so a tile of
// two stages makes one LDS round trip.
template <typename T, int DIR>
__device__ __forceinline__ void tile_first(const int radix, const int n, const int C, TwPtr<T> xb, const int64_t ist,
                                           const bool valid, LdsPtr<T> lds, const int c, const int jt,
                                           const int jstep) {
#define FFTJH_RF(R_) first_stage<T, DIR, R_>(n, C, xb, ist, valid, lds, c, jt, jstep)
  switch (radix) {
    case 2: FFTJH_RF(2); break;
    case 3: FFTJH_RF(3); break;
    case 4: FFTJH_RF(4); break;
    case 5: FFTJH_RF(5); break;
    case 6: FFTJH_RF(6); break;
    case 9: FFTJH_RF(9); break;
    case 10: FFTJH_RF(10); break;
    case 12: FFTJH_RF(12); break;
    case 14: FFTJH_RF(14); break;
    case 15: FFTJH_RF(15); break;
    case 7: FFTJH_RF(7); break;
    case 8: FFTJH_RF(8); break;
    case 11: FFTJH_RF(11); break;
    case 13: FFTJH_RF(13); break;
    default: FFTJH_RF(16); break;
  }
#undef FFTJH_RF
}

template <typename T, int DIR>
__device__ __forceinline__ void tile_last(const StageDesc& sl, const int n, const int C, LdsPtr<T> lds, TwPtr<T> tw,
                                          GPtr<T> yb, const int64_t ost, const bool valid, const T scale,
                                          const bool twon, const double2 w0, const double2 Sr, const double2 Sb,
                                          const int c, const int jt, const int jstep) {
  const bool scaled = scale != T(1);
#define FFTJH_RL(R_) \
  last_stage<T, DIR, R_>(n, C, sl, lds, tw, yb, ost, valid, scale, scaled, twon, w0, Sr, Sb, c, jt, jstep)
  switch (sl.radix) {
    case 2: FFTJH_RL(2); break;
    case 3: FFTJH_RL(3); break;
    case 4: FFTJH_RL(4); break;
    case 5: FFTJH_RL(5); break;
    case 6: FFTJH_RL(6); break;
    case 9: FFTJH_RL(9); break;
    case 10: FFTJH_RL(10); break;
    case 12: FFTJH_RL(12); break;
    case 14: FFTJH_RL(14); break;
    case 15: FFTJH_RL(15); break;
    case 7: FFTJH_RL(7); break;
    case 8: FFTJH_RL(8); break;
    case 11: FFTJH_RL(11); break;
    case 13: FFTJH_RL(13); break;
    default: FFTJH_RL(16); break;
  }
#undef FFTJH_RL
}

// the stages [FROM, S-1) of a chain in LDS (FROM = 1: after a fused first stage)
template <typename T, int DIR, int kRT, int FROM = 1>
__device__ void tile_mid(LdsPtr<T> lds, const RaggedSched& sc, int C, int logC, TwPtr<T> tw) {
  const int tid = static_cast<int>(threadIdx.x);
  const int c = tid & (C - 1), jt = tid >> logC, jstep = kRT >> logC;
  for (int s = FROM; s < sc.nstages - 1; ++s) {
    const StageDesc& st = sc.st[s];
    switch (st.radix) {
      case 2: run_stage<T, DIR, 2>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 3: run_stage<T, DIR, 3>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 4: run_stage<T, DIR, 4>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 5: run_stage<T, DIR, 5>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 6: run_stage<T, DIR, 6>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 9: run_stage<T, DIR, 9>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 10: run_stage<T, DIR, 10>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 12: run_stage<T, DIR, 12>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 14: run_stage<T, DIR, 14>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 15: run_stage<T, DIR, 15>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 7: run_stage<T, DIR, 7>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 8: run_stage<T, DIR, 8>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 11: run_stage<T, DIR, 11>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      case 13: run_stage<T, DIR, 13>(sc.n, C, st, lds, tw, c, jt, jstep); break;
      default: run_stage<T, DIR, 16>(sc.n, C, st, lds, tw, c, jt, jstep); break;
    }
  }
}

#ifndef FFTJH_RAGGED_PASS_WAVES
#define FFTJH_RAGGED_PASS_WAVES 4
#endif
// Coalesced load of a row tile: the tile's rows r0..r0+C-1 of length n1 are
// `total` contiguous elements at src (C*n1 in all; the rows past the item's
// end load as zeros); element f = row*n1 + e goes to LDS slot e*C + row.
// (Lanes along the rows: one row per lane would touch a line per lane.)
template <typename T, int kRT, int kTile>
__device__ __noinline__ void tile_load_rows(GPtr<T> src_, const int n1_, const int C_, const int total_,
                                            LdsPtr<T> lds_) {
  constexpr int kU = kTile / kRT;
  const int tid = static_cast<int>(threadIdx.x);
  const GPtr<T> src = uni_ptr(src_);
  const int n1 = uni(n1_), C = uni(C_), total = uni(total_);
  const LdsPtr<T> lds = uni_lds<T>(lds_);
  const int all = C * n1;
  // row = f / n1 by a multiply-high: exact for f < 2^20 (f, n1 <= kTile)
  const unsigned magic = 0xffffffffu / static_cast<unsigned>(n1) + 1u;
  V2<T> v[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int f = tid + u * kRT;
    v[u] = f < total ? ldv<T>(src, f) : mk<T>(T(0), T(0));
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int f = tid + u * kRT;
    if (f < all) {
      const int row = static_cast<int>(__umulhi(static_cast<unsigned>(f), magic));
      const int e = f - row * n1;
      stv(lds, lpad(e * C + row), v[u]);
    }
  }
}

// w_n^e (e < 2^32) from the item's FP64 two-level tables: w_n^(i*2^lo_bits)
// at hi[i], w_n^i at lo[i]; e reduced exactly mod n
__device__ __forceinline__ double2 ragged_w(const double2* __restrict__ t64, const RaggedItem& it, uint32_t e) {
  e %= static_cast<uint32_t>(it.n);
  const uint32_t mask = (1u << it.lo_bits) - 1u;
  return cmul64(t64[it.t64_hi + (e >> it.lo_bits)], t64[it.t64_lo + (e & mask)]);
}

template <typename T, int DIR, int PH>
__global__ void __launch_bounds__(kRaggedThreads) __attribute__((amdgpu_waves_per_eu(FFTJH_RAGGED_PASS_WAVES)))
k_ragged_pass(const RaggedItem* __restrict__ items, const RaggedTile* __restrict__ tiles,
              const RaggedSched* __restrict__ scheds, const V2<T>* tw, const double2* __restrict__ tw64,
              const V2<T>* x, V2<T>* y, V2<T>* ws) {
  constexpr int kRT = kRaggedThreads, kTile = kRaggedTile;
  extern __shared__ __align__(16) unsigned char smem[];
  V2<T>* lds = reinterpret_cast<V2<T>*>(smem);
  const LdsPtr<T> l3 = as_lds<T>(lds);
  const TwPtr<T> t1 = as_tw<T>(tw);
  const int tid = static_cast<int>(threadIdx.x);
  const RaggedTile tl = tiles[blockIdx.x];
  if (tl.item < 0) return;  // an empty slot of the XCD-local order
  const RaggedItem& it = items[tl.item];
  const int n1 = it.n1, n2 = it.n2;
  V2<T>* mid = ws + it.woff;
  if constexpr (PH == 0) {
    const int C = it.cA, logC = it.logcA;
    const int c = tid & (C - 1), jt = tid >> logC, jstep = kRT >> logC;
    const int c0 = tl.first;
    const RaggedSched& sc = scheds[it.sA];
    const bool valid = c0 + c < n1;
    if (sc.nstages < 2) {
      tile_load<T, kRT, kTile>(as_tw<T>(x + it.off), c0 + c + static_cast<int64_t>(n1) * jt,
                               static_cast<int64_t>(n1) * jstep, jt, jstep, n2, valid, l3);
      __syncthreads();
      tile_stages<T, DIR, kRT>(l3, sc, C, logC, t1);
      tile_store_tw<T, kRT, kTile>(l3, as_g<T>(mid), t1 + it.tw_hi, t1 + it.tw_lo, it.lo_bits, n1, C * n2, C, logC,
                                   c0);
      return;
    }
    tile_first<T, DIR>(sc.st[0].radix, n2, C, as_tw<T>(x + it.off + c0 + c), n1, valid, l3, c, jt, jstep);
    tile_mid<T, DIR, kRT>(l3, sc, C, logC, t1);
    // outputs k2 = j + r*ns of column j1 = c0 + c times w_n^(j1*k2)
    const StageDesc& sl = sc.st[sc.nstages - 1];
    const uint32_t j1 = static_cast<uint32_t>(c0 + c);  // (j1 < n1 <= 4096, jt, ns, jstep < 4096)
    const double2 w0 = ragged_w(tw64, it, j1 * static_cast<uint32_t>(jt));
    const double2 Sr = ragged_w(tw64, it, j1 * static_cast<uint32_t>(sl.ns));
    const double2 Sb = ragged_w(tw64, it, j1 * static_cast<uint32_t>(jstep));
    tile_last<T, DIR>(sl, n2, C, l3, t1, as_g<T>(mid + c0 + c), n1, valid, T(1), true, w0, Sr, Sb, c, jt, jstep);
  } else {
    const int C = it.cB, logC = it.logcB;
    const int c = tid & (C - 1), jt = tid >> logC, jstep = kRT >> logC;
    const int r0 = tl.first;
    const RaggedSched& sc = scheds[it.sB];
    const bool valid = r0 + c < n2;
    const T scale = static_cast<T>(it.scale);
    V2<T>* yi = y + it.off;
    if (sc.nstages < 2) {
      tile_load<T, kRT, kTile>(as_g<T>(mid), static_cast<int64_t>(r0 + c) * n1 + jt, jstep, jt, jstep, n1, valid, l3);
      __syncthreads();
      tile_stages<T, DIR, kRT>(l3, sc, C, logC, t1);
      for (int f = tid; f < C * n1; f += kRT) {
        const int cc = f & (C - 1), k1 = f >> logC;
        const int row = r0 + cc;
        if (row < n2) yi[row + static_cast<int64_t>(n2) * k1] = cscale<T>(lds[lpad(k1 * C + cc)], scale);
      }
      return;
    }
    const int rows = n2 - r0 < C ? n2 - r0 : C;
    tile_load_rows<T, kRT, kTile>(as_g<T>(mid + static_cast<int64_t>(r0) * n1), n1, C, rows * n1, l3);
    __syncthreads();
    tile_mid<T, DIR, kRT, 0>(l3, sc, C, logC, t1);
    const double2 one = make_double2(1.0, 0.0);
    tile_last<T, DIR>(sc.st[sc.nstages - 1], n1, C, l3, t1, as_g<T>(yi + r0 + c), n2, valid, scale, false, one, one,
                      one, c, jt, jstep);
  }
}

template <typename T, int DIR, int PH>
static hipError_t launch_pass_t(const RaggedPassLaunch& L, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ragged_pass<T, DIR, PH>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    attr = true;
  }
  const size_t lds = static_cast<size_t>(lds_elems(kRaggedTile)) * sizeof(V2<T>);
  hipLaunchKernelGGL((k_ragged_pass<T, DIR, PH>), dim3(L.tiles), dim3(kRaggedThreads), lds, st,
                     static_cast<const RaggedItem*>(L.items), static_cast<const RaggedTile*>(L.tile_list),
                     static_cast<const RaggedSched*>(L.scheds), static_cast<const V2<T>*>(L.tw),
                     static_cast<const double2*>(L.tw64), static_cast<const V2<T>*>(L.x), static_cast<V2<T>*>(L.y),
                     static_cast<V2<T>*>(L.ws));
  return hipGetLastError();
}

template <typename T, int DIR>
static hipError_t launch_pass_d(const RaggedPassLaunch& L, hipStream_t st) {
  return L.phase == 0 ? launch_pass_t<T, DIR, 0>(L, st) : launch_pass_t<T, DIR, 1>(L, st);
}

hipError_t launch_ragged_pass(int dtype, int dir, const RaggedPassLaunch& L, hipStream_t st) {
  if (L.tiles <= 0) return hipSuccess;
  if (dtype == 0) return dir < 0 ? launch_pass_d<float, -1>(L, st) : launch_pass_d<float, 1>(L, st);
  return dir < 0 ? launch_pass_d<double, -1>(L, st) : launch_pass_d<double, 1>(L, st);
}

template <typename T, int DIR, int kRT, int kTile>
static hipError_t launch_ragged_t(const RaggedLaunch& L, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ragged<T, DIR, kRT, kTile>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    attr = true;
  }
  const size_t lds = static_cast<size_t>(lds_elems(kTile)) * sizeof(V2<T>);
  hipLaunchKernelGGL((k_ragged<T, DIR, kRT, kTile>), dim3(L.groups), dim3(kRT), lds, st,
                     static_cast<const RaggedItem*>(L.items), static_cast<const int32_t*>(L.wg_ptr),
                     static_cast<const int32_t*>(L.wg_items), static_cast<const RaggedSched*>(L.scheds),
                     static_cast<const V2<T>*>(L.tw), static_cast<const T*>(L.corr),
                     static_cast<const V2<T>*>(L.x), static_cast<V2<T>*>(L.y), static_cast<V2<T>*>(L.ws),
                     static_cast<const int64_t*>(L.wg_ws));
  return hipGetLastError();
}

template <typename T, int DIR>
static hipError_t launch_ragged_d(const RaggedLaunch& L, hipStream_t st) {
  constexpr int dt = sizeof(T) == 4 ? 0 : 1;
  if (L.big) return launch_ragged_t<T, DIR, ragged_big_threads(dt), ragged_big_tile(dt)>(L, st);
  return launch_ragged_t<T, DIR, kRaggedThreads, kRaggedTile>(L, st);
}

hipError_t launch_ragged(int dtype, int dir, const RaggedLaunch& L, hipStream_t st) {
  if (dtype == 0) return dir < 0 ? launch_ragged_d<float, -1>(L, st) : launch_ragged_d<float, 1>(L, st);
  return dir < 0 ? launch_ragged_d<double, -1>(L, st) : launch_ragged_d<double, 1>(L, st);
}

template <typename T, int DIR, int kRT, int kTile>
static int resident_t() {
  const void* fn = reinterpret_cast<const void*>(&k_ragged<T, DIR, kRT, kTile>);
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
  const size_t lds = static_cast<size_t>(lds_elems(kTile)) * sizeof(V2<T>);
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kRT, lds) != hipSuccess) return 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return per_cu * cus;
}

template <typename T, int DIR>
static int resident_d(int big) {
  constexpr int dt = sizeof(T) == 4 ? 0 : 1;
  return big ? resident_t<T, DIR, ragged_big_threads(dt), ragged_big_tile(dt)>()
             : resident_t<T, DIR, kRaggedThreads, kRaggedTile>();
}

int ragged_resident_groups(int dtype, int dir, int big) {
  if (dtype == 0) return dir < 0 ? resident_d<float, -1>(big) : resident_d<float, 1>(big);
  return dir < 0 ? resident_d<double, -1>(big) : resident_d<double, 1>(big);
}

}  // namespace fftjh
