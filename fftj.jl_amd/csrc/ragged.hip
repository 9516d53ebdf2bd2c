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
