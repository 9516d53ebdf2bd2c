// Descriptors of the ragged persistent kernel (host planner <-> device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "common.h"

namespace fftjh {

constexpr int kRaggedThreads = 256;  // workgroup size of the small-item launch
constexpr int kRaggedTile = 4096;    // complex elements one LDS tile holds (small items)
// smooth items longer than this run as two tile-granular passes (default)
constexpr int kRaggedPassMin = 4096;
// the large-item launch: one workgroup per CU whose LDS tile holds 16384
// complex FP32 (8192 FP64) elements, so lengths up to that run in one tile
// and longer ones take 64-column tiles in their two phases
constexpr int ragged_big_threads(int dtype) { return dtype == 0 ? 1024 : 512; }
constexpr int ragged_big_tile(int dtype) { return dtype == 0 ? 16384 : 8192; }

enum RaggedKind : int32_t { kRaggedDirect = 0, kRaggedTwoLevel = 1, kRaggedRsDirect = 2, kRaggedRsTwo = 3 };

// stage chain of one tile length (tw_off into the shared twiddle pool)
struct RaggedSched {
  int32_t n;
  int32_t nstages;
  StageDesc st[kMaxStages];
};

struct RaggedItem {
  int64_t off;       // complex offset in x and y
  int32_t n;         // transform length
  int32_t kind;      // RaggedKind
  int32_t sA, sB;    // schedules: direct / column tiles (sA), row tiles (sB)
  int32_t n1, n2;    // two-level split of n (or of m): rows n1, columns n2
  int32_t cA, logcA; // columns per phase-A tile
  int32_t cB, logcB; // rows per phase-B tile
  int32_t lo_bits;   // two-level twiddle w^e = pool[tw_hi + (e >> lo_bits)] * pool[tw_lo + (e & mask)]
  int32_t m, w;      // resampling grid and taps
  int32_t pad_;
  int64_t tw_hi, tw_lo;
  int64_t corr_off;  // into the correction pool (reals)
  double beta;
  double scale;
  int64_t woff;      // tile-granular passes: the item's region of the workspace (n elements)
  int64_t t64_hi, t64_lo;  // tile-granular passes: FP64 two-level tables of w_n (shared per length)
};

// One tile of a tile-granular pass over long items: phase A tiles are
// column blocks c0 of an item's n1 columns, phase B tiles row blocks r0 of
// its n2 rows.
struct RaggedTile {
  int32_t item;
  int32_t first;  // c0 (phase A) or r0 (phase B)
};

struct RaggedPassLaunch {
  int phase;      // 0: columns of x into the workspace, 1: rows of the workspace into y
  int tiles;      // one workgroup per tile
  const void* items;
  const void* tile_list;
  const void* scheds;
  const void* tw;
  const void* tw64;  // FP64 two-level tables (double2) of the phase-A twiddle
  const void* x;
  void* y;
  void* ws;
};

struct RaggedLaunch {
  int big;  // 0: the small-item kernel, 1: the large-item kernel
  int groups;
  const void* items;
  const void* wg_ptr;
  const void* wg_items;
  const void* scheds;
  const void* tw;
  const void* corr;
  const void* x;
  void* y;
  void* ws;
  const void* wg_ws;  // int64 per workgroup: its workspace offset (complex elements)
};

hipError_t launch_ragged(int dtype, int dir, const RaggedLaunch& L, hipStream_t st);
hipError_t launch_ragged_pass(int dtype, int dir, const RaggedPassLaunch& L, hipStream_t st);
int ragged_resident_groups(int dtype, int dir, int big);

}  // namespace fftjh
