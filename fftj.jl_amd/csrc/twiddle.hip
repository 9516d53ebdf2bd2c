// Plan-time twiddle tables, generated on the GPU in FP64 and rounded once to
// the plan's precision.
//
// Every entry is exp(sign * 2*pi*i * e / L) for an exact integer numerator e:
// the numerator is reduced modulo L in 64-bit integers, folded to
// (-L/2, L/2], and only then turned into an FP64 angle in units of pi for
// sincospi -- the reference's idiom (plan.jl:76-77, fourstep.jl:284,
// real.jl:31: integer numerator, one division, sincospi in Float64).
#include "butterflies.h"
#include "common.h"
#include "launch.h"
#include "stages_mat.h"

namespace fftjh {

__device__ __forceinline__ void expi_exact(int64_t e, int64_t L, int dir, double* c, double* s) {
  e %= L;
  if (e < 0) e += L;
  if (2 * e > L) e -= L;
  const double a = (2.0 * static_cast<double>(e)) / static_cast<double>(L);
  double sv, cv;
  sincospi(a, &sv, &cv);
  *c = cv;
  *s = dir < 0 ? -sv : sv;
}

// stage table: entry (r-1)*ns + k = w^{r*k}, w = exp(dir*2*pi*i/(ns*R))
template <typename T>
__global__ void k_tw_stage(V2<T>* out, int ns, int R, int dir) {
  const int64_t count = static_cast<int64_t>(R - 1) * ns;
  for (int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; idx < count;
       idx += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t r = idx / ns + 1, k = idx % ns;
    double c, s;
    expi_exact(r * k, static_cast<int64_t>(ns) * R, dir, &c, &s);
    out[idx] = mk<T>(static_cast<T>(c), static_cast<T>(s));
  }
}

// power table: out[i] = exp(dir*2*pi*i * (i*mult) / L)
template <typename T>
__global__ void k_tw_pow(V2<T>* out, int64_t count, int64_t mult, int64_t L, int dir) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double c, s;
    expi_exact((i % L) * (mult % L), L, dir, &c, &s);
    out[i] = mk<T>(static_cast<T>(c), static_cast<T>(s));
  }
}

// split-precision DFT-matrix operands (layout: stages_mat.h): the entry of
// output row m and input point p is w^{m p}, w = exp(dir*2*pi*i/R), from
// an exact reduction of m*p mod R; each value is split in FP64 into
// `terms` FP16 or BF16 parts (hi first) whose sum carries its significance
__global__ void k_mat_table(uint16_t* out, int mode, int dir) {
  const int terms = mat_terms(mode);
  const int total = mat_entries(terms) * 8;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
    const int e = idx & 7, ent = idx >> 3;
    int R, rel;
    if (ent < mat_base(16, terms)) {
      R = 9;
      rel = ent;
    } else if (ent < mat_base(32, terms)) {
      R = 16;
      rel = ent - mat_base(16, terms);
    } else {
      R = 32;
      rel = ent - mat_base(32, terms);
    }
    const int KT = mat_kt(R);
    const int lane = rel & 63;
    int q = rel >> 6;
    const int term = q % terms;
    q /= terms;
    const int comp = q & 1;
    q >>= 1;
    const int t = q % KT, h = q / KT;
    const int m = 16 * h + (lane & 15);
    const int p = 16 * t + 4 * (lane >> 4) + (e >> 1);
    const int part = e & 1;
    double v = 0.0;
    if (m < R && p < R) {
      double c, s;
      expi_exact(static_cast<int64_t>(m) * p, R, dir, &c, &s);
      // Re(y) = sum Re(w) xr - Im(w) xi ;  Im(y) = sum Im(w) xr + Re(w) xi
      v = comp == 0 ? (part == 0 ? c : -s) : (part == 0 ? s : c);
    }
    uint16_t bits = 0;
    for (int k = 0; k <= term; ++k) {
      if (mode == kMatF16) {
        const _Float16 hv = static_cast<_Float16>(static_cast<float>(v));
        bits = __builtin_bit_cast(uint16_t, hv);
        v -= static_cast<double>(hv);
      } else {
        const __bf16 bv = static_cast<__bf16>(static_cast<float>(v));
        bits = __builtin_bit_cast(uint16_t, bv);
        v -= static_cast<double>(static_cast<float>(bv));
      }
    }
    out[idx] = bits;
  }
}

int64_t mat_table_bytes(int mode) { return static_cast<int64_t>(mat_entries(mat_terms(mode))) * 16; }

hipError_t launch_mat_table(int mode, int dir, void* out, hipStream_t stream) {
  if (mode < kMatF16 || mode > kMatBF16x3) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mat_table, dim3(64), dim3(256), 0, stream, static_cast<uint16_t*>(out), mode, dir);
  return hipGetLastError();
}

static dim3 grid_for(int64_t count) {
  int64_t g = (count + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return dim3(static_cast<unsigned>(g));
}

hipError_t launch_tw_stage(int dtype, void* out, int ns, int R, int dir, hipStream_t stream) {
  const dim3 g = grid_for(static_cast<int64_t>(R - 1) * ns);
  if (dtype == 0)
    hipLaunchKernelGGL(k_tw_stage<float>, g, dim3(256), 0, stream, static_cast<float2*>(out), ns, R, dir);
  else
    hipLaunchKernelGGL(k_tw_stage<double>, g, dim3(256), 0, stream, static_cast<double2*>(out), ns, R, dir);
  return hipGetLastError();
}

hipError_t launch_tw_pow(int dtype, void* out, int64_t count, int64_t mult, int64_t L, int dir,
                         hipStream_t stream) {
  const dim3 g = grid_for(count);
  if (dtype == 0)
    hipLaunchKernelGGL(k_tw_pow<float>, g, dim3(256), 0, stream, static_cast<float2*>(out), count, mult, L, dir);
  else
    hipLaunchKernelGGL(k_tw_pow<double>, g, dim3(256), 0, stream, static_cast<double2*>(out), count, mult, L, dir);
  return hipGetLastError();
}

}  // namespace fftjh
