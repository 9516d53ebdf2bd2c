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
