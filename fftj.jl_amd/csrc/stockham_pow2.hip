// Public entry points of the power-of-two pass kernels; the kernels are
// instantiated per (precision, direction) in stockham_pow2_<t><dir>.hip so
// they compile in parallel.
#include "stockham_pow2.h"

namespace fftjh {

hipError_t launch_pow2_f32_fwd(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                               hipStream_t st);
hipError_t launch_pow2_f32_inv(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                               hipStream_t st);
hipError_t launch_pow2_f64_fwd(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                               hipStream_t st);
hipError_t launch_pow2_f64_inv(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                               hipStream_t st);
int resident_pow2_f32_fwd(const PassDesc& d);
int resident_pow2_f32_inv(const PassDesc& d);
int resident_pow2_f64_fwd(const PassDesc& d);
int resident_pow2_f64_inv(const PassDesc& d);

bool pow2_chain(int n, int* radices, int* count) {
  switch (n) {
#define X(N, A, B, C_, D_)                  \
  case N: {                                 \
    const int r[4] = {A, B, C_, D_};        \
    int k = 0;                              \
    for (int i = 0; i < 4; ++i)             \
      if (r[i] > 1) radices[k++] = r[i];    \
    *count = k;                             \
    return true;                            \
  }
    FFTJH_POW2_CHAINS(X)
#undef X
  }
  return false;
}

bool fixed_chain(int n, int* radices, int* count, int* elems) {
  if (pow2_chain(n, radices, count)) {
    *elems = 16;
    return true;
  }
  switch (n) {
#define X(N, A, B, C_, D_, E)               \
  case N: {                                 \
    const int r[4] = {A, B, C_, D_};        \
    int k = 0;                              \
    for (int i = 0; i < 4; ++i)             \
      if (r[i] > 1) radices[k++] = r[i];    \
    *count = k;                             \
    *elems = E;                             \
    return true;                            \
  }
    FFTJH_FIXED_CHAINS(X)
#undef X
  }
  return false;
}

hipError_t launch_pass_pow2(int dtype, int dir, const PassDesc& d, const PassDesc* dd, const void* x, void* y,
                            const void* tw, hipStream_t st) {
  if (dtype == 0)
    return dir < 0 ? launch_pow2_f32_fwd(d, dd, x, y, tw, st) : launch_pow2_f32_inv(d, dd, x, y, tw, st);
  return dir < 0 ? launch_pow2_f64_fwd(d, dd, x, y, tw, st) : launch_pow2_f64_inv(d, dd, x, y, tw, st);
}

int pass_pow2_resident_groups(int dtype, int dir, const PassDesc& d) {
  if (dtype == 0) return dir < 0 ? resident_pow2_f32_fwd(d) : resident_pow2_f32_inv(d);
  return dir < 0 ? resident_pow2_f64_fwd(d) : resident_pow2_f64_inv(d);
}

int pow2_max_threads(int dtype) { return dtype == 0 ? MaxThreadsP2<float>::value : MaxThreadsP2<double>::value; }

}  // namespace fftjh
