// Instantiations of the power-of-two pass kernels for double, direction 1.
#include "stockham_pow2.h"

namespace fftjh {

hipError_t launch_pow2_f64_inv(const PassDesc& d, const PassDesc* dd, const void* x, void* y, const void* tw,
                               hipStream_t st) {
  return dispatch_launch<double, 1>(d, dd, x, y, tw, st);
}

int resident_pow2_f64_inv(const PassDesc& d) { return dispatch_resident<double, 1>(d); }

}  // namespace fftjh
