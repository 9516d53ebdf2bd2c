// The pass kernel with split-precision MFMA stages (FP32 plans): radix 9,
// 16 and 32 stages run as DFT-matrix products on gfx950 MFMA with hi/lo
// split operands (stages_mat.h); the other radices of the chain, the loads,
// the stores and the four-step twiddles are those of the plain pass kernel.
#include "launch.h"
#include "pass_kernel.h"

namespace fftjh {

template <int DIR>
static hipError_t launch_mat_dir(int mode, const PassDesc& d, const PassDesc* dev_desc, const void* x, void* y,
                                 const void* tw, const void* mat, hipStream_t stream) {
  switch (mode) {
    case kMatF16: return launch_pass_t<float, DIR, kMatF16>(d, dev_desc, x, y, tw, mat, stream);
    case kMatBF16: return launch_pass_t<float, DIR, kMatBF16>(d, dev_desc, x, y, tw, mat, stream);
    case kMatBF16x3: return launch_pass_t<float, DIR, kMatBF16x3>(d, dev_desc, x, y, tw, mat, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pass_mat(int mode, int dir, const PassDesc& d, const PassDesc* dev_desc, const void* x, void* y,
                           const void* tw, const void* mat, hipStream_t stream) {
  return dir < 0 ? launch_mat_dir<-1>(mode, d, dev_desc, x, y, tw, mat, stream)
                 : launch_mat_dir<1>(mode, d, dev_desc, x, y, tw, mat, stream);
}

template <int DIR>
static int resident_mat_dir(int mode, int nthreads, size_t lds) {
  switch (mode) {
    case kMatF16: return resident_groups_t<float, DIR, kMatF16>(nthreads, lds);
    case kMatBF16: return resident_groups_t<float, DIR, kMatBF16>(nthreads, lds);
    case kMatBF16x3: return resident_groups_t<float, DIR, kMatBF16x3>(nthreads, lds);
    default: return 0;
  }
}

int pass_mat_resident_groups(int mode, int dir, const PassDesc& d) {
  const size_t lds = static_cast<size_t>(lds_elems(d.C * d.n)) * 8;
  return dir < 0 ? resident_mat_dir<-1>(mode, d.nthreads, lds) : resident_mat_dir<1>(mode, d.nthreads, lds);
}

}  // namespace fftjh
