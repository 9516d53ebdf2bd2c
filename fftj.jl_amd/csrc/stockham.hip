// The pass kernel: C whole length-n transforms per workgroup, all Stockham
// stages of the chain in one launch.
//
// Per workgroup:
//   load   HBM -> LDS, element i of column c to lds[i*C + c]; a wave reads
//          C columns x 64/C consecutive rows (row transforms) or 64/C rows x
//          C adjacent columns (strided axis passes, four-step pass A), so
//          every access is a contiguous run;
//   stages LDS -> registers -> LDS, in place, two barriers per stage;
//   store  LDS -> HBM with the scale and the optional four-step twiddle
//          folded in, same access pattern as the load.
// HBM is touched exactly once per element in each direction per pass.
//
// Reference: the stage loops of stockham.jl (stage4! :26-76, stage8! :81-130,
// stagep! :351-392, stage2! :394-428) run one radix pass per sweep over split
// planes; run_chain! (plan.jl:311-321) strings them together.  Here the whole
// chain runs inside one workgroup.
#include "launch.h"
#include "pass_kernel.h"

namespace fftjh {

int max_threads(int dtype) { return dtype == 0 ? MaxThreads<float>::value : MaxThreads<double>::value; }

// chains of two or more stages take the kernel with fused first and last stages
static bool fused(const PassDesc& d) { return d.fuse_io && d.nstages >= 2; }

int pass_resident_groups(int dtype, int dir, const PassDesc& d) {
  const size_t lds = static_cast<size_t>(lds_elems(d.C * d.n)) * (dtype == 0 ? 8 : 16);
  if (fused(d)) {
    if (dtype == 0)
      return dir < 0 ? resident_groups_fused_t<float, -1>(d.nthreads, lds)
                     : resident_groups_fused_t<float, 1>(d.nthreads, lds);
    return dir < 0 ? resident_groups_fused_t<double, -1>(d.nthreads, lds)
                   : resident_groups_fused_t<double, 1>(d.nthreads, lds);
  }
  if (dtype == 0)
    return dir < 0 ? resident_groups_t<float, -1, kMatNone>(d.nthreads, lds)
                   : resident_groups_t<float, 1, kMatNone>(d.nthreads, lds);
  return dir < 0 ? resident_groups_t<double, -1, kMatNone>(d.nthreads, lds)
                 : resident_groups_t<double, 1, kMatNone>(d.nthreads, lds);
}

hipError_t launch_pass(int dtype, int dir, const PassDesc& d, const PassDesc* dev_desc, const void* x,
                       void* y, const void* tw, hipStream_t stream) {
  if (fused(d)) {
    if (dtype == 0)
      return dir < 0 ? launch_pass_fused_t<float, -1>(d, dev_desc, x, y, tw, stream)
                     : launch_pass_fused_t<float, 1>(d, dev_desc, x, y, tw, stream);
    return dir < 0 ? launch_pass_fused_t<double, -1>(d, dev_desc, x, y, tw, stream)
                   : launch_pass_fused_t<double, 1>(d, dev_desc, x, y, tw, stream);
  }
  if (dtype == 0)
    return dir < 0 ? launch_pass_t<float, -1, kMatNone>(d, dev_desc, x, y, tw, nullptr, stream)
                   : launch_pass_t<float, 1, kMatNone>(d, dev_desc, x, y, tw, nullptr, stream);
  return dir < 0 ? launch_pass_t<double, -1, kMatNone>(d, dev_desc, x, y, tw, nullptr, stream)
                 : launch_pass_t<double, 1, kMatNone>(d, dev_desc, x, y, tw, nullptr, stream);
}

}  // namespace fftjh
