// vmc_rrw_f32.hip -- multi-wavefront register-resident kernels, mixed precision (FP32 working set).
#include "vmc_rrw_impl.hpp"

namespace vmc {
bool rrw_launch_f32(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                   const ReinvLaunch &reinv) {
    return rrw_launch<float>(d, st, grid, n_sweeps, s0, flags, bs, stream, timer, reinv);
}
}  // namespace vmc
