// vmc_rrw_f64.hip -- multi-wavefront register-resident kernels, FP64.
#include "vmc_rrw_impl.hpp"

namespace vmc {
bool rrw_launch_f64(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                   const ReinvLaunch &reinv) {
    return rrw_launch<double>(d, st, grid, n_sweeps, s0, flags, bs, stream, timer, reinv);
}
bool rrw_batch_step_f64(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                       const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream) {
    return rrw_batch_step<double>(shape, items, n_items, grid, sweep_ctr, flags, bs, reinv, stream);
}
}  // namespace vmc
