// vmc_rr_f64.hip -- register-resident sweep, plain FP64.
#include "vmc_rr_impl.hpp"

namespace vmc {
bool rr_launch_f64(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, int variant, hipStream_t stream,
                   KernelTimer &timer) {
    return rr_launch<double>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream, timer);
}
}  // namespace vmc
