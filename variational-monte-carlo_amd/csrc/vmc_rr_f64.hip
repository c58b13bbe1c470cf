// vmc_rr_f64.hip -- register-resident sweep, plain FP64.
#include "vmc_rr_impl.hpp"

namespace vmc {
bool rr_launch_f64(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, int variant, hipStream_t stream,
                   KernelTimer &timer) {
    return rr_launch<double>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream, timer);
}
bool rr_batch_step_f64(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                       const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream) {
    return rr_batch_step<double>(shape, items, n_items, grid, sweep_ctr, flags, bs, reinv, stream);
}
bool rr_given_f64(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
                  hipStream_t stream) {
    return rr_given<double>(d, st, grid, given, stream);
}
bool rr_parts_f64(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream) {
    return rr_parts<double>(d, st, grid, parts, stream);
}
}  // namespace vmc
