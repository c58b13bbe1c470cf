// vmc_rr_f32.hip -- register-resident sweep, mixed precision (FP32 working set).
#include "vmc_rr_impl.hpp"

namespace vmc {
bool rr_launch_f32(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, int variant, hipStream_t stream,
                   KernelTimer &timer) {
    return rr_launch<float>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream, timer);
}
bool rr_batch_step_f32(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                       const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream) {
    return rr_batch_step<float>(shape, items, n_items, grid, sweep_ctr, flags, bs, reinv, stream);
}
bool rr_given_f32(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
                  hipStream_t stream) {
    return rr_given<float>(d, st, grid, given, stream);
}
bool rr_parts_f32(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream) {
    return rr_parts<float>(d, st, grid, parts, stream);
}
}  // namespace vmc
