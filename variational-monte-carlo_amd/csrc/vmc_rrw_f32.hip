// vmc_rrw_f32.hip -- multi-wavefront register-resident kernels, mixed precision (FP32 working set).
#include "vmc_rrw_impl.hpp"

namespace vmc {
bool rrw_launch_f32(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                   const ReinvLaunch &reinv, double *inv64) {
    return rrw_launch<float>(d, st, grid, n_sweeps, s0, flags, bs, stream, timer, reinv, inv64);
}
bool rrw_batch_step_f32(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                       const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream,
                       const BatchItem *items_refresh, const BatchItem *items_m64) {
    return rrw_batch_step<float>(shape, items, n_items, grid, sweep_ctr, flags, bs, reinv, stream, items_refresh,
                               items_m64);
}
bool rrw_given_f32(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
                   hipStream_t stream, const ReinvLaunch &reinv) {
    return rrw_given<float>(d, st, grid, given, stream, reinv);
}
bool rrw_parts_f32(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream) {
    return rrw_parts<float>(d, st, grid, parts, stream);
}
bool rrw_parts_exact_f32(const DevSys &d, const DevState &st, unsigned grid, double *parts, double *inv64,
                         hipStream_t stream, const ReinvLaunch &reinv) {
    return rrw_parts_exact<float>(d, st, grid, parts, inv64, stream, reinv);
}
}  // namespace vmc
