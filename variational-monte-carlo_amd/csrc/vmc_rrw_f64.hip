// vmc_rrw_f64.hip -- multi-wavefront register-resident kernels, FP64.
#include "vmc_rrw_impl.hpp"

namespace vmc {
bool rrw_launch_f64(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                   const ReinvLaunch &reinv, double *inv64) {
    return rrw_launch<double>(d, st, grid, n_sweeps, s0, flags, bs, stream, timer, reinv, inv64);
}
bool rrw_batch_step_f64(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                       const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream,
                       const BatchItem *items_refresh, const BatchItem *items_m64) {
    return rrw_batch_step<double>(shape, items, n_items, grid, sweep_ctr, flags, bs, reinv, stream, items_refresh,
                               items_m64);
}
bool rrw_given_f64(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
                   hipStream_t stream, const ReinvLaunch &reinv) {
    return rrw_given<double>(d, st, grid, given, stream, reinv);
}
bool rrw_parts_f64(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream) {
    return rrw_parts<double>(d, st, grid, parts, stream);
}
bool rrw_measure64(const DevSys &d, const DevState &st64, unsigned grid, int bs, double *parts,
                   hipStream_t stream) {
    return rrw_measure64_impl(d, st64, grid, bs, parts, stream);
}
bool rrw_measure64_batch(const DevSys &shape, const BatchItem *items_m64, int n_items, unsigned grid, int bs,
                         hipStream_t stream) {
    return rrw_measure64_batch_impl(shape, items_m64, n_items, grid, bs, stream);
}
}  // namespace vmc
