// vmc_rr_launch.hpp -- launchers of the register-resident sweep, one
// translation unit per working precision (vmc_rr_f32.hip, vmc_rr_f64.hip).
#pragma once

#include "vmc_device.hpp"

#include <hip/hip_runtime.h>

#include <vector>

namespace vmc {

// Kernel kinds timed by KernelTimer (vmc_ensemble_read_timing order).
enum { kTimeSweep = 0, kTimeReinv = 1, kTimeMeasure = 2, kTimeKinds = 3 };

// Device time per kernel kind from HIP events recorded on the launch stream
// around each launch while enabled; summed when read.
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> start, stop;
    std::vector<int> kind;
    std::vector<hipEvent_t> pool;

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t ev = pool.back();
            pool.pop_back();
            return ev;
        }
        hipEvent_t ev = nullptr;
        return hipEventCreate(&ev) == hipSuccess ? ev : nullptr;
    }
    // Records the start of a launch; returns false when not timing.
    bool begin(hipStream_t s, hipEvent_t &ev) {
        if (!on || !(ev = take())) return false;
        return hipEventRecord(ev, s) == hipSuccess;
    }
    void end(hipStream_t s, hipEvent_t ev0, int k) {
        hipEvent_t ev1 = take();
        if (!ev1 || hipEventRecord(ev1, s) != hipSuccess) return;
        start.push_back(ev0);
        stop.push_back(ev1);
        kind.push_back(k);
    }
    ~KernelTimer() {
        for (hipEvent_t ev : start) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : stop) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : pool) (void)hipEventDestroy(ev);
    }
};

// Sizes with a register-resident instantiation: 7, 19 or 27 electrons per
// spin, unpolarised with at most 5 G vectors per lane or polarised with at
// most 12.
inline bool rr_supported_sizes(int ndet, int n_up, int n_dn, int ng) {
    const bool n_ok = n_up == 7 || n_up == 19 || n_up == 27;
    if (ndet == 2) return n_ok && n_dn == n_up && ng <= 5 * kWave;
    return n_ok && ng <= 12 * kWave;
}

// Spins of 33, 57 or 81 electrons (equal spins or polarised) have the
// multi-wavefront register-resident instantiations (one wavefront per spin, two
// for 81), and in FP32 also 123 and 147 (two and three wavefronts); structure
// factor over at most 768 G vectors.
inline bool rrw_supported_sizes(int ndet, int n_up, int n_dn, int ng, int real_bytes) {
    const bool n_ok = n_up == 33 || n_up == 57 || n_up == 81 ||
                      (real_bytes == 4 && (n_up == 123 || n_up == 147));
    if (ndet == 2) return n_ok && n_dn == n_up && ng <= 12 * kWave;
    return n_ok && ng <= 12 * kWave;
}

// Newton-Schulz re-inversion on the matrix cores for FP32 inverses of
// determinants up to kNSMaxOrder electrons (vmc_reinv_ns.hip); above 80 it
// needs the ensemble's residual buffer (ns_resid_floats per walker).
constexpr int kNSMaxOrder = 160;
inline int64_t ns_resid_floats(int nmax, int real_bytes) {
    if (real_bytes != 4 || nmax <= 80 || nmax > kNSMaxOrder) return 0;
    const int64_t p = nmax > 128 ? 160 : nmax > 96 ? 128 : 96;
    return p * p;
}
bool reinv_ns_supported(const DevSys &d, int real_bytes);
void reinv_ns_launch(const DevSys &d, const DevState &st, unsigned grid, hipStream_t stream);
void reinv_ns_batch_launch(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                           hipStream_t stream);
// The re-inversion of a batch group of one size (Newton-Schulz where
// supported, else Gauss-Jordan), `grid` = the group's walkers.
void reinv_batch_launch(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                        hipStream_t stream, int real_bytes);

// The FP64 re-inversion after a sweep of the multi-wavefront register and the
// generic kernels: Newton-Schulz where supported, else Gauss-Jordan (defined
// with the generic kernels).
void reinv_generic_launch(const DevSys &d, const DevState &st, unsigned grid, hipStream_t stream,
                          int real_bytes);
struct ReinvLaunch {
    int real_bytes;
    void launch(const DevSys &d, const DevState &st, unsigned grid, hipStream_t s) const {
        reinv_generic_launch(d, st, grid, s, real_bytes);
    }
};

// inv64 (mixed precision): the ensemble's FP64 inverse buffer; measured
// sweeps then re-invert every walker into it and measure in FP64.
bool rrw_launch_f32(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                    unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                    const ReinvLaunch &reinv, double *inv64);
bool rrw_launch_f64(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                    unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                    const ReinvLaunch &reinv, double *inv64);
// The FP64 multi-wavefront local energy from FP64 inverses (st64.inv): into
// the statistics, or with `parts` into parts[w][kPartCount].
bool rrw_measure64(const DevSys &d, const DevState &st64, unsigned grid, int bs, double *parts,
                   hipStream_t stream);
// Mixed-precision parts: FP64 re-inversion of every walker into inv64, then
// rrw_measure64.
bool rrw_parts_exact_f32(const DevSys &d, const DevState &st, unsigned grid, double *parts, double *inv64,
                         hipStream_t stream, const ReinvLaunch &reinv);

// One sweep of a batch group of register-resident systems of one size
// (`shape` = any member's tables, `grid` = the group's walkers): the sweep,
// the re-inversion when `reinv`, and with VMC_SWEEP_MEASURE the local energy,
// the sweep index read from *sweep_ctr.  False when the size has no
// instantiation.
bool rr_batch_step_f32(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                       const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream);
bool rr_batch_step_f64(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                       const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream);
// items_refresh / items_m64 (mixed precision): the group's items with
// st.inv64 set and with st.inv = inv64; a measured sweep then re-inverts
// every walker into inv64 and measures in FP64 (rrw_measure64_batch).
bool rrw_batch_step_f32(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                        const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream,
                        const BatchItem *items_refresh, const BatchItem *items_m64);
bool rrw_batch_step_f64(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                        const unsigned long long *sweep_ctr, int flags, int bs, bool reinv, hipStream_t stream,
                        const BatchItem *items_refresh, const BatchItem *items_m64);
bool rrw_measure64_batch(const DevSys &shape, const BatchItem *items_m64, int n_items, unsigned grid, int bs,
                         hipStream_t stream);

// One sweep of the register kernels with the caller's moves, each recorded
// (then the re-inversion, as after a sweep); and the per-walker parts of the
// local energy by the measurement kernel, statistics untouched.  False when
// the size has no instantiation.
bool rr_given_f32(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
                  hipStream_t stream);
bool rr_given_f64(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
                  hipStream_t stream);
bool rr_parts_f32(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream);
bool rr_parts_f64(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream);
bool rrw_given_f32(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
                   hipStream_t stream, const ReinvLaunch &reinv);
bool rrw_given_f64(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
                   hipStream_t stream, const ReinvLaunch &reinv);
bool rrw_parts_f32(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream);
bool rrw_parts_f64(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream);

// FP64 sweeps with delayed Woodbury updates of HBM-resident inverses
// (vmc_sweep_du.hip): at most 160 electrons per spin, equal spins or
// polarised.  du_launch runs n_sweeps sweeps (no measurement) with the
// Philox draws, or with `given` set one sweep of the caller's moves.
bool du_supported(const DevSys &d);
bool du_launch(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps, unsigned long long s0, int flags,
               const SweepGiven &given, hipStream_t stream);

// Launches n_sweeps sweeps (and, with VMC_SWEEP_MEASURE, one measurement per
// sweep).  variant 0: no occupancy bound; 1: at most 128 VGPRs.
bool rr_launch_f32(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, int variant, hipStream_t stream,
                   KernelTimer &timer);
bool rr_launch_f64(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, int variant, hipStream_t stream,
                   KernelTimer &timer);

}  // namespace vmc
