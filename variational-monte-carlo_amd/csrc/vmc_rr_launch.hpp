// vmc_rr_launch.hpp -- launchers of the register-resident sweep, one
// translation unit per working precision (vmc_rr_f32.hip, vmc_rr_f64.hip).
#pragma once

#include "vmc_device.hpp"

#include <hip/hip_runtime.h>

namespace vmc {

// Sizes with a register-resident instantiation: 7, 19 or 27 electrons per
// spin, unpolarised with at most 5 G vectors per lane or polarised with at
// most 12.
inline bool rr_supported_sizes(int ndet, int n_up, int n_dn, int ng) {
    const bool n_ok = n_up == 7 || n_up == 19 || n_up == 27;
    if (ndet == 2) return n_ok && n_dn == n_up && ng <= 5 * kWave;
    return n_ok && ng <= 12 * kWave;
}

// Launches n_sweeps sweeps (and, with VMC_SWEEP_MEASURE, one measurement per
// sweep).  variant 0: no occupancy bound; 1: at most 128 VGPRs.
bool rr_launch_f32(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, int variant, hipStream_t stream);
bool rr_launch_f64(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                   unsigned long long s0, int flags, int bs, int variant, hipStream_t stream);

}  // namespace vmc
