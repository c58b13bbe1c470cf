// vmc_rr_impl.hpp -- launcher body shared by vmc_rr_f32.hip / vmc_rr_f64.hip.
#pragma once

#include "vmc_rr_launch.hpp"
#include "vmc_sweep_rr.hpp"

#include "../../include/vmc_hip.h"

namespace vmc {

// Sweeps in launches of `per_launch`, each followed by the post-sweep kernel
// when it has work: the local energy after every sweep (MEASURE) and the
// FP64 re-inversion of walkers that reached K accepted moves.  Without
// measurement a launch holds at most K / N sweeps, so no walker runs more
// than K accepted moves past its due re-inversion.
template <typename Real, int NU, int ND, int GPL>
void rr_launch_g(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                 unsigned long long s0, int flags, int bs, int variant, hipStream_t stream) {
    const bool measure = flags & VMC_SWEEP_MEASURE;
    const bool reinv = d.reinv_k > 0;
    int per_launch = n_sweeps;
    if (measure) per_launch = 1;
    else if (reinv) per_launch = d.reinv_k / d.N > 1 ? d.reinv_k / d.N : 1;
    const int post_flags = (measure ? kPostMeasure : 0) | (reinv ? kPostReinvert : 0);
    for (int i = 0; i < n_sweeps; i += per_launch) {
        const int n = n_sweeps - i < per_launch ? n_sweeps - i : per_launch;
        const unsigned long long si = s0 + (unsigned long long)i;
        if (variant == 1)
            hipLaunchKernelGGL((k_sweep_rr_o4<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0,
                               stream, d, st, n, si, flags);
        else
            hipLaunchKernelGGL((k_sweep_rr<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream,
                               d, st, n, si, flags);
        if (post_flags)
            hipLaunchKernelGGL((k_post_rr<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream,
                               d, st, bs, post_flags);
    }
}

template <typename Real>
bool rr_launch(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
               unsigned long long s0, int flags, int bs, int variant, hipStream_t stream) {
    if (!rr_supported_sizes(d.ndet, d.n_up, d.n_dn, d.ng)) return false;
    if (d.ndet == 2) {
        if (d.n_up == 27) rr_launch_g<Real, 27, 27, 5>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream);
        else if (d.n_up == 19) rr_launch_g<Real, 19, 19, 5>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream);
        else rr_launch_g<Real, 7, 7, 5>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream);
    } else {
        if (d.n_up == 27) rr_launch_g<Real, 27, 0, 12>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream);
        else if (d.n_up == 19) rr_launch_g<Real, 19, 0, 12>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream);
        else rr_launch_g<Real, 7, 0, 12>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream);
    }
    return true;
}

}  // namespace vmc
