// vmc_rrw_impl.hpp -- launchers of the multi-wavefront register-resident
// kernels, shared by vmc_rrw_f32.hip / vmc_rrw_f64.hip.
#pragma once

#include "vmc_rr_launch.hpp"
#include "vmc_sweep_rrw.hpp"

namespace vmc {

template <typename Real, int NU, int ND, int WPS>
void rrw_launch_g(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                  unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                  const ReinvLaunch &reinv) {
    constexpr int NT = RRW<Real, NU, ND, WPS>::NT;
    const bool measure = flags & VMC_SWEEP_MEASURE;
    for (int i = 0; i < n_sweeps; ++i) {
        const unsigned long long si = s0 + (unsigned long long)i;
        hipEvent_t ev = nullptr;
        bool timed = timer.begin(stream, ev);
        hipLaunchKernelGGL((k_sweep_rrw<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, d, st, si,
                           flags);
        if (timed) timer.end(stream, ev, kTimeSweep);
        if (d.reinv_k > 0) {
            timed = timer.begin(stream, ev);
            reinv.launch(d, st, grid, stream);
            if (timed) timer.end(stream, ev, kTimeReinv);
        }
        if (measure) {
            timed = timer.begin(stream, ev);
            hipLaunchKernelGGL((k_measure_rrw<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, d,
                               st, bs);
            if (timed) timer.end(stream, ev, kTimeMeasure);
        }
    }
}

template <typename Real>
bool rrw_launch(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                const ReinvLaunch &reinv) {
    if (!rrw_supported_sizes(d.ndet, d.n_up, d.n_dn, d.ng, (int)sizeof(Real))) return false;
#define VMC_RRW(NU_, ND_, W_) \
    rrw_launch_g<Real, NU_, ND_, W_>(d, st, grid, n_sweeps, s0, flags, bs, stream, timer, reinv)
    if (d.ndet == 2) {
        if (d.n_up == 33) VMC_RRW(33, 33, 1);
        else if (d.n_up == 57) VMC_RRW(57, 57, 1);
        else if (d.n_up == 81) VMC_RRW(81, 81, 2);
        else if constexpr (sizeof(Real) == 4) {
            if (d.n_up == 123) VMC_RRW(123, 123, 2);
            else VMC_RRW(147, 147, 3);
        }
    } else {
        if (d.n_up == 33) VMC_RRW(33, 0, 1);
        else if (d.n_up == 57) VMC_RRW(57, 0, 1);
        else if (d.n_up == 81) VMC_RRW(81, 0, 2);
        else if constexpr (sizeof(Real) == 4) {
            if (d.n_up == 123) VMC_RRW(123, 0, 2);
            else VMC_RRW(147, 0, 3);
        }
    }
#undef VMC_RRW
    return true;
}

}  // namespace vmc
