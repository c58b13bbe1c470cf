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

// Calls f.template operator()<NU, ND, WPS>() for the system's instantiation;
// false when there is none.
template <typename Real, typename F>
bool rrw_dispatch(const DevSys &d, F &&f) {
    if (!rrw_supported_sizes(d.ndet, d.n_up, d.n_dn, d.ng, (int)sizeof(Real))) return false;
    if (d.ndet == 2) {
        if (d.n_up == 33) f.template operator()<33, 33, 1>();
        else if (d.n_up == 57) f.template operator()<57, 57, 1>();
        else if (d.n_up == 81) f.template operator()<81, 81, 2>();
        else if constexpr (sizeof(Real) == 4) {
            if (d.n_up == 123) f.template operator()<123, 123, 2>();
            else f.template operator()<147, 147, 3>();
        }
    } else {
        if (d.n_up == 33) f.template operator()<33, 0, 1>();
        else if (d.n_up == 57) f.template operator()<57, 0, 1>();
        else if (d.n_up == 81) f.template operator()<81, 0, 2>();
        else if constexpr (sizeof(Real) == 4) {
            if (d.n_up == 123) f.template operator()<123, 0, 2>();
            else f.template operator()<147, 0, 3>();
        }
    }
    return true;
}

template <typename Real>
bool rrw_launch(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                const ReinvLaunch &reinv) {
    return rrw_dispatch<Real>(d, [&]<int NU, int ND, int WPS>() {
        rrw_launch_g<Real, NU, ND, WPS>(d, st, grid, n_sweeps, s0, flags, bs, stream, timer, reinv);
    });
}

template <typename Real>
bool rrw_batch_step(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                    const unsigned long long *ctr, int flags, int bs, bool reinv, hipStream_t stream) {
    return rrw_dispatch<Real>(shape, [&]<int NU, int ND, int WPS>() {
        constexpr int NT = RRW<Real, NU, ND, WPS>::NT;
        hipLaunchKernelGGL((k_sweep_rrw_b<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, items,
                           n_items, ctr, flags);
        if (reinv) reinv_batch_launch(shape, items, n_items, grid, stream, (int)sizeof(Real));
        if (flags & VMC_SWEEP_MEASURE)
            hipLaunchKernelGGL((k_measure_rrw_b<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, items,
                               n_items, bs);
    });
}

}  // namespace vmc
