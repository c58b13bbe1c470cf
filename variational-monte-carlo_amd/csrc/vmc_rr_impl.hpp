// vmc_rr_impl.hpp -- launcher body shared by vmc_rr_f32.hip / vmc_rr_f64.hip.
#pragma once

#include "vmc_rr_launch.hpp"
#include "vmc_sweep_rr.hpp"

#include "../../include/vmc_hip.h"

namespace vmc {

// One launch of the sweep kernel per sweep, each followed by the FP64
// re-inversion of the walkers that reached K accepted moves (when K > 0) and,
// with MEASURE, the local energy.  In mixed precision the local energy is
// the FP64 one of k_measure_rr_mixed, which refines and refreshes every
// walker's inverse itself: a measured sweep takes no separate re-inversion.
template <typename Real, int NU, int ND, int GPL>
void rr_launch_g(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                 unsigned long long s0, int flags, int bs, int variant, hipStream_t stream,
                 KernelTimer &timer) {
    constexpr bool kMixed = sizeof(Real) == 4;
    const bool measure = flags & VMC_SWEEP_MEASURE;
    const bool reinv = d.reinv_k > 0 && !(kMixed && measure);
    // The sweep lists the due walkers; the re-inversion runs on at most
    // kListGrid workgroups (eight per CU) over that list.
    constexpr unsigned kListGrid = 2048;
    const unsigned items = grid * (ND > 0 ? 2 : 1);
    const unsigned rgrid = items < kListGrid ? items : kListGrid;
    if (reinv) flags |= kSweepListDue;
    for (int i = 0; i < n_sweeps; ++i) {
        const unsigned long long si = s0 + (unsigned long long)i;
        hipEvent_t ev = nullptr;
        bool timed = timer.begin(stream, ev);
        if (variant == 1)
            hipLaunchKernelGGL((k_sweep_rr_o4<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0,
                               stream, d, st, si, flags);
        else
            hipLaunchKernelGGL((k_sweep_rr<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream,
                               d, st, si, flags);
        if (timed) timer.end(stream, ev, kTimeSweep);
        if (reinv) {
            timed = timer.begin(stream, ev);
            hipLaunchKernelGGL((k_reinv_rr_list<Real, NU, ND, GPL>), dim3(rgrid), dim3(kWave), 0, stream, d,
                               st);
            if (!measure) hipLaunchKernelGGL(k_due_reset<Real>, dim3(1), dim3(1), 0, stream, st);
            if (timed) timer.end(stream, ev, kTimeReinv);
        }
        if (measure) {
            timed = timer.begin(stream, ev);
            if constexpr (kMixed)
                hipLaunchKernelGGL((k_measure_rr_mixed<NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream, d, st,
                                   bs);
            else
                hipLaunchKernelGGL((k_measure_rr<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0,
                                   stream, d, st, bs);
            if (timed) timer.end(stream, ev, kTimeMeasure);
        }
    }
}

// Calls f.template operator()<NU, ND, GPL>() for the system's instantiation;
// false when there is none.
template <typename Real, typename F>
bool rr_dispatch(const DevSys &d, F &&f) {
    if (!rr_supported_sizes(d.ndet, d.n_up, d.n_dn, d.ng)) return false;
    if (d.ndet == 2) {
        if (d.n_up == 27) f.template operator()<27, 27, 5>();
        else if (d.n_up == 19) f.template operator()<19, 19, 5>();
        else f.template operator()<7, 7, 5>();
    } else {
        if (d.n_up == 27) f.template operator()<27, 0, 12>();
        else if (d.n_up == 19) f.template operator()<19, 0, 12>();
        else f.template operator()<7, 0, 12>();
    }
    return true;
}

template <typename Real>
bool rr_launch(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
               unsigned long long s0, int flags, int bs, int variant, hipStream_t stream,
               KernelTimer &timer) {
    return rr_dispatch<Real>(d, [&]<int NU, int ND, int GPL>() {
        rr_launch_g<Real, NU, ND, GPL>(d, st, grid, n_sweeps, s0, flags, bs, variant, stream, timer);
    });
}

// One sweep with the caller's moves (each recorded), followed like a sweep by
// the re-inversion of the walkers it made due.
template <typename Real>
bool rr_given(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given,
              hipStream_t stream) {
    return rr_dispatch<Real>(d, [&]<int NU, int ND, int GPL>() {
        const bool reinv = d.reinv_k > 0;
        constexpr unsigned kListGrid = 2048;
        const unsigned items = grid * (ND > 0 ? 2 : 1);
        hipLaunchKernelGGL((k_sweep_rr_given<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream, d, st,
                           reinv ? kSweepListDue : 0, given);
        if (reinv) {
            hipLaunchKernelGGL((k_reinv_rr_list<Real, NU, ND, GPL>), dim3(items < kListGrid ? items : kListGrid),
                               dim3(kWave), 0, stream, d, st);
            hipLaunchKernelGGL(k_due_reset<Real>, dim3(1), dim3(1), 0, stream, st);
        }
    });
}

template <typename Real>
bool rr_parts(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream) {
    return rr_dispatch<Real>(d, [&]<int NU, int ND, int GPL>() {
        if constexpr (sizeof(Real) == 4)
            hipLaunchKernelGGL((k_measure_rr_mixed_parts<NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream, d, st,
                               parts);
        else
            hipLaunchKernelGGL((k_measure_rr_parts<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream, d, st,
                               parts);
    });
}

template <typename Real>
bool rr_batch_step(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                   const unsigned long long *ctr, int flags, int bs, bool reinv, hipStream_t stream) {
    return rr_dispatch<Real>(shape, [&]<int NU, int ND, int GPL>() {
        constexpr bool kMixed = sizeof(Real) == 4;
        const bool measure = flags & VMC_SWEEP_MEASURE;
        hipLaunchKernelGGL((k_sweep_rr_b<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream, items,
                           n_items, ctr, flags);
        // (the mixed local energy refreshes every inverse itself)
        if (reinv && !(kMixed && measure))
            hipLaunchKernelGGL((k_reinv_rr_b<Real, NU, ND, GPL>), dim3(grid * (ND > 0 ? 2 : 1)), dim3(kWave), 0,
                               stream, items, n_items);
        if (measure) {
            if constexpr (kMixed)
                hipLaunchKernelGGL((k_measure_rr_mixed_b<NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream,
                                   items, n_items, bs);
            else
                hipLaunchKernelGGL((k_measure_rr_b<Real, NU, ND, GPL>), dim3(grid), dim3(kWave), 0, stream,
                                   items, n_items, bs);
        }
    });
}

}  // namespace vmc
