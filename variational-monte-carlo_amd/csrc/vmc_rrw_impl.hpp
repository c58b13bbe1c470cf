// vmc_rrw_impl.hpp -- launchers of the multi-wavefront register-resident
// kernels, shared by vmc_rrw_f32.hip / vmc_rrw_f64.hip.
#pragma once

#include "vmc_rr_launch.hpp"
#include "vmc_sweep_rrw.hpp"

namespace vmc {

// Sweeps, each followed by the re-inversion of the walkers due for it and,
// with MEASURE, the local energy.  In mixed precision (inv64 set) a
// measured sweep re-inverts every walker with an FP64 result in inv64 and
// takes the FP64 local energy from it.
template <typename Real, int NU, int ND, int WPS>
void rrw_launch_g(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps,
                  unsigned long long s0, int flags, int bs, hipStream_t stream, KernelTimer &timer,
                  const ReinvLaunch &reinv, double *inv64) {
    constexpr int NT = RRW<Real, NU, ND, WPS>::NT;
    const bool measure = flags & VMC_SWEEP_MEASURE;
    const bool exact = measure && inv64;
    DevState st_refresh = st, st64 = st;
    st_refresh.inv64 = inv64;
    st64.inv = inv64;
    for (int i = 0; i < n_sweeps; ++i) {
        const unsigned long long si = s0 + (unsigned long long)i;
        hipEvent_t ev = nullptr;
        bool timed = timer.begin(stream, ev);
        hipLaunchKernelGGL((k_sweep_rrw<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, d, st, si,
                           flags);
        if (timed) timer.end(stream, ev, kTimeSweep);
        if (d.reinv_k > 0 || exact) {
            timed = timer.begin(stream, ev);
            reinv.launch(d, exact ? st_refresh : st, grid, stream);
            if (timed) timer.end(stream, ev, kTimeReinv);
        }
        if (measure) {
            timed = timer.begin(stream, ev);
            if (exact)
                rrw_measure64(d, st64, grid, bs, nullptr, stream);
            else
                hipLaunchKernelGGL((k_measure_rrw<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, d,
                                   st, bs);
            if (timed) timer.end(stream, ev, kTimeMeasure);
        }
    }
}

// The FP64 local energy of a mixed-precision measured sweep, from the FP64
// inverses its re-inversion left in st64.inv (every multi-wavefront size,
// 123 and 147 per spin included: their rows are read from memory).
template <typename F>
bool rrw_measure64_dispatch(const DevSys &d, F &&f) {
    const int n = d.n_up;
    if (!((n == 33 || n == 57 || n == 81 || n == 123 || n == 147) && (d.ndet == 1 || d.n_dn == n) &&
          d.ng <= 12 * kWave))
        return false;
    if (d.ndet == 2) {
        if (n == 33) f.template operator()<33, 33, 1>();
        else if (n == 57) f.template operator()<57, 57, 1>();
        else if (n == 81) f.template operator()<81, 81, 2>();
        else if (n == 123) f.template operator()<123, 123, 2>();
        else f.template operator()<147, 147, 3>();
    } else {
        if (n == 33) f.template operator()<33, 0, 1>();
        else if (n == 57) f.template operator()<57, 0, 1>();
        else if (n == 81) f.template operator()<81, 0, 2>();
        else if (n == 123) f.template operator()<123, 0, 2>();
        else f.template operator()<147, 0, 3>();
    }
    return true;
}

inline bool rrw_measure64_impl(const DevSys &d, const DevState &st64, unsigned grid, int bs, double *parts,
                               hipStream_t stream) {
    return rrw_measure64_dispatch(d, [&]<int NU, int ND, int WPS>() {
        constexpr int NT = RRW<double, NU, ND, WPS>::NT;
        if (parts)
            hipLaunchKernelGGL((k_measure_rrw_parts<double, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, d,
                               st64, parts);
        else
            hipLaunchKernelGGL((k_measure_rrw<double, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, d, st64,
                               bs);
    });
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
                const ReinvLaunch &reinv, double *inv64) {
    return rrw_dispatch<Real>(d, [&]<int NU, int ND, int WPS>() {
        rrw_launch_g<Real, NU, ND, WPS>(d, st, grid, n_sweeps, s0, flags, bs, stream, timer, reinv, inv64);
    });
}

// One sweep with the caller's moves (each recorded), then the re-inversion
// as after a sweep.
template <typename Real>
bool rrw_given(const DevSys &d, const DevState &st, unsigned grid, const SweepGiven &given, hipStream_t stream,
               const ReinvLaunch &reinv) {
    return rrw_dispatch<Real>(d, [&]<int NU, int ND, int WPS>() {
        constexpr int NT = RRW<Real, NU, ND, WPS>::NT;
        hipLaunchKernelGGL((k_sweep_rrw_given<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, d, st, 0,
                           given);
        if (d.reinv_k > 0) reinv.launch(d, st, grid, stream);
    });
}

template <typename Real>
bool rrw_parts(const DevSys &d, const DevState &st, unsigned grid, double *parts, hipStream_t stream) {
    return rrw_dispatch<Real>(d, [&]<int NU, int ND, int WPS>() {
        constexpr int NT = RRW<Real, NU, ND, WPS>::NT;
        hipLaunchKernelGGL((k_measure_rrw_parts<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, d, st,
                           parts);
    });
}

// The mixed-precision parts: every walker re-inverted with its FP64 result
// in inv64 (the walkers' state, the stored FP32 inverse among it, left
// alone), then the FP64 local energy's parts from it.
template <typename Real>
bool rrw_parts_exact(const DevSys &d, const DevState &st, unsigned grid, double *parts, double *inv64,
                     hipStream_t stream, const ReinvLaunch &reinv) {
    if (!rrw_supported_sizes(d.ndet, d.n_up, d.n_dn, d.ng, (int)sizeof(Real))) return false;
    DevState st_refresh = st, st64 = st;
    st_refresh.inv64 = inv64;
    st_refresh.inv64_only = 1;
    st64.inv = inv64;
    reinv.launch(d, st_refresh, grid, stream);
    return rrw_measure64(d, st64, grid, 1, parts, stream);
}

template <typename Real>
bool rrw_batch_step(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                    const unsigned long long *ctr, int flags, int bs, bool reinv, hipStream_t stream,
                    const BatchItem *items_refresh, const BatchItem *items_m64) {
    return rrw_dispatch<Real>(shape, [&]<int NU, int ND, int WPS>() {
        constexpr int NT = RRW<Real, NU, ND, WPS>::NT;
        const bool measure = flags & VMC_SWEEP_MEASURE;
        hipLaunchKernelGGL((k_sweep_rrw_b<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, items,
                           n_items, ctr, flags);
        if (measure && items_m64) {
            reinv_batch_launch(shape, items_refresh, n_items, grid, stream, (int)sizeof(Real));
            rrw_measure64_batch(shape, items_m64, n_items, grid, bs, stream);
            return;
        }
        if (reinv) reinv_batch_launch(shape, items, n_items, grid, stream, (int)sizeof(Real));
        if (measure)
            hipLaunchKernelGGL((k_measure_rrw_b<Real, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, items,
                               n_items, bs);
    });
}

inline bool rrw_measure64_batch_impl(const DevSys &shape, const BatchItem *items_m64, int n_items, unsigned grid,
                                     int bs, hipStream_t stream) {
    return rrw_measure64_dispatch(shape, [&]<int NU, int ND, int WPS>() {
        constexpr int NT = RRW<double, NU, ND, WPS>::NT;
        hipLaunchKernelGGL((k_measure_rrw_b<double, NU, ND, WPS>), dim3(grid), dim3(NT), 0, stream, items_m64,
                           n_items, bs);
    });
}

}  // namespace vmc
