// vmc_hip.hip -- gfx950 kernels and the C ABI (include/vmc_hip.h) of the VMC
// sweep for the 3D homogeneous electron gas.
//
// Layout in HBM, per ensemble (walker-major SoA):
//   pos     double [W][3][N]   x block, y block, z block
//   inv     Real   [W][ninv]   transposed inverses of the determinants
//   sigma   double [W]         proposal width
//   stats   double [W][16]     FP64 accumulators (VMC_STAT_*)
//   logdet  double [W][2]
//   since   int    [W]         accepted moves since the last FP64 re-inversion
//   flags   int    [W]         bit0: singular determinant met; bit1: jrow current
//   scratch double [W][nmax^2] FP64 Gauss-Jordan workspace (and Newton-Schulz E
//                              above 80 electrons per spin)
//   drift   double [W]         largest |I - D^-1 D| met at a re-inversion
//   jrow    double [W][N]      Jastrow row sums kept by the register sweeps
//   due_list int   [W]         walkers a register sweep found due for
//                              re-inversion (count in due[0])
//   resid   float  [W][P^2]    Newton-Schulz residual, mixed above 80 per spin
// The generic kernel in this file runs one walker per 64-, 128- or 256-thread
// workgroup; its positions, Jastrow row sums and draws stay in LDS for a whole
// launch, and so do the inverses when they fit.  The register-resident
// kernels (vmc_sweep_rr.hpp, vmc_sweep_rrw.hpp) and the Newton-Schulz
// re-inversion (vmc_reinv_ns.hip) live in their own translation units.
#include "../../include/vmc_hip.h"
#include "vmc_device.hpp"
#include "vmc_tables.hpp"
#include "vmc_walker.hpp"
#include "vmc_rr_launch.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numbers>
#include <string>
#include <vector>

namespace vmc {

constexpr size_t kMaxLdsBytes = 64 * 1024;

// ----------------------------------------------------------------- kernels

// Start configurations: Philox draws (tag INIT, counter {i, walker, attempt})
// retried while a determinant is singular, mirroring
// Simulation::initialize_positions (simulation.cpp:43-76).  With draw == 0
// the current positions are kept and only the FP64 inverses are rebuilt.
template <typename Real>
__global__ __launch_bounds__(kWave) void k_init(DevSys s, DevState st, int draw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = blockIdx.x;
    const int lane = threadIdx.x;
    const int N = s.N;
    Real *inv_hbm = reinterpret_cast<Real *>(st.inv) + (size_t)w * s.ninv;
    double *A = st.scratch + (size_t)w * s.nmax * s.nmax;
    WalkerView<Real> v = carve<Real>(smem, s, false, inv_hbm, A);
    double *gp = st.pos + (size_t)w * 3 * N;
    const int attempts = draw ? 100 : 1;
    int ok = 0;
    double ld[2] = {0.0, 0.0};
    for (int att = 0; att < attempts && !ok; ++att) {
        for (int i = lane; i < N; i += kWave) {
            if (draw) {
                const uint4 r = philox4x32_10(make_uint4((uint32_t)i, (uint32_t)w, (uint32_t)att,
                                                         kTagInit),
                                              s.key0, s.key1);
                v.px[i] = u01(r.x) * s.L;
                v.py[i] = u01(r.y) * s.L;
                v.pz[i] = u01(r.z) * s.L;
            } else {
                v.px[i] = gp[i];
                v.py[i] = gp[N + i];
                v.pz[i] = gp[2 * N + i];
            }
        }
        wave_sync();
        ok = 1;
        for (int d = 0; d < s.ndet; ++d) {
            ld[d] = reinvert<Real>(s, d, v, lane);
            if (!(ld[d] > -INFINITY)) ok = 0;
        }
    }
    for (int i = lane; i < N; i += kWave) {
        gp[i] = v.px[i];
        gp[N + i] = v.py[i];
        gp[2 * N + i] = v.pz[i];
    }
    if (lane == 0) {
        st.logdet[2 * w + 0] = ld[0];
        st.logdet[2 * w + 1] = ld[1];
        st.flags[w] = ok ? 0 : 1;
        st.since[w] = 0;
    }
}

// n_sweeps sequential sweeps of walker blockIdx.x (simulation.cpp:80-263 as
// an ensemble): per electron p, in the sweep order of sweep_visit, a Philox
// proposal (tag MOVE, counter {p, walker, sweep}), ratio and Jastrow change, Metropolis test with a
// Philox uniform (tag ACCP), Sherman-Morrison on accept; after each sweep the
// optional step adaptation and local-energy measurement.
// With given.prop set, one sweep takes the caller's moves (absolute
// proposals and accept uniforms) and records each move in given.trace.
template <typename Real>
__device__ void sweep_walker(const DevSys &s, const DevState &st, int w, bool kInvLds,
                             int n_sweeps, unsigned long long sweep0, int flags, int block_size,
                             unsigned char *smem, SweepGiven given = SweepGiven{},
                             StepTrace trace = StepTrace{-1, nullptr, nullptr, nullptr, nullptr}) {
    const int lane = threadIdx.x;
    const int N = s.N;
    Real *inv_hbm = reinterpret_cast<Real *>(st.inv) + (size_t)w * s.ninv;
    double *A = st.scratch + (size_t)w * s.nmax * s.nmax;
    WalkerView<Real> v = carve<Real>(smem, s, kInvLds, inv_hbm, A);
    if (kInvLds)
        for (int e = lane; e < s.ninv; e += (int)blockDim.x) v.inv[e] = inv_hbm[e];
    double *gp = st.pos + (size_t)w * 3 * N;
    for (int i = lane; i < N; i += (int)blockDim.x) {
        v.px[i] = gp[i];
        v.py[i] = gp[N + i];
        v.pz[i] = gp[2 * N + i];
    }
    wave_sync();
    jastrow_rows<Real>(s, v, lane);

    double *ws = st.stats + (size_t)w * kStatCount;
    double stat[kStatCount];
#pragma unroll
    for (int k = 0; k < kStatCount; ++k) stat[k] = ws[k];
    double sigma = st.sigma[w];
    int since = st.since[w];
    if (s.reinv_k > 0 && since >= s.reinv_k) since = 0;  // re-inverted after the last sweep
    int wflags = st.flags[w];
    double ld0 = st.logdet[2 * w + 0], ld1 = st.logdet[2 * w + 1];
    double drift = st.drift[w];

    for (int sw = 0; sw < n_sweeps; ++sw) {
        const uint32_t sweep = (uint32_t)(sweep0 + (unsigned long long)sw);
        for (int q = lane; q < N && given.prop; q += (int)blockDim.x) {
            const size_t g = (size_t)w * N + q;
            v.draw[q] = given.prop[3 * g];
            v.draw[N + q] = given.prop[3 * g + 1];
            v.draw[2 * N + q] = given.prop[3 * g + 2];
            v.draw[3 * N + q] = given.u[g];
        }
        for (int q = lane; q < N && !given.prop; q += (int)blockDim.x) {
            const uint4 r = philox4x32_10(make_uint4((uint32_t)q, (uint32_t)w, sweep, kTagMove),
                                          s.key0, s.key1);
            const uint4 ra = philox4x32_10(make_uint4((uint32_t)q, (uint32_t)w, sweep, kTagAccept),
                                           s.key0, s.key1);
            double dx, dy, dz;
            if (s.move_kind == 1) {
                const double rho0 = sqrt(-2.0 * log(u01(r.x)));
                const double rho1 = sqrt(-2.0 * log(u01(r.z)));
                double s0, c0, s1, c1;
                sincospi(2.0 * u01(r.y), &s0, &c0);
                sincospi(2.0 * u01(r.w), &s1, &c1);
                dx = rho0 * c0;
                dy = rho0 * s0;
                dz = rho1 * c1;
            } else {
                dx = 2.0 * u01(r.x) - 1.0;
                dy = 2.0 * u01(r.y) - 1.0;
                dz = 2.0 * u01(r.z) - 1.0;
            }
            v.draw[q] = dx;
            v.draw[N + q] = dy;
            v.draw[2 * N + q] = dz;
            v.draw[3 * N + q] = u01(ra.x);
        }
        wave_sync();
        int acc = 0;
        for (int t = 0; t < N; ++t) {
            const int p = sweep_visit(s.n_up, s.n_dn, t);
            double x, y, z;
            if (given.prop) {
                x = v.draw[p];
                y = v.draw[N + p];
                z = v.draw[2 * N + p];
            } else {
                x = v.px[p] + sigma * v.draw[p];
                y = v.py[p] + sigma * v.draw[N + p];
                z = v.pz[p] + sigma * v.draw[2 * N + p];
                x -= s.L * floor(x * s.inv_L);
                y -= s.L * floor(y * s.inv_L);
                z -= s.L * floor(z * s.inv_L);
            }
            const Proposal pr = propose<Real>(s, v, p, x, y, z, lane);
            const double prob = pr.ratio * pr.ratio * exp(2.0 * pr.dj);
            const int accept = wave_uniform((int)(v.draw[3 * N + p] < prob));
            if (given.trace && threadIdx.x == 0) {
                double *q = given.trace + ((size_t)w * N + p) * kTraceCount;
                q[0] = pr.ratio;
                q[1] = pr.dj;
                q[2] = (double)accept;
            }
            if (accept) {
                accept_move<Real>(s, v, p, x, y, z, pr, lane);
                ++acc;
                const int d = spin_of(s, p);
                if (d == 0) ld0 += log(fabs(pr.ratio)); else ld1 += log(fabs(pr.ratio));
                ++since;  // the FP64 re-inversion follows the sweep (k_reinv_generic)
            }
            if ((long long)w == trace.walker) {
                // E_L after every step (simulation.cpp:212-221); measure reads
                // the positions and the inverse only, so the chain goes on as
                // it would untraced.
                wave_sync();
                const Measured m = measure<Real>(s, v, lane, nullptr, nullptr);
                if (lane == 0) {
                    trace.e_local[t] = m.e_local;
                    trace.moved[t] = p;
                    trace.accepted[t] = accept;
                }
                for (int i = lane; i < N && trace.pos; i += (int)blockDim.x) {
                    double *q = trace.pos + ((size_t)t * N + i) * 3;
                    q[0] = v.px[i];
                    q[1] = v.py[i];
                    q[2] = v.pz[i];
                }
                wave_sync();
            }
        }
        stat[6] += (double)acc;
        stat[7] += (double)N;
        if (flags & VMC_SWEEP_ADAPT) {
            sigma *= exp(0.25 * ((double)acc / (double)N - 0.5));
            if (sigma > 0.5 * s.L) sigma = 0.5 * s.L;
        }
        if (flags & VMC_SWEEP_MEASURE) {
            jastrow_rows<Real>(s, v, lane);
            const Measured m = measure<Real>(s, v, lane, nullptr, nullptr);
            accumulate_measure(stat, m, s.e_const, block_size);
        }
    }
    if (kInvLds)
        for (int e = lane; e < s.ninv; e += (int)blockDim.x) inv_hbm[e] = v.inv[e];
    for (int i = lane; i < N; i += (int)blockDim.x) {
        gp[i] = v.px[i];
        gp[N + i] = v.py[i];
        gp[2 * N + i] = v.pz[i];
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kStatCount; ++k) ws[k] = stat[k];
        st.sigma[w] = sigma;
        st.since[w] = since;
        st.flags[w] = wflags & ~kFlagJrowValid;  // this kernel keeps no row sums
        st.logdet[2 * w + 0] = ld0;
        st.logdet[2 * w + 1] = ld1;
        st.drift[w] = drift;
    }
    wave_sync();
}

// One walker per workgroup of 64, 128 or 256 threads (1, 2 or 4 wavefronts).
template <typename Real, bool kInvLds>
__global__ __launch_bounds__(256) void k_sweep(DevSys s, DevState st, int n_sweeps,
                                                 unsigned long long sweep0, int flags,
                                                 int block_size) {
    extern __shared__ __align__(16) unsigned char smem[];
    sweep_walker<Real>(s, st, blockIdx.x, kInvLds, n_sweeps, sweep0, flags, block_size, smem);
}

// One sweep with walker trace.walker's local energy after each of its moves.
template <bool kInvLds>
__global__ __launch_bounds__(256) void k_sweep_traced(DevSys s, DevState st, unsigned long long sweep0, int flags,
                                                      int block_size, StepTrace trace) {
    extern __shared__ __align__(16) unsigned char smem[];
    sweep_walker<double>(s, st, blockIdx.x, kInvLds, 1, sweep0, flags, block_size, smem, SweepGiven{}, trace);
}

// One sweep with the caller's moves, each recorded.
template <typename Real, bool kInvLds>
__global__ __launch_bounds__(256) void k_sweep_given(DevSys s, DevState st, SweepGiven given) {
    extern __shared__ __align__(16) unsigned char smem[];
    sweep_walker<Real>(s, st, blockIdx.x, kInvLds, 1, 0, 0, 1, smem, given);
}

// The local energy of every walker by the generic routines (the measurement
// of k_sweep): its parts into parts[w][kPartCount] (statistics untouched),
// or with parts null into the statistics.  The FP64 local energy of a mixed
// measured sweep runs this in FP64 with st.inv the sweep's FP64 inverses.
template <typename Real>
__device__ void measure_walker(const DevSys &s, const DevState &st, int w, bool kInvLds, double *parts,
                               int block_size, unsigned char *smem) {
    const int lane = threadIdx.x;
    const int N = s.N;
    Real *inv_hbm = reinterpret_cast<Real *>(st.inv) + (size_t)w * s.ninv;
    double *A = st.scratch + (size_t)w * s.nmax * s.nmax;
    WalkerView<Real> v = carve<Real>(smem, s, kInvLds, inv_hbm, A);
    if (kInvLds)
        for (int e = lane; e < s.ninv; e += (int)blockDim.x) v.inv[e] = inv_hbm[e];
    const double *gp = st.pos + (size_t)w * 3 * N;
    for (int i = lane; i < N; i += (int)blockDim.x) {
        v.px[i] = gp[i];
        v.py[i] = gp[N + i];
        v.pz[i] = gp[2 * N + i];
    }
    wave_sync();
    const Measured m = measure<Real>(s, v, lane, nullptr, nullptr);
    if (!parts) {
        if (lane == 0) accumulate_measure(st.stats + (size_t)w * kStatCount, m, s.e_const, block_size);
        return;
    }
    jastrow_rows<Real>(s, v, lane);
    double u = 0.0;
    for (int i = lane; i < N; i += (int)blockDim.x) u += v.U[i];
    const double J = 0.5 * block_sum(u);
    if (lane == 0) {
        double *q = parts + (size_t)w * kPartCount;
        q[0] = m.kinetic;
        q[1] = m.vreal;
        q[2] = m.vrecip;
        q[3] = m.e_local;
        q[4] = st.logdet[2 * w + 0];
        q[5] = s.ndet == 2 ? st.logdet[2 * w + 1] : 0.0;
        q[6] = J;
    }
}

template <typename Real, bool kInvLds>
__global__ __launch_bounds__(256) void k_measure_parts(DevSys s, DevState st, double *parts, int block_size) {
    extern __shared__ __align__(16) unsigned char smem[];
    measure_walker<Real>(s, st, blockIdx.x, kInvLds, parts, block_size, smem);
}

// The FP64 local energy of a mixed batch's generic group after its measured
// sweep: one workgroup per walker of the group, st.inv the FP64 inverses.
__global__ __launch_bounds__(256) void k_measure_batch64(const BatchItem *items, int n_items, int block_size) {
    extern __shared__ __align__(16) unsigned char smem[];
    const BatchItem &it = items[batch_find(items, n_items, blockIdx.x)];
    measure_walker<double>(it.s, it.st, blockIdx.x - it.walker_begin, false, nullptr, block_size, smem);
}

// One sweep of every walker of every system of a batch in one persistent
// launch: each workgroup takes global walker indices from a device queue
// until the batch is exhausted, so systems of uneven size and cost share the
// machine.  The sweep index (the Philox counter) and the queue live in
// device memory and k_batch_advance steps them, so a captured graph of
// sweeps replays with fresh random numbers.  Walkers use their
// system-local index and their system's seed: a system's chain is the same
// alone and in a batch.
template <typename Real>
__global__ __launch_bounds__(256) void k_sweep_batch(const BatchItem *items, int n_items,
                                                       int total, unsigned int *queue,
                                                       const unsigned long long *sweep_ctr,
                                                       int flags, int block_size) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int next;
    const unsigned long long sweep0 = *sweep_ctr;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int g = next;
        __syncthreads();
        if (g >= total) break;
        const BatchItem &it = items[batch_find(items, n_items, g)];
        sweep_walker<Real>(it.s, it.st, g - it.walker_begin, it.inv_lds != 0, 1, sweep0, flags,
                           block_size, smem);
    }
}

// After a batch sweep: next sweep index, queue back to zero.
__global__ void k_batch_advance(unsigned int *queue, unsigned long long *sweep_ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (queue) *queue = 0u;
        *sweep_ctr += 1ull;
    }
}

// ------------------------------------------------ FP64 re-inversion (generic)
// Walkers whose accepted moves since their last re-inversion reached K
// rebuild each determinant's inverse in FP64 after the sweep, one 256-thread
// workgroup per walker.  E = D^T (E[m][i] = phi_m(r_i)) is built in FP64 in
// LDS (in the walker's HBM scratch when n^2 doubles do not fit); the drift
// max|I - E invT| of the inverse being replaced is measured first, four
// columns of invT at a time; Gauss-Jordan with partial pivoting then inverts
// E in place (the reference's LU tolerance 1e-12, matrix.hpp:35-39) and the
// result replaces invT in the working precision.
constexpr int kReinvThreads = 256;
constexpr int kReinvCols = 4;

__host__ __device__ inline size_t reinv_lds_bytes(int nmax, bool a_in_lds) {
    return (a_in_lds ? (size_t)nmax * nmax * 8 : 0) + (size_t)nmax * 8 /* gjcol */ +
           (size_t)nmax * 4 /* perm */ + (size_t)kReinvCols * nmax * 8 /* columns */ + 64 * 8;
}

template <typename Real>
__device__ void reinv_walker(const DevSys &s, const DevState &st, int w, bool a_in_lds,
                             unsigned char *smem) {
    // st.inv64 set: every walker, with the FP64 result kept there.
    if (!st.inv64 && (s.reinv_k <= 0 || st.since[w] < s.reinv_k)) return;
    const int t = threadIdx.x, nt = blockDim.x;
    size_t off = 0;
    double *A = a_in_lds ? reinterpret_cast<double *>(smem) : st.scratch + (size_t)w * s.nmax * s.nmax;
    if (a_in_lds) off += (size_t)s.nmax * s.nmax * 8;
    double *gjcol = reinterpret_cast<double *>(smem + off); off += (size_t)s.nmax * 8;
    double *cols = reinterpret_cast<double *>(smem + off); off += (size_t)kReinvCols * s.nmax * 8;
    double *red = reinterpret_cast<double *>(smem + off); off += 64 * 8;
    int *perm = reinterpret_cast<int *>(smem + off);
    const double *gp = st.pos + (size_t)w * 3 * s.N;
    Real *inv = reinterpret_cast<Real *>(st.inv) + (size_t)w * s.ninv;
    double drift = 0.0, ld[2] = {st.logdet[2 * w + 0], st.logdet[2 * w + 1]};
    bool singular = false;
    for (int d = 0; d < s.ndet; ++d) {
        const int n = s.n[d], first = s.first[d];
        Real *Y = inv + s.inv_off[d];
        for (int e = t; e < n * n; e += nt) {
            const int j = e / n, i = e - (e / n) * n;
            const int4 o = s.orb[d * s.nmax + j];
            double sn, cs;
            sincos_2pi<double>(frac_phase(o.x, o.y, o.z, gp[first + i] * s.inv_L,
                                          gp[s.N + first + i] * s.inv_L,
                                          gp[2 * s.N + first + i] * s.inv_L), &sn, &cs);
            A[e] = o.w ? sn : cs;
        }
        __syncthreads();
        // Drift of the inverse being replaced, kReinvCols columns at a time.
        double dmax = 0.0;
        for (int c0 = 0; c0 < n; c0 += kReinvCols) {
            const int nc = n - c0 < kReinvCols ? n - c0 : kReinvCols;
            for (int e = t; e < nc * n; e += nt) {
                const int c = e / n, i = e - (e / n) * n;
                cols[c * n + i] = (double)Y[i * n + c0 + c];
            }
            __syncthreads();
            for (int e = t; e < nc * n; e += nt) {
                const int c = e / n, m = e - (e / n) * n;
                double z = 0.0;
                for (int i = 0; i < n; ++i) z += A[m * n + i] * cols[c * n + i];
                const double r = fabs((m == c0 + c ? 1.0 : 0.0) - z);
                dmax = (r > dmax || r != r) ? r : dmax;
            }
            __syncthreads();
        }
        dmax = block_max(dmax, red);
        drift = (dmax > drift || dmax != dmax) ? dmax : drift;
        // Gauss-Jordan in place: A becomes (D^T)^-1 = invT.
        const double logdet = gj_invert(A, n, gjcol, perm, red);
        if (!(logdet > -INFINITY)) {
            singular = true;
            break;
        }
        for (int e = t; e < n * n && !st.inv64_only; e += nt) Y[e] = (Real)A[e];
        if (st.inv64)
            for (int e = t; e < n * n; e += nt) st.inv64[(size_t)w * s.ninv + s.inv_off[d] + e] = A[e];
        ld[d] = logdet;
        __syncthreads();
    }
    if (t == 0 && !st.inv64_only) {
        st.since[w] = 0;
        st.stats[(size_t)w * kStatCount + 15] += 1.0;
        if (singular) {
            st.flags[w] |= 1;
        } else {
            st.logdet[2 * w + 0] = ld[0];
            st.logdet[2 * w + 1] = ld[1];
        }
        if (drift > st.drift[w] || drift != drift) st.drift[w] = drift;
    }
    __syncthreads();
}

template <typename Real>
__global__ __launch_bounds__(kReinvThreads) void k_reinv_generic(DevSys s, DevState st, int a_in_lds) {
    extern __shared__ __align__(16) unsigned char smem[];
    reinv_walker<Real>(s, st, blockIdx.x, a_in_lds != 0, smem);
}

// The batch's re-inversion: one workgroup per global walker, each finding
// its system; walkers not due leave at once.
template <typename Real>
__global__ __launch_bounds__(kReinvThreads) void k_reinv_batch(const BatchItem *items, int n_items,
                                                               size_t lds_budget) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int g = blockIdx.x;
    const BatchItem &it = items[batch_find(items, n_items, g)];
    if (it.s.reinv_k <= 0) return;
    const bool a_in_lds = reinv_lds_bytes(it.s.nmax, true) <= lds_budget;
    reinv_walker<Real>(it.s, it.st, g - it.walker_begin, a_in_lds, smem);
}

void reinv_generic_launch(const DevSys &d, const DevState &st, unsigned grid, hipStream_t stream,
                          int real_bytes) {
    if (reinv_ns_supported(d, real_bytes)) {
        reinv_ns_launch(d, st, grid, stream);
        return;
    }
    const bool a_lds = reinv_lds_bytes(d.nmax, true) <= kMaxLdsBytes;
    const size_t lds = reinv_lds_bytes(d.nmax, a_lds);
    if (real_bytes == 8)
        hipLaunchKernelGGL((k_reinv_generic<double>), dim3(grid), dim3(kReinvThreads), lds, stream, d, st,
                           a_lds ? 1 : 0);
    else
        hipLaunchKernelGGL((k_reinv_generic<float>), dim3(grid), dim3(kReinvThreads), lds, stream, d, st,
                           a_lds ? 1 : 0);
}

void reinv_batch_launch(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                        hipStream_t stream, int real_bytes) {
    if (reinv_ns_supported(shape, real_bytes)) {
        reinv_ns_batch_launch(shape, items, n_items, grid, stream);
        return;
    }
    const bool a_lds = reinv_lds_bytes(shape.nmax, true) <= kMaxLdsBytes;
    const size_t lds = reinv_lds_bytes(shape.nmax, a_lds);
    if (real_bytes == 8)
        hipLaunchKernelGGL((k_reinv_batch<double>), dim3(grid), dim3(kReinvThreads), lds, stream, items, n_items,
                           lds);
    else
        hipLaunchKernelGGL((k_reinv_batch<float>), dim3(grid), dim3(kReinvThreads), lds, stream, items, n_items,
                           lds);
}

// Frozen-configuration evaluation (one configuration per wavefront).
template <typename Real, bool kInvLds>
__global__ __launch_bounds__(kWave) void k_frozen(DevSys s, int C, const double *pos,
                                                  const int *moved, const double *newpos,
                                                  double *scalars, double *gl, double *inv_before,
                                                  double *inv_after, double *sk, Real *inv_g,
                                                  double *A_g, double *sk_scratch) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = blockIdx.x;
    if (c >= C) return;
    const int lane = threadIdx.x;
    const int N = s.N;
    Real *inv_hbm = inv_g + (size_t)c * s.ninv;
    double *A = A_g + (size_t)c * s.nmax * s.nmax;
    WalkerView<Real> v = carve<Real>(smem, s, kInvLds, inv_hbm, A);
    const double *cp = pos + (size_t)c * 3 * N;
    for (int i = lane; i < N; i += kWave) {
        v.px[i] = cp[3 * i + 0];
        v.py[i] = cp[3 * i + 1];
        v.pz[i] = cp[3 * i + 2];
    }
    wave_sync();
    double *out = scalars + (size_t)c * VMC_FZ_COUNT;
    double ld[2] = {0.0, 0.0};
    int singular = 0;
    for (int d = 0; d < s.ndet; ++d) {
        ld[d] = reinvert<Real>(s, d, v, lane);
        if (!(ld[d] > -INFINITY)) singular = 1;
    }
    if (inv_before)
        for (int e = lane; e < s.ninv; e += kWave) inv_before[(size_t)c * s.ninv + e] = (double)v.inv[e];
    jastrow_rows<Real>(s, v, lane);
    double jp = 0.0;
    for (int i = lane; i < N; i += kWave) jp += v.U[i];
    const double J = 0.5 * wave_sum(jp);
    double *skc = sk_scratch + (size_t)c * 2 * s.ng;
    const Measured m0 = measure<Real>(s, v, lane, gl ? gl + (size_t)c * 4 * N : nullptr, skc);
    wave_sync();
    if (sk)
        for (int g = lane; g < 2 * s.ng; g += kWave) sk[(size_t)c * 2 * s.ng + g] = skc[g];

    const int p = moved[c];
    const double x = newpos[3 * c + 0], y = newpos[3 * c + 1], z = newpos[3 * c + 2];
    const double ox = v.px[p], oy = v.py[p], oz = v.pz[p];
    const Proposal pr = propose<Real>(s, v, p, x, y, z, lane);

    // Incremental Ewald updates of the move (energy_tracking.cpp:193-251, 280-365).
    double dq = 0.0;
    for (int g = lane; g < s.ng; g += kWave) {
        const int4 G = s.gvec[g];
        Real sn, cs, so, co;
        sincos_2pi<Real>(frac_phase(G.x, G.y, G.z, x * s.inv_L, y * s.inv_L, z * s.inv_L), &sn, &cs);
        sincos_2pi<Real>(frac_phase(G.x, G.y, G.z, ox * s.inv_L, oy * s.inv_L, oz * s.inv_L), &so,
                         &co);
        const double dr = (double)cs - (double)co, di = (double)sn - (double)so;
        dq += s.gw[g] * (2.0 * (skc[2 * g] * dr + skc[2 * g + 1] * di) + dr * dr + di * di);
    }
    const double vrecip_inc = m0.vrecip + s.pref * wave_sum(dq);
    double dv = 0.0;
    for (int j = lane; j < N; j += kWave) {
        if (j == p) continue;
        const Real rn = pair_dist<Real>(s, x, y, z, v, j, nullptr, nullptr, nullptr);
        const Real ro = pair_dist<Real>(s, ox, oy, oz, v, j, nullptr, nullptr, nullptr);
        dv += (double)ewald_pair<Real>(rn, s) - (double)ewald_pair<Real>(ro, s);
    }
    const double vreal_inc = m0.vreal + wave_sum(dv);

    accept_move<Real>(s, v, p, x, y, z, pr, lane);
    if (inv_after)
        for (int e = lane; e < s.ninv; e += kWave) inv_after[(size_t)c * s.ninv + e] = (double)v.inv[e];
    const Measured m1 = measure<Real>(s, v, lane, nullptr, nullptr);
    if (lane == 0) {
        out[VMC_FZ_LOGDET0] = ld[0];
        out[VMC_FZ_LOGDET1] = ld[1];
        out[VMC_FZ_JASTROW] = J;
        out[VMC_FZ_KINETIC] = m0.kinetic;
        out[VMC_FZ_VREAL] = m0.vreal;
        out[VMC_FZ_VRECIP] = m0.vrecip;
        out[VMC_FZ_ELOCAL] = m0.e_local;
        out[VMC_FZ_RATIO] = pr.ratio;
        out[VMC_FZ_DELTA_J] = pr.dj;
        out[VMC_FZ_VREAL_INC] = vreal_inc;
        out[VMC_FZ_VRECIP_INC] = vrecip_inc;
        out[VMC_FZ_ELOCAL_AFTER] = m1.e_local;
        out[VMC_FZ_KINETIC_AFTER] = m1.kinetic;
        out[VMC_FZ_JASTROW_AFTER] = J + pr.dj;
        out[VMC_FZ_SINGULAR] = (double)singular;
    }
}

}  // namespace vmc

// ================================================================== host side

using namespace vmc;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define VMC_HIP_TRY(expr)                                                                    \
    do {                                                                                     \
        hipError_t err__ = (expr);                                                           \
        if (err__ != hipSuccess)                                                             \
            return fail(VMC_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(err__));    \
    } while (0)

// Host copy of the constant tables and the device-side DevSys.
struct SystemTables {
    vmc_system desc{};
    DevSys dev{};
    std::vector<int4> orb;
    std::vector<int4> gvec;
    std::vector<double> gw;
    int4 *d_orb = nullptr;
    int4 *d_gvec = nullptr;
    double *d_gw = nullptr;
    int4 *d_sfseg = nullptr;
    double *d_sfw = nullptr;
    int real_bytes = 8;

    ~SystemTables() {
        if (d_orb) (void)hipFree(d_orb);
        if (d_gvec) (void)hipFree(d_gvec);
        if (d_gw) (void)hipFree(d_gw);
        if (d_sfseg) (void)hipFree(d_sfseg);
        if (d_sfw) (void)hipFree(d_sfw);
    }
};

int validate(const vmc_system *s) {
    if (!s) return fail(VMC_ERR_INVALID, "system descriptor is NULL");
    if (s->n_up < 1) return fail(VMC_ERR_INVALID, "n_up must be >= 1");
    if (s->n_dn < 0) return fail(VMC_ERR_INVALID, "n_dn must be >= 0");
    if (!(s->box_length > 0.0) || !std::isfinite(s->box_length))
        return fail(VMC_ERR_INVALID, "box_length must be finite and > 0");
    if (s->precision != VMC_PREC_FP64 && s->precision != VMC_PREC_MIXED)
        return fail(VMC_ERR_INVALID, "unknown precision %d", s->precision);
    if (s->move_kind != VMC_MOVE_UNIFORM && s->move_kind != VMC_MOVE_GAUSSIAN)
        return fail(VMC_ERR_INVALID, "unknown move kind %d", s->move_kind);
    if (s->erfc_kind != VMC_ERFC_AS7126 && s->erfc_kind != VMC_ERFC_EXACT)
        return fail(VMC_ERR_INVALID, "unknown erfc kind %d", s->erfc_kind);
    if (s->reinvert_interval < 0 || s->reinvert_interval > (1 << 24))
        return fail(VMC_ERR_INVALID, "reinvert_interval must be in [0, 2^24]");
    if (s->recip_nmax2 < 0) return fail(VMC_ERR_INVALID, "recip_nmax2 must be >= 0");
    if (s->n_up + s->n_dn > 4096) return fail(VMC_ERR_INVALID, "at most 4096 electrons");
    return VMC_OK;
}

int build_tables(const vmc_system *sys, uint64_t seed, SystemTables &t) {
    int rc = validate(sys);
    if (rc) return rc;
    t.desc = *sys;
    DevSys &d = t.dev;
    d.n_up = sys->n_up;
    d.n_dn = sys->n_dn;
    d.N = sys->n_up + sys->n_dn;
    d.ndet = sys->n_dn > 0 ? 2 : 1;
    d.n[0] = sys->n_up;
    d.n[1] = sys->n_dn;
    d.first[0] = 0;
    d.first[1] = sys->n_up;
    d.inv_off[0] = 0;
    d.inv_off[1] = sys->n_up * sys->n_up;
    d.ninv = sys->n_up * sys->n_up + sys->n_dn * sys->n_dn;
    d.nmax = std::max(sys->n_up, sys->n_dn);
    d.erfc_kind = sys->erfc_kind;
    d.move_kind = sys->move_kind;
    d.use_cut = sys->real_cutoff > 0.0 ? 1 : 0;
    // K counts accepted moves per determinant; a walker's counter adds up the
    // accepted moves of all its determinants, so it is due at K per
    // determinant on average.
    d.reinv_k = sys->reinvert_interval * d.ndet;
    d.L = sys->box_length;
    d.inv_L = 1.0 / sys->box_length;
    d.half_L = 0.5 * sys->box_length;
    d.a_same = sys->jastrow_a_same;
    d.a_anti = sys->jastrow_a_anti;
    d.b = sys->jastrow_b;
    d.rcut = sys->real_cutoff;
    constexpr double pi = std::numbers::pi;
    const GTable g = ewald_gvectors(sys->box_length, sys->ewald_alpha, sys->recip_nmax2);
    d.alpha = g.alpha;
    const double N = static_cast<double>(d.N), L = sys->box_length;
    // Self and background terms (energy_tracking.cpp:8-9 for alpha = 6/L).
    double self_term, background;
    if (sys->ewald_alpha > 0.0) {
        self_term = -g.alpha * N / std::sqrt(pi);
        background = -pi * N * N / (2.0 * g.alpha * g.alpha * L * L * L);
    } else {
        self_term = -6.0 * N / (std::sqrt(pi) * L);
        background = -pi * N * N / (72.0 * L);
    }
    d.e_const = self_term + background;
    d.pref = 1.0 / (2.0 * pi * L * L * L);
    d.kunit = 2.0 * pi / L;
    d.key0 = static_cast<uint32_t>(seed);
    d.key1 = static_cast<uint32_t>(seed >> 32);
    d.ng = static_cast<int>(g.m.size());
    t.orb.assign(static_cast<size_t>(d.ndet) * d.nmax, make_int4(0, 0, 0, 0));
    for (int det = 0; det < d.ndet; ++det) {
        const ShellTable sh = fill_shells(d.n[det]);
        for (int j = 0; j < d.n[det]; ++j) {
            const auto &nv = sh.nvec[static_cast<size_t>(sh.orb_k[static_cast<size_t>(j)])];
            t.orb[static_cast<size_t>(det * d.nmax + j)] =
                make_int4(nv[0], nv[1], nv[2], sh.orb_type[static_cast<size_t>(j)]);
        }
    }
    t.gvec.resize(g.m.size());
    for (size_t i = 0; i < g.m.size(); ++i) t.gvec[i] = make_int4(g.m[i][0], g.m[i][1], g.m[i][2], 0);
    t.gw = g.w;
    t.real_bytes = sys->precision == VMC_PREC_FP64 ? 8 : 4;
    VMC_HIP_TRY(hipMalloc(&t.d_orb, sizeof(int4) * std::max<size_t>(t.orb.size(), 1)));
    VMC_HIP_TRY(hipMemcpy(t.d_orb, t.orb.data(), sizeof(int4) * t.orb.size(), hipMemcpyHostToDevice));
    VMC_HIP_TRY(hipMalloc(&t.d_gvec, sizeof(int4) * std::max<size_t>(t.gvec.size(), 1)));
    VMC_HIP_TRY(hipMalloc(&t.d_gw, sizeof(double) * std::max<size_t>(t.gw.size(), 1)));
    if (!t.gvec.empty()) {
        VMC_HIP_TRY(hipMemcpy(t.d_gvec, t.gvec.data(), sizeof(int4) * t.gvec.size(), hipMemcpyHostToDevice));
        VMC_HIP_TRY(hipMemcpy(t.d_gw, t.gw.data(), sizeof(double) * t.gw.size(), hipMemcpyHostToDevice));
    }
    d.orb = t.d_orb;
    d.gvec = t.d_gvec;
    d.gw = t.d_gw;
    const SfSegments sg = sf_segments(g);
    std::vector<int4> seg(sg.seg.size());
    for (size_t i = 0; i < seg.size(); ++i) seg[i] = make_int4(sg.seg[i][0], sg.seg[i][1], sg.seg[i][2], sg.seg[i][3]);
    VMC_HIP_TRY(hipMalloc(&t.d_sfseg, sizeof(int4) * std::max<size_t>(seg.size(), 1)));
    VMC_HIP_TRY(hipMalloc(&t.d_sfw, sizeof(double) * std::max<size_t>(sg.w.size(), 1)));
    if (!seg.empty()) {
        VMC_HIP_TRY(hipMemcpy(t.d_sfseg, seg.data(), sizeof(int4) * seg.size(), hipMemcpyHostToDevice));
        VMC_HIP_TRY(hipMemcpy(t.d_sfw, sg.w.data(), sizeof(double) * sg.w.size(), hipMemcpyHostToDevice));
    }
    d.nseg = static_cast<int>(seg.size());
    d.sfseg = t.d_sfseg;
    d.sfw = t.d_sfw;
    return VMC_OK;
}

template <typename T>
int dev_alloc(T **p, size_t count) {
    if (hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
        *p = nullptr;
        return fail(VMC_ERR_MEMORY, "hipMalloc of %zu bytes failed", sizeof(T) * count);
    }
    return VMC_OK;
}

size_t lds_for(const DevSys &d, int real_bytes, bool inv_lds) {
    return walker_lds_bytes(d.N, d.nmax, d.ninv, real_bytes, inv_lds);
}

}  // namespace

constexpr int kMaxWalkerGroups = 8;

struct vmc_ensemble {
    SystemTables tables;
    DevState st{};
    int64_t W = 0;
    uint64_t seed = 0;
    uint64_t sweeps_done = 0;
    double step0 = 0.0;
    bool positions_ready = false;
    int kernel = VMC_KERNEL_AUTO;
    int rr_variant = 0;
    KernelTimer timer;
    int inv_placement = 0;  // generic kernel: 0 LDS when it fits, 1 LDS, 2 HBM
    int walker_threads = 64;  // generic kernel: threads (1, 2 or 4 wavefronts) per walker
    bool in_batch = false;  // owned by a vmc_batch: swept and destroyed through it
    // Mixed precision: [W][ninv] FP64 inverses that a measured sweep of the
    // multi-wavefront or generic kernels re-inverts into for its FP64 local
    // energy (DevState::inv64 of that sweep's re-inversion).
    double *inv64 = nullptr;
    // Walker groups of the register kernels' measured sweeps: the walkers are
    // split into `groups` contiguous ranges, each swept and measured on a
    // stream of its own (forked from and joined back into the caller's
    // stream), so one group's sweep overlaps another's local energy.
    // Events are made on first use.
    int groups = 1;
    hipEvent_t gjoin[kMaxWalkerGroups] = {};
    hipEvent_t gfork = nullptr;

    ~vmc_ensemble() {
        for (int g = 0; g < kMaxWalkerGroups; ++g)
            if (gjoin[g]) (void)hipEventDestroy(gjoin[g]);
        if (gfork) (void)hipEventDestroy(gfork);
        if (inv64) (void)hipFree(inv64);
        if (st.pos) (void)hipFree(st.pos);
        if (st.inv) (void)hipFree(st.inv);
        if (st.sigma) (void)hipFree(st.sigma);
        if (st.stats) (void)hipFree(st.stats);
        if (st.logdet) (void)hipFree(st.logdet);
        if (st.since) (void)hipFree(st.since);
        if (st.flags) (void)hipFree(st.flags);
        if (st.scratch) (void)hipFree(st.scratch);
        if (st.drift) (void)hipFree(st.drift);
        if (st.resid) (void)hipFree(st.resid);
        if (st.jrow) (void)hipFree(st.jrow);
        if (st.due) (void)hipFree(st.due);
        if (st.due_list) (void)hipFree(st.due_list);
    }
};

namespace {

template <typename Real>
int launch_init(vmc_ensemble *e, int draw, hipStream_t stream) {
    const DevSys &d = e->tables.dev;
    const size_t lds = lds_for(d, sizeof(Real), false);
    hipLaunchKernelGGL((k_init<Real>), dim3(static_cast<unsigned>(e->W)), dim3(kWave), lds, stream,
                       d, e->st, draw);
    VMC_HIP_TRY(hipGetLastError());
    return VMC_OK;
}

bool rr_supported(const DevSys &d) { return rr_supported_sizes(d.ndet, d.n_up, d.n_dn, d.ng); }
bool rrw_supported(const DevSys &d, int real_bytes) {
    return rrw_supported_sizes(d.ndet, d.n_up, d.n_dn, d.ng, real_bytes);
}

// Sweep kernel family of an ensemble: the register-resident kernels where an
// instantiation exists, in FP64 the delayed-update kernel where the inverses
// do not fit the generic kernel's LDS, else the generic kernel; or the
// family the caller forced (vmc_ensemble_set_kernel validated it).
enum { kFamGeneric = 0, kFamRegister = 1, kFamMultiwave = 2, kFamDelayed = 3 };
int sweep_family(const vmc_ensemble *e) {
    const DevSys &d = e->tables.dev;
    const int rb = e->tables.real_bytes;
    if (e->kernel == VMC_KERNEL_GENERIC) return kFamGeneric;
    if (e->kernel == VMC_KERNEL_DELAYED) return kFamDelayed;
    if (rr_supported(d)) return kFamRegister;
    if (rrw_supported(d, rb)) return kFamMultiwave;
    if (e->kernel == VMC_KERNEL_AUTO && rb == 8 && du_supported(d) && lds_for(d, 8, true) > kMaxLdsBytes)
        return kFamDelayed;
    return kFamGeneric;
}

// Delayed-update sweeps (FP64), one launch per sweep when each is measured
// or followed by the periodic re-inversion.  The local energy is the FP64
// multi-wavefront measurement where that has an instantiation (123 and 147
// per spin among them: it reads the inverse rows from memory), else the
// generic one; both take the inverse the sweep left current in HBM.
int launch_sweep_du(vmc_ensemble *e, int n_sweeps, int flags, int block_size, hipStream_t stream) {
    const DevSys &d = e->tables.dev;
    const unsigned grid = static_cast<unsigned>(e->W);
    const bool measure = flags & VMC_SWEEP_MEASURE;
    const bool reinv = d.reinv_k > 0;
    const int per_launch = (measure || reinv) ? 1 : n_sweeps;
    for (int i = 0; i < n_sweeps; i += per_launch) {
        const int n = n_sweeps - i < per_launch ? n_sweeps - i : per_launch;
        const unsigned long long s0 = (unsigned long long)e->sweeps_done + (unsigned long long)i;
        hipEvent_t ev = nullptr;
        bool timed = e->timer.begin(stream, ev);
        if (!du_launch(d, e->st, grid, n, s0, flags & ~VMC_SWEEP_MEASURE, SweepGiven{}, stream))
            return fail(VMC_ERR_INVALID, "no delayed-update kernel for this system");
        if (timed) e->timer.end(stream, ev, kTimeSweep);
        if (reinv) {
            timed = e->timer.begin(stream, ev);
            reinv_generic_launch(d, e->st, grid, stream, 8);
            if (timed) e->timer.end(stream, ev, kTimeReinv);
        }
        if (measure) {
            timed = e->timer.begin(stream, ev);
            if (!rrw_measure64(d, e->st, grid, block_size, nullptr, stream))
                hipLaunchKernelGGL((k_measure_parts<double, false>), dim3(grid),
                                   dim3(static_cast<unsigned>(e->walker_threads)), lds_for(d, 8, false), stream, d,
                                   e->st, static_cast<double *>(nullptr), block_size);
            if (timed) e->timer.end(stream, ev, kTimeMeasure);
        }
    }
    VMC_HIP_TRY(hipGetLastError());
    return VMC_OK;
}

// The state of walkers [b, b + n) of an ensemble: every per-walker array
// offset to walker b, w0 = b for the Philox counters.  The due-walker count
// stays with the group holding walker 0 (the measurement clears it there).
DevState walker_view(const vmc_ensemble *e, int64_t b, int64_t n) {
    const DevSys &d = e->tables.dev;
    const DevState &a = e->st;
    const size_t u = static_cast<size_t>(b);
    DevState v = a;
    v.W = static_cast<int>(n);
    v.w0 = a.w0 + static_cast<int>(b);
    v.pos = a.pos + u * 3 * d.N;
    v.inv = static_cast<unsigned char *>(a.inv) + u * d.ninv * e->tables.real_bytes;
    v.sigma = a.sigma + u;
    v.stats = a.stats + u * kStatCount;
    v.logdet = a.logdet + 2 * u;
    v.since = a.since + u;
    v.flags = a.flags + u;
    v.scratch = a.scratch ? a.scratch + u * d.nmax * d.nmax : nullptr;
    v.drift = a.drift + u;
    if (a.resid) v.resid = a.resid + u * static_cast<size_t>(ns_resid_floats(d.nmax, e->tables.real_bytes));
    v.jrow = a.jrow + u * d.N;
    v.due = b == 0 ? a.due : nullptr;
    v.due_list = a.due_list + u;
    return v;
}

// The walker-group streams of the current device, kept for the life of the
// process and shared by every ensemble on that device (ensembles used from
// several host threads at once serialise on them, in stream order).  Each
// is made with a CU mask of every CU of the device, which gives it a
// hardware queue of its own: plain streams share the process's few hardware
// queues, and two groups on one queue run one after the other (measured:
// the groups' overlap came and went with the order streams were made in).
int group_streams(int G, hipStream_t out[kMaxWalkerGroups]) {
    static std::mutex mu;
    static std::vector<std::array<hipStream_t, kMaxWalkerGroups>> per_device;
    int dev = 0;
    VMC_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (per_device.size() <= static_cast<size_t>(dev)) per_device.resize(static_cast<size_t>(dev) + 1);
    auto &ss = per_device[static_cast<size_t>(dev)];
    for (int g = 0; g < G; ++g) {
        hipStream_t &st = ss[static_cast<size_t>(g)];
        if (!st) {
            int cus = 0;
            VMC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            std::vector<uint32_t> mask(static_cast<size_t>((cus + 31) / 32), 0u);
            for (int c = 0; c < cus; ++c) mask[static_cast<size_t>(c / 32)] |= 1u << (c % 32);
            VMC_HIP_TRY(hipExtStreamCreateWithCUMask(&st, static_cast<uint32_t>(mask.size()), mask.data()));
        }
        out[g] = st;
    }
    return VMC_OK;
}

// Walker groups for a register-kernel launch: the caller's count (one by
// default), and one for sweeps that need a listed re-inversion (FP64 with
// K > 0) or are timed per kernel.  Only measured sweeps are split.
int walker_groups(const vmc_ensemble *e, int flags) {
    const DevSys &d = e->tables.dev;
    if (!(flags & VMC_SWEEP_MEASURE) || e->timer.on) return 1;
    if (e->tables.real_bytes == 8 && d.reinv_k > 0) return 1;
    return static_cast<int>(std::min<int64_t>(e->groups < 1 ? 1 : e->groups, e->W));
}

template <typename Real>
int launch_sweep(vmc_ensemble *e, int n_sweeps, int flags, int block_size, hipStream_t stream) {
    const DevSys &d = e->tables.dev;
    const int fam = sweep_family(e);
    if constexpr (sizeof(Real) == 8)
        if (fam == kFamDelayed) return launch_sweep_du(e, n_sweeps, flags, block_size, stream);
    if (fam == kFamRegister) {
        const unsigned long long s0 = (unsigned long long)e->sweeps_done;
        const auto run = [&](const DevState &st, unsigned grid, hipStream_t s) {
            return sizeof(Real) == 4
                       ? rr_launch_f32(d, st, grid, n_sweeps, s0, flags, block_size, e->rr_variant, s, e->timer)
                       : rr_launch_f64(d, st, grid, n_sweeps, s0, flags, block_size, e->rr_variant, s, e->timer);
        };
        const int G = walker_groups(e, flags);
        if (G <= 1) {
            if (!run(e->st, static_cast<unsigned>(e->W), stream))
                return fail(VMC_ERR_INVALID, "no register-resident instantiation for this system");
            VMC_HIP_TRY(hipGetLastError());
            return VMC_OK;
        }
        // Every group on a library stream of its own, made together so that
        // they sit on distinct hardware queues (the caller's stream, wherever
        // it sits, only forks and joins).  Fork: the group streams wait for
        // the caller's stream; join: the caller's stream waits for them.
        hipStream_t gs[kMaxWalkerGroups];
        if (const int rc = group_streams(G, gs)) return rc;
        if (!e->gfork) VMC_HIP_TRY(hipEventCreateWithFlags(&e->gfork, hipEventDisableTiming));
        for (int g = 0; g < G; ++g)
            if (!e->gjoin[g]) VMC_HIP_TRY(hipEventCreateWithFlags(&e->gjoin[g], hipEventDisableTiming));
        VMC_HIP_TRY(hipEventRecord(e->gfork, stream));
        for (int g = 0; g < G; ++g) VMC_HIP_TRY(hipStreamWaitEvent(gs[g], e->gfork, 0));
        for (int g = 0; g < G; ++g) {
            const int64_t b = e->W * g / G, n = e->W * (g + 1) / G - b;
            if (!run(walker_view(e, b, n), static_cast<unsigned>(n), gs[g]))
                return fail(VMC_ERR_INVALID, "no register-resident instantiation for this system");
        }
        for (int g = 0; g < G; ++g) {
            VMC_HIP_TRY(hipEventRecord(e->gjoin[g], gs[g]));
            VMC_HIP_TRY(hipStreamWaitEvent(stream, e->gjoin[g], 0));
        }
        VMC_HIP_TRY(hipGetLastError());
        return VMC_OK;
    }
    if (fam == kFamMultiwave) {
        const unsigned grid = static_cast<unsigned>(e->W);
        const unsigned long long s0 = (unsigned long long)e->sweeps_done;
        const ReinvLaunch rl{static_cast<int>(sizeof(Real))};
        const bool ok = sizeof(Real) == 4
                            ? rrw_launch_f32(d, e->st, grid, n_sweeps, s0, flags, block_size, stream, e->timer, rl,
                                             e->inv64)
                            : rrw_launch_f64(d, e->st, grid, n_sweeps, s0, flags, block_size, stream, e->timer, rl,
                                             nullptr);
        if (!ok) return fail(VMC_ERR_INVALID, "no register-resident instantiation for this system");
        VMC_HIP_TRY(hipGetLastError());
        return VMC_OK;
    }
    const size_t lds_in = lds_for(d, sizeof(Real), true);
    const unsigned grid = static_cast<unsigned>(e->W);
    const bool inv_lds = lds_in <= kMaxLdsBytes && e->inv_placement != 2;
    const size_t lds = inv_lds ? lds_in : lds_for(d, sizeof(Real), false);
    if (lds > kMaxLdsBytes) return fail(VMC_ERR_INVALID, "system too large for one walker per wavefront");
    // With periodic re-inversion, one sweep per launch, each followed by the
    // FP64 re-inversion of the walkers that came due.  In mixed precision a
    // measured sweep re-inverts every walker with its FP64 result in inv64
    // and takes the local energy from that in FP64 (k_measure_parts).
    const bool exact = sizeof(Real) == 4 && (flags & VMC_SWEEP_MEASURE) && e->inv64;
    const bool reinv = d.reinv_k > 0 || exact;
    const int per_launch = reinv ? 1 : n_sweeps;
    const int sweep_flags = exact ? flags & ~VMC_SWEEP_MEASURE : flags;
    DevState st_refresh = e->st, st64 = e->st;
    st_refresh.inv64 = e->inv64;
    st64.inv = e->inv64;
    const size_t lds64 = lds_for(d, 8, false);
    for (int i = 0; i < n_sweeps; i += per_launch) {
        const int n = n_sweeps - i < per_launch ? n_sweeps - i : per_launch;
        const unsigned long long s0 = (unsigned long long)e->sweeps_done + (unsigned long long)i;
        hipEvent_t ev = nullptr;
        bool timed = e->timer.begin(stream, ev);
        if (inv_lds)
            hipLaunchKernelGGL((k_sweep<Real, true>), dim3(grid), dim3(e->walker_threads), lds, stream, d, e->st, n,
                               s0, sweep_flags, block_size);
        else
            hipLaunchKernelGGL((k_sweep<Real, false>), dim3(grid), dim3(e->walker_threads), lds, stream, d, e->st, n,
                               s0, sweep_flags, block_size);
        if (timed) e->timer.end(stream, ev, kTimeSweep);
        if (reinv) {
            timed = e->timer.begin(stream, ev);
            reinv_generic_launch(d, exact ? st_refresh : e->st, grid, stream, static_cast<int>(sizeof(Real)));
            if (timed) e->timer.end(stream, ev, kTimeReinv);
        }
        if (exact) {
            timed = e->timer.begin(stream, ev);
            hipLaunchKernelGGL((k_measure_parts<double, false>), dim3(grid), dim3(e->walker_threads), lds64, stream,
                               d, st64, static_cast<double *>(nullptr), block_size);
            if (timed) e->timer.end(stream, ev, kTimeMeasure);
        }
    }
    VMC_HIP_TRY(hipGetLastError());
    return VMC_OK;
}

}  // namespace

namespace vmc {
// Sets the message vmc_last_error() returns on this thread (for the
// library's other translation units); returns code.
int set_last_error(int code, const char *msg) { return fail(code, "%s", msg); }
}  // namespace vmc

extern "C" {

const char *vmc_last_error(void) { return g_last_error.c_str(); }

int vmc_device_check(int32_t *device_count) {
    int n = 0;
    VMC_HIP_TRY(hipGetDeviceCount(&n));
    if (device_count) *device_count = n;
    if (n < 1) return fail(VMC_ERR_DEVICE, "no HIP device visible");
    hipDeviceProp_t prop;
    int dev = 0;
    VMC_HIP_TRY(hipGetDevice(&dev));
    VMC_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(VMC_ERR_DEVICE, "device %d is %s, this library is built for gfx950", dev, prop.gcnArchName);
    return VMC_OK;
}

int vmc_system_sizes(const vmc_system *sys, int32_t *n_k, int32_t *n_g, int32_t *n_inv) {
    int rc = validate(sys);
    if (rc) return rc;
    if (n_k) {
        n_k[0] = fill_shells(sys->n_up).nk;
        n_k[1] = sys->n_dn > 0 ? fill_shells(sys->n_dn).nk : 0;
    }
    if (n_g) *n_g = static_cast<int32_t>(ewald_gvectors(sys->box_length, sys->ewald_alpha, sys->recip_nmax2).m.size());
    if (n_inv) *n_inv = sys->n_up * sys->n_up + sys->n_dn * sys->n_dn;
    return VMC_OK;
}

int vmc_ensemble_footprint(const vmc_system *sys, int64_t n_walkers, int64_t *bytes) {
    if (!bytes) return fail(VMC_ERR_INVALID, "bytes is NULL");
    int32_t nk[2], ng = 0, ninv = 0;
    int rc = vmc_system_sizes(sys, nk, &ng, &ninv);
    if (rc) return rc;
    if (n_walkers < 1) return fail(VMC_ERR_INVALID, "n_walkers must be >= 1");
    const int64_t N = sys->n_up + sys->n_dn;
    const int64_t nmax = std::max(sys->n_up, sys->n_dn);
    const int64_t rb = sys->precision == VMC_PREC_MIXED ? 4 : 8;
    const int64_t per_walker = 3 * N * 8 + ninv * rb + nmax * nmax * 8 + kStatCount * 8 +
                               8 /* sigma */ + 16 /* logdet */ + 12 /* since, flags, due list */ + 8 /* drift */ +
                               4 * ns_resid_floats(static_cast<int>(nmax), static_cast<int>(rb)) +
                               (rb == 4 ? ninv * 8 : 0) /* FP64 inverses of a measured sweep */ +
                               8 * N /* Jastrow row sums */;
    const int64_t tables = 2 * nmax * 16 + static_cast<int64_t>(ng) * (16 + 8) + 8 /* due counters */;
    *bytes = n_walkers * per_walker + tables;
    return VMC_OK;
}

int vmc_ensemble_create(const vmc_system *sys, int64_t n_walkers, uint64_t seed, double step0,
                        vmc_ensemble **out) {
    if (!out) return fail(VMC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n_walkers < 1 || n_walkers > (int64_t(1) << 24)) return fail(VMC_ERR_INVALID, "n_walkers out of range");
    if (!(step0 >= 0.0) || !std::isfinite(step0)) return fail(VMC_ERR_INVALID, "step0 must be finite and >= 0");
    auto *e = new vmc_ensemble();
    int rc = build_tables(sys, seed, e->tables);
    if (rc) { delete e; return rc; }
    const DevSys &d = e->tables.dev;
    if (lds_for(d, e->tables.real_bytes, false) > kMaxLdsBytes) {
        delete e;
        return fail(VMC_ERR_INVALID, "N=%d exceeds the per-walker LDS budget", d.N);
    }
    e->W = n_walkers;
    e->seed = seed;
    e->step0 = step0;
    const size_t W = static_cast<size_t>(n_walkers);
    DevState &st = e->st;
    st.W = static_cast<int>(n_walkers);
    if ((rc = dev_alloc(&st.pos, W * 3 * d.N)) || (rc = dev_alloc(reinterpret_cast<unsigned char **>(&st.inv), W * d.ninv * e->tables.real_bytes)) ||
        (rc = dev_alloc(&st.sigma, W)) || (rc = dev_alloc(&st.stats, W * kStatCount)) ||
        (rc = dev_alloc(&st.logdet, W * 2)) || (rc = dev_alloc(&st.since, W)) ||
        (rc = dev_alloc(&st.flags, W)) || (rc = dev_alloc(&st.scratch, W * d.nmax * d.nmax)) || (rc = dev_alloc(&st.drift, W)) ||
        (rc = dev_alloc(&st.jrow, W * d.N)) || (rc = dev_alloc(&st.due, 2)) || (rc = dev_alloc(&st.due_list, W))) {
        delete e;
        return rc;
    }
    // (every mixed ensemble re-inverts: periodically when K > 0, and on each
    // measured sweep of the multi-wavefront and generic kernels)
    if (const int64_t nr = ns_resid_floats(d.nmax, e->tables.real_bytes);
        nr > 0 && (rc = dev_alloc(&st.resid, W * static_cast<size_t>(nr)))) {
        delete e;
        return rc;
    }
    if (e->tables.real_bytes == 4 && (rc = dev_alloc(&e->inv64, W * d.ninv))) {
        delete e;
        return rc;
    }
    std::vector<double> sig(W, step0);
    if (hipMemcpy(st.sigma, sig.data(), sizeof(double) * W, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(st.stats, 0, sizeof(double) * W * kStatCount) != hipSuccess ||
        hipMemset(st.since, 0, sizeof(int) * W) != hipSuccess ||
        hipMemset(st.flags, 0, sizeof(int) * W) != hipSuccess ||
        hipMemset(st.logdet, 0, sizeof(double) * W * 2) != hipSuccess ||
        hipMemset(st.drift, 0, sizeof(double) * W) != hipSuccess ||
        hipMemset(st.due, 0, sizeof(unsigned) * 2) != hipSuccess) {
        delete e;
        return fail(VMC_ERR_DEVICE, "ensemble initialisation copy failed");
    }
    *out = e;
    return VMC_OK;
}

int vmc_ensemble_destroy(vmc_ensemble *e) {
    if (e && e->in_batch) return fail(VMC_ERR_STATE, "ensemble belongs to a batch");
    delete e;
    return VMC_OK;
}

static int finish_init(vmc_ensemble *e, hipStream_t stream) {
    VMC_HIP_TRY(hipStreamSynchronize(stream));
    std::vector<int> fl(static_cast<size_t>(e->W));
    VMC_HIP_TRY(hipMemcpy(fl.data(), e->st.flags, sizeof(int) * fl.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < fl.size(); ++i)
        if (fl[i]) return fail(VMC_ERR_SINGULAR, "walker %zu: no non-singular configuration", i);
    e->positions_ready = true;
    return VMC_OK;
}

int vmc_ensemble_init_positions(vmc_ensemble *e, void *stream) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = e->tables.real_bytes == 8 ? launch_init<double>(e, 1, s) : launch_init<float>(e, 1, s);
    if (rc) return rc;
    return finish_init(e, s);
}

int vmc_ensemble_set_positions(vmc_ensemble *e, const double *pos, void *stream) {
    if (!e || !pos) return fail(VMC_ERR_INVALID, "NULL argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int N = e->tables.dev.N;
    const size_t W = static_cast<size_t>(e->W);
    std::vector<double> soa(W * 3 * N);
    const double L = e->tables.dev.L;
    for (size_t w = 0; w < W; ++w)
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) {
                const double x = pos[(w * N + i) * 3 + c];
                if (!(x >= 0.0 && x < L)) return fail(VMC_ERR_INVALID, "position outside [0, L)");
                soa[(w * 3 + c) * N + i] = x;
            }
    VMC_HIP_TRY(hipMemcpyAsync(e->st.pos, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice, s));
    int rc = e->tables.real_bytes == 8 ? launch_init<double>(e, 0, s) : launch_init<float>(e, 0, s);
    if (rc) return rc;
    return finish_init(e, s);
}

int vmc_ensemble_get_positions(vmc_ensemble *e, double *pos, void *stream) {
    if (!e || !pos) return fail(VMC_ERR_INVALID, "NULL argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int N = e->tables.dev.N;
    const size_t W = static_cast<size_t>(e->W);
    std::vector<double> soa(W * 3 * N);
    VMC_HIP_TRY(hipMemcpyAsync(soa.data(), e->st.pos, sizeof(double) * soa.size(), hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    for (size_t w = 0; w < W; ++w)
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) pos[(w * N + i) * 3 + c] = soa[(w * 3 + c) * N + i];
    return VMC_OK;
}

int vmc_ensemble_sweep(vmc_ensemble *e, int32_t n_sweeps, int32_t flags, int32_t block_size,
                       void *stream) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    if (!e->positions_ready) return fail(VMC_ERR_STATE, "positions not initialised");
    if (e->in_batch) return fail(VMC_ERR_STATE, "ensemble belongs to a batch; sweep the batch");
    if (n_sweeps < 0) return fail(VMC_ERR_INVALID, "n_sweeps must be >= 0");
    if (block_size < 1) return fail(VMC_ERR_INVALID, "block_size must be >= 1");
    if (flags & ~(VMC_SWEEP_MEASURE | VMC_SWEEP_ADAPT)) return fail(VMC_ERR_INVALID, "unknown sweep flags");
    if (n_sweeps == 0) return VMC_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = e->tables.real_bytes == 8 ? launch_sweep<double>(e, n_sweeps, flags, block_size, s)
                                       : launch_sweep<float>(e, n_sweeps, flags, block_size, s);
    if (rc) return rc;
    e->sweeps_done += static_cast<uint64_t>(n_sweeps);
    return VMC_OK;
}

int vmc_ensemble_sweep_traced(vmc_ensemble *e, int32_t flags, int32_t block_size, int64_t walker,
                              double *e_local, int32_t *moved, int32_t *accepted, double *pos, void *stream) {
    if (!e || !e_local || !moved || !accepted) return fail(VMC_ERR_INVALID, "NULL argument");
    if (!e->positions_ready) return fail(VMC_ERR_STATE, "positions not initialised");
    if (e->in_batch) return fail(VMC_ERR_STATE, "ensemble belongs to a batch; sweep the batch");
    if (e->tables.real_bytes != 8) return fail(VMC_ERR_INVALID, "per-move local energies need an FP64 system");
    if (walker < 0 || walker >= e->W) return fail(VMC_ERR_INVALID, "walker %lld out of range", (long long)walker);
    if (block_size < 1) return fail(VMC_ERR_INVALID, "block_size must be >= 1");
    if (flags & ~(VMC_SWEEP_MEASURE | VMC_SWEEP_ADAPT)) return fail(VMC_ERR_INVALID, "unknown sweep flags");
    const DevSys &d = e->tables.dev;
    const size_t N = static_cast<size_t>(d.N);
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *d_e = nullptr, *d_pos = nullptr;
    int *d_mv = nullptr, *d_acc = nullptr;
    struct Guard {
        std::vector<void *> p;
        ~Guard() { for (void *q : p) if (q) (void)hipFree(q); }
    } guard;
    int rc;
    if ((rc = dev_alloc(&d_e, N)) || (rc = dev_alloc(&d_mv, N)) || (rc = dev_alloc(&d_acc, N)) ||
        (pos && (rc = dev_alloc(&d_pos, N * N * 3)))) {
        guard.p = {d_e, d_mv, d_acc, d_pos};
        return rc;
    }
    guard.p = {d_e, d_mv, d_acc, d_pos};
    const StepTrace tr{static_cast<long long>(walker), d_e, d_mv, d_acc, d_pos};
    const size_t lds_in = lds_for(d, 8, true);
    const bool inv_lds = lds_in <= kMaxLdsBytes && e->inv_placement != 2;
    const size_t lds = inv_lds ? lds_in : lds_for(d, 8, false);
    if (lds > kMaxLdsBytes) return fail(VMC_ERR_INVALID, "system too large for one walker per wavefront");
    const dim3 grid(static_cast<unsigned>(e->W)), block(static_cast<unsigned>(e->walker_threads));
    const unsigned long long s0 = e->sweeps_done;
    if (inv_lds) hipLaunchKernelGGL((k_sweep_traced<true>), grid, block, lds, s, d, e->st, s0, flags, block_size, tr);
    else hipLaunchKernelGGL((k_sweep_traced<false>), grid, block, lds, s, d, e->st, s0, flags, block_size, tr);
    if (d.reinv_k > 0) reinv_generic_launch(d, e->st, grid.x, s, 8);
    VMC_HIP_TRY(hipGetLastError());
    VMC_HIP_TRY(hipMemcpyAsync(e_local, d_e, sizeof(double) * N, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipMemcpyAsync(moved, d_mv, sizeof(int) * N, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipMemcpyAsync(accepted, d_acc, sizeof(int) * N, hipMemcpyDeviceToHost, s));
    if (pos) VMC_HIP_TRY(hipMemcpyAsync(pos, d_pos, sizeof(double) * N * N * 3, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    e->sweeps_done += 1;
    return VMC_OK;
}

int vmc_ensemble_sweep_given(vmc_ensemble *e, const double *proposals, const double *uniforms, double *trace,
                             void *stream) {
    if (!e || !proposals || !uniforms || !trace) return fail(VMC_ERR_INVALID, "NULL argument");
    if (!e->positions_ready) return fail(VMC_ERR_STATE, "positions not initialised");
    if (e->in_batch) return fail(VMC_ERR_STATE, "ensemble belongs to a batch; sweep the batch");
    const DevSys &d = e->tables.dev;
    const size_t n = static_cast<size_t>(e->W) * static_cast<size_t>(d.N);
    for (size_t i = 0; i < 3 * n; ++i)
        if (!(proposals[i] >= 0.0 && proposals[i] < d.L)) return fail(VMC_ERR_INVALID, "proposal outside [0, L)");
    for (size_t i = 0; i < n; ++i)
        if (uniforms[i] != uniforms[i]) return fail(VMC_ERR_INVALID, "accept uniform is NaN");
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *d_prop = nullptr, *d_u = nullptr, *d_tr = nullptr;
    struct Guard {
        std::vector<void *> p;
        ~Guard() { for (void *q : p) if (q) (void)hipFree(q); }
    } guard;
    int rc;
    if ((rc = dev_alloc(&d_prop, 3 * n)) || (rc = dev_alloc(&d_u, n)) || (rc = dev_alloc(&d_tr, n * kTraceCount))) {
        guard.p = {d_prop, d_u, d_tr};
        return rc;
    }
    guard.p = {d_prop, d_u, d_tr};
    VMC_HIP_TRY(hipMemcpyAsync(d_prop, proposals, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    VMC_HIP_TRY(hipMemcpyAsync(d_u, uniforms, sizeof(double) * n, hipMemcpyHostToDevice, s));
    const SweepGiven g{d_prop, d_u, d_tr};
    const unsigned grid = static_cast<unsigned>(e->W);
    const int rb = e->tables.real_bytes;
    bool done = false;
    const int fam = sweep_family(e);
    if (fam == kFamRegister) {
        done = rb == 4 ? rr_given_f32(d, e->st, grid, g, s) : rr_given_f64(d, e->st, grid, g, s);
    } else if (fam == kFamMultiwave) {
        const ReinvLaunch rl{rb};
        done = rb == 4 ? rrw_given_f32(d, e->st, grid, g, s, rl) : rrw_given_f64(d, e->st, grid, g, s, rl);
    } else if (fam == kFamDelayed) {
        done = du_launch(d, e->st, grid, 1, 0, 0, g, s);
        if (done && d.reinv_k > 0) reinv_generic_launch(d, e->st, grid, s, rb);
    }
    if (!done) {
        const size_t lds_in = lds_for(d, rb, true);
        const bool inv_lds = lds_in <= kMaxLdsBytes && e->inv_placement != 2;
        const size_t lds = inv_lds ? lds_in : lds_for(d, rb, false);
        const dim3 block(static_cast<unsigned>(e->walker_threads));
        if (rb == 8) {
            if (inv_lds) hipLaunchKernelGGL((k_sweep_given<double, true>), dim3(grid), block, lds, s, d, e->st, g);
            else hipLaunchKernelGGL((k_sweep_given<double, false>), dim3(grid), block, lds, s, d, e->st, g);
        } else {
            if (inv_lds) hipLaunchKernelGGL((k_sweep_given<float, true>), dim3(grid), block, lds, s, d, e->st, g);
            else hipLaunchKernelGGL((k_sweep_given<float, false>), dim3(grid), block, lds, s, d, e->st, g);
        }
        if (d.reinv_k > 0) reinv_generic_launch(d, e->st, grid, s, rb);
    }
    VMC_HIP_TRY(hipGetLastError());
    VMC_HIP_TRY(hipMemcpyAsync(trace, d_tr, sizeof(double) * n * kTraceCount, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    return VMC_OK;
}

}  // extern "C"

namespace {

// The measurement kernels of vmc_ensemble_measure writing every walker's
// parts into d_parts [W][kPartCount] on stream s.  Allocates nothing: the
// mixed-precision FP64 inverses and the re-inversion workspaces belong to
// the ensemble.
int launch_parts(vmc_ensemble *e, double *d_parts, hipStream_t s) {
    const DevSys &d = e->tables.dev;
    const unsigned grid = static_cast<unsigned>(e->W);
    const int rb = e->tables.real_bytes;
    bool done = false;
    const int fam = sweep_family(e);
    if (fam == kFamRegister)
        done = rb == 4 ? rr_parts_f32(d, e->st, grid, d_parts, s) : rr_parts_f64(d, e->st, grid, d_parts, s);
    else if (fam == kFamDelayed)
        done = rrw_measure64(d, e->st, grid, 1, d_parts, s);
    else if (fam == kFamMultiwave)
        done = rb == 4 ? rrw_parts_exact_f32(d, e->st, grid, d_parts, e->inv64, s, ReinvLaunch{rb})
                       : rrw_parts_f64(d, e->st, grid, d_parts, s);
    if (!done && rb == 4) {
        // Mixed precision, generic kernel: FP64 re-inversion of every walker
        // into inv64 (the walkers' state untouched), then the FP64 local
        // energy from it, as a measured sweep takes it.
        DevState st_refresh = e->st, st64 = e->st;
        st_refresh.inv64 = e->inv64;
        st_refresh.inv64_only = 1;
        st64.inv = e->inv64;
        reinv_generic_launch(d, st_refresh, grid, s, rb);
        hipLaunchKernelGGL((k_measure_parts<double, false>), dim3(grid), dim3(static_cast<unsigned>(e->walker_threads)),
                           lds_for(d, 8, false), s, d, st64, d_parts, 1);
        done = true;
    }
    if (!done) {
        const size_t lds_in = lds_for(d, rb, true);
        const bool inv_lds = lds_in <= kMaxLdsBytes && e->inv_placement != 2;
        const size_t lds = inv_lds ? lds_in : lds_for(d, rb, false);
        const dim3 block(static_cast<unsigned>(e->walker_threads));
        if (rb == 8) {
            if (inv_lds) hipLaunchKernelGGL((k_measure_parts<double, true>), dim3(grid), block, lds, s, d, e->st, d_parts, 1);
            else hipLaunchKernelGGL((k_measure_parts<double, false>), dim3(grid), block, lds, s, d, e->st, d_parts, 1);
        } else {
            if (inv_lds) hipLaunchKernelGGL((k_measure_parts<float, true>), dim3(grid), block, lds, s, d, e->st, d_parts, 1);
            else hipLaunchKernelGGL((k_measure_parts<float, false>), dim3(grid), block, lds, s, d, e->st, d_parts, 1);
        }
    }
    VMC_HIP_TRY(hipGetLastError());
    return VMC_OK;
}

// Staging of vmc_ensemble_measure_noalloc: the parts of up to
// kNoAllocWalkers walkers, one copy per device for the life of the process.
constexpr int64_t kNoAllocWalkers = VMC_NOALLOC_MAX_WALKERS;
static __device__ __align__(256) double scratch[kNoAllocWalkers * kPartCount];

// The device address of `scratch` on each device, looked up once per device
// with that device current (the array's host-side name is no device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(double **out) {
    int dev = -1;
    VMC_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(VMC_ERR_INVALID, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        VMC_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(VMC_ERR_DEVICE, "no device address for the measurement staging array");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return VMC_OK;
}

}  // namespace

extern "C" {

int vmc_ensemble_measure(vmc_ensemble *e, double *parts, void *stream) {
    if (!e || !parts) return fail(VMC_ERR_INVALID, "NULL argument");
    if (!e->positions_ready) return fail(VMC_ERR_STATE, "positions not initialised");
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *d_parts = nullptr;
    const size_t n = static_cast<size_t>(e->W) * kPartCount;
    int rc;
    if ((rc = dev_alloc(&d_parts, n))) return rc;
    struct Guard {
        double *p;
        ~Guard() { (void)hipFree(p); }
    } guard{d_parts};
    if ((rc = launch_parts(e, d_parts, s))) return rc;
    VMC_HIP_TRY(hipMemcpyAsync(parts, d_parts, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    return VMC_OK;
}

int vmc_ensemble_measure_noalloc(vmc_ensemble *e, double *parts, void *stream) {
    if (!e || !parts) return fail(VMC_ERR_INVALID, "NULL argument");
    if (e->W > kNoAllocWalkers)
        return fail(VMC_ERR_INVALID, "vmc_ensemble_measure_noalloc takes at most %lld walkers, ensemble has %lld",
                    static_cast<long long>(kNoAllocWalkers), static_cast<long long>(e->W));
    if (!e->positions_ready) return fail(VMC_ERR_STATE, "positions not initialised");
    double *d_parts = nullptr;
    int rc;
    if ((rc = scratch_for_current_device(&d_parts))) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((rc = launch_parts(e, d_parts, s))) return rc;
    VMC_HIP_TRY(hipMemcpyAsync(parts, d_parts, sizeof(double) * static_cast<size_t>(e->W) * kPartCount,
                               hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    return VMC_OK;
}

int vmc_ensemble_get_inverse(vmc_ensemble *e, double *inv, void *stream) {
    if (!e || !inv) return fail(VMC_ERR_INVALID, "NULL argument");
    const DevSys &d = e->tables.dev;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t n = static_cast<size_t>(e->W) * static_cast<size_t>(d.ninv);
    if (e->tables.real_bytes == 8) {
        VMC_HIP_TRY(hipMemcpyAsync(inv, e->st.inv, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        VMC_HIP_TRY(hipStreamSynchronize(s));
        return VMC_OK;
    }
    std::vector<float> tmp(n);
    VMC_HIP_TRY(hipMemcpyAsync(tmp.data(), e->st.inv, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i) inv[i] = static_cast<double>(tmp[i]);
    return VMC_OK;
}

int vmc_ensemble_set_kernel(vmc_ensemble *e, int32_t kernel) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    if (kernel != VMC_KERNEL_AUTO && kernel != VMC_KERNEL_GENERIC && kernel != VMC_KERNEL_REGISTER &&
        kernel != VMC_KERNEL_DELAYED)
        return fail(VMC_ERR_INVALID, "unknown kernel %d", kernel);
    if (kernel == VMC_KERNEL_REGISTER && !rr_supported(e->tables.dev) && !rrw_supported(e->tables.dev, e->tables.real_bytes))
        return fail(VMC_ERR_INVALID, "register kernel needs 7, 19, 27, 33, 57 or 81 electrons per spin (FP32 also 123, 147)");
    if (kernel == VMC_KERNEL_DELAYED && (e->tables.real_bytes != 8 || !du_supported(e->tables.dev)))
        return fail(VMC_ERR_INVALID, "delayed-update kernel needs an FP64 system of at most 160 electrons per spin");
    e->kernel = kernel;
    return VMC_OK;
}

int vmc_ensemble_set_walker_groups(vmc_ensemble *e, int32_t groups) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    if (groups < 1 || groups > kMaxWalkerGroups)
        return fail(VMC_ERR_INVALID, "walker groups must be in [1, %d]", kMaxWalkerGroups);
    e->groups = groups;
    return VMC_OK;
}

int vmc_ensemble_set_rr_variant(vmc_ensemble *e, int32_t variant) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    if (variant != 0 && variant != 1) return fail(VMC_ERR_INVALID, "unknown variant %d", variant);
    e->rr_variant = variant;
    return VMC_OK;
}

int vmc_ensemble_set_inverse_placement(vmc_ensemble *e, int32_t placement) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    if (placement < 0 || placement > 2) return fail(VMC_ERR_INVALID, "unknown placement %d", placement);
    if (placement == 1 && lds_for(e->tables.dev, e->tables.real_bytes, true) > kMaxLdsBytes)
        return fail(VMC_ERR_INVALID, "inverse does not fit in LDS");
    e->inv_placement = placement;
    return VMC_OK;
}

int vmc_ensemble_set_walker_threads(vmc_ensemble *e, int32_t threads) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    if (threads != 64 && threads != 128 && threads != 256)
        return fail(VMC_ERR_INVALID, "threads per walker must be 64, 128 or 256");
    e->walker_threads = threads;
    return VMC_OK;
}

int vmc_ensemble_get_kernel(const vmc_ensemble *e, int32_t *kernel) {
    if (!e || !kernel) return fail(VMC_ERR_INVALID, "NULL argument");
    const int fam = sweep_family(e);
    *kernel = fam == kFamRegister || fam == kFamMultiwave ? VMC_KERNEL_REGISTER
              : fam == kFamDelayed                        ? VMC_KERNEL_DELAYED
                                                          : VMC_KERNEL_GENERIC;
    return VMC_OK;
}

int vmc_ensemble_reset_stats(vmc_ensemble *e, void *stream) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    VMC_HIP_TRY(hipMemsetAsync(e->st.stats, 0, sizeof(double) * e->W * kStatCount, static_cast<hipStream_t>(stream)));
    return VMC_OK;
}

int vmc_ensemble_read_stats(vmc_ensemble *e, double *stats, void *stream) {
    if (!e || !stats) return fail(VMC_ERR_INVALID, "NULL argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    VMC_HIP_TRY(hipMemcpyAsync(stats, e->st.stats, sizeof(double) * e->W * kStatCount, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    return VMC_OK;
}

int vmc_ensemble_set_timing(vmc_ensemble *e, int32_t enable) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    e->timer.on = enable != 0;
    return VMC_OK;
}

int vmc_ensemble_read_timing(vmc_ensemble *e, double *ms, int64_t *launches) {
    if (!e || !ms || !launches) return fail(VMC_ERR_INVALID, "NULL argument");
    for (int k = 0; k < kTimeKinds; ++k) {
        ms[k] = 0.0;
        launches[k] = 0;
    }
    KernelTimer &t = e->timer;
    int rc = VMC_OK;
    for (size_t i = 0; i < t.start.size(); ++i) {
        float dt = 0.0f;
        if (hipEventSynchronize(t.stop[i]) != hipSuccess ||
            hipEventElapsedTime(&dt, t.start[i], t.stop[i]) != hipSuccess)
            rc = fail(VMC_ERR_DEVICE, "kernel timing event failed");
        ms[t.kind[i]] += dt;
        launches[t.kind[i]] += 1;
        t.pool.push_back(t.start[i]);
        t.pool.push_back(t.stop[i]);
    }
    t.start.clear();
    t.stop.clear();
    t.kind.clear();
    return rc;
}

int vmc_ensemble_read_drift(vmc_ensemble *e, double *drift, void *stream) {
    if (!e || !drift) return fail(VMC_ERR_INVALID, "NULL argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    VMC_HIP_TRY(hipMemcpyAsync(drift, e->st.drift, sizeof(double) * e->W, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    return VMC_OK;
}

int vmc_ensemble_read_steps(vmc_ensemble *e, double *steps, void *stream) {
    if (!e || !steps) return fail(VMC_ERR_INVALID, "NULL argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    VMC_HIP_TRY(hipMemcpyAsync(steps, e->st.sigma, sizeof(double) * e->W, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    return VMC_OK;
}

int vmc_ensemble_read_frame(vmc_ensemble *e, int64_t walker, double *stats, double *pos,
                            void *stream) {
    if (!e || (!stats && !pos)) return fail(VMC_ERR_INVALID, "NULL argument");
    if (walker < 0 || walker >= e->W) return fail(VMC_ERR_INVALID, "walker %lld out of range", (long long)walker);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int N = e->tables.dev.N;
    const size_t w = static_cast<size_t>(walker);
    std::vector<double> soa(pos ? 3 * static_cast<size_t>(N) : 0);
    if (stats)
        VMC_HIP_TRY(hipMemcpyAsync(stats, e->st.stats + w * kStatCount, sizeof(double) * kStatCount,
                                   hipMemcpyDeviceToHost, s));
    if (pos)
        VMC_HIP_TRY(hipMemcpyAsync(soa.data(), e->st.pos + w * 3 * N, sizeof(double) * soa.size(),
                                   hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    if (pos)
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) pos[3 * i + c] = soa[static_cast<size_t>(c) * N + i];
    return VMC_OK;
}

int vmc_ensemble_info(const vmc_ensemble *e, int64_t *n_walkers, uint64_t *sweeps_done) {
    if (!e) return fail(VMC_ERR_INVALID, "ensemble is NULL");
    if (n_walkers) *n_walkers = e->W;
    if (sweeps_done) *sweeps_done = e->sweeps_done;
    return VMC_OK;
}

int vmc_frozen_eval(const vmc_system *sys, int32_t n_configs, const double *pos,
                    const int32_t *moved, const double *newpos, double *scalars, double *grad_lap,
                    double *inv_before, double *inv_after, double *s_k, void *stream) {
    if (n_configs < 1 || !pos || !moved || !newpos || !scalars) return fail(VMC_ERR_INVALID, "bad frozen-eval arguments");
    SystemTables t;
    int rc = build_tables(sys, 0, t);
    if (rc) return rc;
    const DevSys &d = t.dev;
    const size_t C = static_cast<size_t>(n_configs);
    for (size_t c = 0; c < C; ++c) {
        if (moved[c] < 0 || moved[c] >= d.N) return fail(VMC_ERR_INVALID, "moved index out of range");
        for (int k = 0; k < 3; ++k)
            if (!(newpos[3 * c + k] >= 0.0 && newpos[3 * c + k] < d.L)) return fail(VMC_ERR_INVALID, "newpos outside [0, L)");
        for (int i = 0; i < 3 * d.N; ++i)
            if (!(pos[c * 3 * d.N + i] >= 0.0 && pos[c * 3 * d.N + i] < d.L)) return fail(VMC_ERR_INVALID, "position outside [0, L)");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *d_pos = nullptr, *d_new = nullptr, *d_sc = nullptr, *d_gl = nullptr, *d_ib = nullptr,
           *d_ia = nullptr, *d_sk = nullptr, *d_A = nullptr, *d_skc = nullptr;
    int *d_mv = nullptr;
    unsigned char *d_inv = nullptr;
    struct Guard {
        std::vector<void *> p;
        ~Guard() { for (void *q : p) if (q) (void)hipFree(q); }
    } guard;
    const size_t rb = static_cast<size_t>(t.real_bytes);
    if ((rc = dev_alloc(&d_pos, C * 3 * d.N)) || (rc = dev_alloc(&d_new, C * 3)) ||
        (rc = dev_alloc(&d_mv, C)) || (rc = dev_alloc(&d_sc, C * VMC_FZ_COUNT)) ||
        (rc = dev_alloc(&d_gl, C * 4 * d.N)) || (rc = dev_alloc(&d_ib, C * d.ninv)) ||
        (rc = dev_alloc(&d_ia, C * d.ninv)) || (rc = dev_alloc(&d_sk, C * 2 * std::max(d.ng, 1))) ||
        (rc = dev_alloc(&d_A, C * d.nmax * d.nmax)) || (rc = dev_alloc(&d_skc, C * 2 * std::max(d.ng, 1))) ||
        (rc = dev_alloc(&d_inv, C * d.ninv * rb))) {
        guard.p = {d_pos, d_new, d_mv, d_sc, d_gl, d_ib, d_ia, d_sk, d_A, d_skc, d_inv};
        return rc;
    }
    guard.p = {d_pos, d_new, d_mv, d_sc, d_gl, d_ib, d_ia, d_sk, d_A, d_skc, d_inv};
    VMC_HIP_TRY(hipMemcpyAsync(d_pos, pos, sizeof(double) * C * 3 * d.N, hipMemcpyHostToDevice, s));
    VMC_HIP_TRY(hipMemcpyAsync(d_new, newpos, sizeof(double) * C * 3, hipMemcpyHostToDevice, s));
    VMC_HIP_TRY(hipMemcpyAsync(d_mv, moved, sizeof(int) * C, hipMemcpyHostToDevice, s));
    const bool inv_lds = lds_for(d, t.real_bytes, true) <= kMaxLdsBytes;
    const size_t lds = lds_for(d, t.real_bytes, inv_lds);
    const dim3 grid(static_cast<unsigned>(C)), block(kWave);
    if (t.real_bytes == 8) {
        if (inv_lds)
            hipLaunchKernelGGL((k_frozen<double, true>), grid, block, lds, s, d, n_configs, d_pos, d_mv, d_new, d_sc, d_gl, d_ib, d_ia, d_sk, reinterpret_cast<double *>(d_inv), d_A, d_skc);
        else
            hipLaunchKernelGGL((k_frozen<double, false>), grid, block, lds, s, d, n_configs, d_pos, d_mv, d_new, d_sc, d_gl, d_ib, d_ia, d_sk, reinterpret_cast<double *>(d_inv), d_A, d_skc);
    } else {
        if (inv_lds)
            hipLaunchKernelGGL((k_frozen<float, true>), grid, block, lds, s, d, n_configs, d_pos, d_mv, d_new, d_sc, d_gl, d_ib, d_ia, d_sk, reinterpret_cast<float *>(d_inv), d_A, d_skc);
        else
            hipLaunchKernelGGL((k_frozen<float, false>), grid, block, lds, s, d, n_configs, d_pos, d_mv, d_new, d_sc, d_gl, d_ib, d_ia, d_sk, reinterpret_cast<float *>(d_inv), d_A, d_skc);
    }
    VMC_HIP_TRY(hipGetLastError());
    VMC_HIP_TRY(hipMemcpyAsync(scalars, d_sc, sizeof(double) * C * VMC_FZ_COUNT, hipMemcpyDeviceToHost, s));
    if (t.real_bytes == 4) {
        // Mixed precision measures in FP64 (as the sweeps' local energy
        // does): log|det|, J and the energies before and after the move, the
        // derivatives and S_G come from the FP64 evaluation of the same
        // configurations; ratio, delta J, the inverses and the incremental
        // Ewald updates stay the FP32 per-move results.
        double *d_sc64 = nullptr, *d_inv64 = nullptr;
        if ((rc = dev_alloc(&d_sc64, C * VMC_FZ_COUNT)) || (rc = dev_alloc(&d_inv64, C * d.ninv))) {
            guard.p.push_back(d_sc64);
            return rc;
        }
        guard.p.push_back(d_sc64);
        guard.p.push_back(d_inv64);
        const bool lds64_in = lds_for(d, 8, true) <= kMaxLdsBytes;
        const size_t lds64 = lds_for(d, 8, lds64_in);
        if (lds64_in)
            hipLaunchKernelGGL((k_frozen<double, true>), grid, block, lds64, s, d, n_configs, d_pos, d_mv, d_new, d_sc64, d_gl, nullptr, nullptr, d_sk, d_inv64, d_A, d_skc);
        else
            hipLaunchKernelGGL((k_frozen<double, false>), grid, block, lds64, s, d, n_configs, d_pos, d_mv, d_new, d_sc64, d_gl, nullptr, nullptr, d_sk, d_inv64, d_A, d_skc);
        VMC_HIP_TRY(hipGetLastError());
        std::vector<double> sc64(C * VMC_FZ_COUNT);
        VMC_HIP_TRY(hipMemcpyAsync(sc64.data(), d_sc64, sizeof(double) * C * VMC_FZ_COUNT, hipMemcpyDeviceToHost, s));
        VMC_HIP_TRY(hipStreamSynchronize(s));
        for (size_t c = 0; c < C; ++c)
            for (int k : {VMC_FZ_LOGDET0, VMC_FZ_LOGDET1, VMC_FZ_JASTROW, VMC_FZ_KINETIC, VMC_FZ_VREAL, VMC_FZ_VRECIP,
                          VMC_FZ_ELOCAL, VMC_FZ_ELOCAL_AFTER, VMC_FZ_KINETIC_AFTER})
                scalars[c * VMC_FZ_COUNT + k] = sc64[c * VMC_FZ_COUNT + k];
    }
    if (grad_lap) VMC_HIP_TRY(hipMemcpyAsync(grad_lap, d_gl, sizeof(double) * C * 4 * d.N, hipMemcpyDeviceToHost, s));
    if (inv_before) VMC_HIP_TRY(hipMemcpyAsync(inv_before, d_ib, sizeof(double) * C * d.ninv, hipMemcpyDeviceToHost, s));
    if (inv_after) VMC_HIP_TRY(hipMemcpyAsync(inv_after, d_ia, sizeof(double) * C * d.ninv, hipMemcpyDeviceToHost, s));
    if (s_k) VMC_HIP_TRY(hipMemcpyAsync(s_k, d_sk, sizeof(double) * C * 2 * d.ng, hipMemcpyDeviceToHost, s));
    VMC_HIP_TRY(hipStreamSynchronize(s));
    return VMC_OK;
}

// ------------------------------------------------------------------ batch

}  // extern "C"

// A batch group: systems swept by one kernel family and, for the register
// families, of one size, so that one launch per kernel covers them all.
struct BatchGroup {
    int family = 0;                  // 0 generic (persistent), 1 register, 2 multi-wavefront register
    DevSys shape{};                  // a member's tables (sizes shared by a register group)
    std::vector<int> members;        // indices into vmc_batch::members
    BatchItem *items = nullptr;      // device copy; walker ranges numbered within the group
    int total = 0;                   // walkers of the group
    bool reinv = false;              // some member re-inverts periodically
    unsigned int *queue = nullptr;   // generic: the persistent kernel's walker queue
    int grid = 0;                    // workgroups of the group's sweep launch
    size_t lds = 0;                  // generic: dynamic LDS of the sweep
    double cost = 0.0;               // sum of W N^2 (launch order, largest first)
    hipStream_t lane = nullptr;      // own stream when groups overlap (not the first group)
    hipEvent_t join = nullptr;
    // generic: the re-inverting members by determinant order, each subset
    // re-inverted by the kernel an ensemble of that order uses alone
    struct Reinv {
        DevSys shape;
        BatchItem *items;
        int n_items, total;
    };
    std::vector<Reinv> reinv_sets;
    // Mixed precision, multi-wavefront and generic groups: the items with
    // st.inv64 set (a measured sweep's re-inversion of every walker, FP64
    // result kept) and with st.inv = inv64 (that sweep's FP64 local energy);
    // generic: the re-inversion sets over every member, and the FP64 local
    // energy's LDS.
    BatchItem *items_refresh = nullptr, *items_m64 = nullptr;
    std::vector<Reinv> refresh_sets;
    size_t lds64 = 0;

    void release() {
        for (auto &r : reinv_sets)
            if (r.items) (void)hipFree(r.items);
        reinv_sets.clear();
        for (auto &r : refresh_sets)
            if (r.items) (void)hipFree(r.items);
        refresh_sets.clear();
        if (items_refresh) (void)hipFree(items_refresh);
        if (items_m64) (void)hipFree(items_m64);
        items_refresh = items_m64 = nullptr;
        if (items) (void)hipFree(items);
        if (queue) (void)hipFree(queue);
        if (lane) (void)hipStreamDestroy(lane);
        if (join) (void)hipEventDestroy(join);
        items = nullptr;
        queue = nullptr;
        lane = nullptr;
        join = nullptr;
    }
};

struct vmc_batch {
    std::vector<vmc_ensemble *> members;
    std::vector<int> placement;      // per member: generic kernel's inverse placement (0 auto, 1 LDS, 2 HBM)
    std::vector<BatchGroup> groups;
    unsigned long long *ctr = nullptr;
    int64_t total = 0;
    int threads = 64;       // threads per walker of the generic persistent sweep
    int real_bytes = 8;
    bool overlap = true;    // groups on their own streams, joined after each sweep
    hipEvent_t fork = nullptr;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_sweeps = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_sweeps = 0;
    }
    void drop_groups() {
        for (auto &g : groups) g.release();
        groups.clear();
    }
    ~vmc_batch() {
        drop_graph();
        drop_groups();
        if (fork) (void)hipEventDestroy(fork);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (ctr) (void)hipFree(ctr);
        for (auto *m : members) delete m;
    }
};

namespace {

// Kernel family of a batch member, by the rule the ensemble's own sweep
// uses: register-resident when an instantiation exists and the member does
// not ask for the generic kernel.
int batch_family(const vmc_ensemble *e) {
    const DevSys &d = e->tables.dev;
    if (e->kernel == VMC_KERNEL_GENERIC) return 0;
    if (rr_supported(d)) return 1;
    if (rrw_supported_sizes(d.ndet, d.n_up, d.n_dn, d.ng, e->tables.real_bytes)) return 2;
    return 0;
}

// (Re)builds the batch's groups from its members' kernels.  The caller has
// synchronised the device (old group buffers may be in use otherwise).
int batch_build_groups(vmc_batch *b) {
    b->drop_graph();
    b->drop_groups();
    for (size_t i = 0; i < b->members.size(); ++i) {
        const vmc_ensemble *e = b->members[i];
        const DevSys &d = e->tables.dev;
        const int fam = batch_family(e);
        BatchGroup *g = nullptr;
        for (auto &h : b->groups)
            if (h.family == fam && (fam == 0 || (h.shape.ndet == d.ndet && h.shape.n_up == d.n_up &&
                                                 h.shape.n_dn == d.n_dn && h.shape.nmax == d.nmax)))
                g = &h;
        if (!g) {
            b->groups.emplace_back();
            g = &b->groups.back();
            g->family = fam;
            g->shape = d;
        }
        g->members.push_back(static_cast<int>(i));
        g->cost += static_cast<double>(e->W) * d.N * d.N;
    }
    std::stable_sort(b->groups.begin(), b->groups.end(),
                     [](const BatchGroup &x, const BatchGroup &y) { return x.cost > y.cost; });
    int dev = 0, cus = 0;
    VMC_HIP_TRY(hipGetDevice(&dev));
    VMC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        BatchGroup &g = b->groups[gi];
        std::vector<BatchItem> items(g.members.size());
        int64_t total = 0;
        for (size_t k = 0; k < g.members.size(); ++k) {
            const int i = g.members[k];
            const vmc_ensemble *e = b->members[static_cast<size_t>(i)];
            const DevSys &d = e->tables.dev;
            items[k].s = d;
            items[k].st = e->st;
            items[k].walker_begin = static_cast<int>(total);
            total += e->W;
            items[k].walker_end = static_cast<int>(total);
            items[k].inv_lds = 0;
            if (d.reinv_k > 0) g.reinv = true;
            if (g.family == 0) {
                const bool fits = lds_for(d, e->tables.real_bytes, true) <= kMaxLdsBytes;
                const bool inv_lds = b->placement[static_cast<size_t>(i)] == 2 ? false : fits;
                items[k].inv_lds = inv_lds ? 1 : 0;
                g.lds = std::max(g.lds, lds_for(d, e->tables.real_bytes, inv_lds));
            }
        }
        g.total = static_cast<int>(total);
        int rc;
        if ((rc = dev_alloc(&g.items, items.size()))) return rc;
        VMC_HIP_TRY(hipMemcpy(g.items, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice));
        if (b->real_bytes == 4 && g.family != 1) {
            std::vector<BatchItem> ir = items, i64 = items;
            for (size_t k = 0; k < items.size(); ++k) {
                const vmc_ensemble *e = b->members[static_cast<size_t>(g.members[k])];
                ir[k].st.inv64 = e->inv64;
                i64[k].st.inv = e->inv64;
                if (g.family == 0) g.lds64 = std::max(g.lds64, lds_for(items[k].s, 8, false));
            }
            if ((rc = dev_alloc(&g.items_refresh, ir.size())) || (rc = dev_alloc(&g.items_m64, i64.size()))) return rc;
            VMC_HIP_TRY(hipMemcpy(g.items_refresh, ir.data(), sizeof(BatchItem) * ir.size(), hipMemcpyHostToDevice));
            VMC_HIP_TRY(hipMemcpy(g.items_m64, i64.data(), sizeof(BatchItem) * i64.size(), hipMemcpyHostToDevice));
            if (g.family == 0) {
                std::vector<std::pair<int, int>> orders;
                for (const auto &it : ir) {
                    const std::pair<int, int> key{it.s.nmax, it.s.ndet};
                    if (std::find(orders.begin(), orders.end(), key) == orders.end()) orders.push_back(key);
                }
                for (const auto &[nm, nd] : orders) {
                    std::vector<BatchItem> sub;
                    int64_t t = 0;
                    for (const auto &it : ir) {
                        if (it.s.nmax != nm || it.s.ndet != nd) continue;
                        BatchItem x = it;
                        const int w = it.walker_end - it.walker_begin;
                        x.walker_begin = static_cast<int>(t);
                        t += w;
                        x.walker_end = static_cast<int>(t);
                        sub.push_back(x);
                    }
                    BatchGroup::Reinv r{sub[0].s, nullptr, static_cast<int>(sub.size()), static_cast<int>(t)};
                    if ((rc = dev_alloc(&r.items, sub.size()))) return rc;
                    g.refresh_sets.push_back(r);
                    VMC_HIP_TRY(hipMemcpy(r.items, sub.data(), sizeof(BatchItem) * sub.size(), hipMemcpyHostToDevice));
                }
            }
        }
        if (g.family == 0) {
            // One re-inversion set per (order, determinant count): a launch
            // takes its determinant count from its set's first member.
            std::vector<std::pair<int, int>> orders;
            for (const auto &it : items) {
                const std::pair<int, int> key{it.s.nmax, it.s.ndet};
                if (it.s.reinv_k > 0 && std::find(orders.begin(), orders.end(), key) == orders.end())
                    orders.push_back(key);
            }
            for (const auto &[nm, nd] : orders) {
                std::vector<BatchItem> sub;
                int64_t t = 0;
                for (const auto &it : items) {
                    if (it.s.reinv_k <= 0 || it.s.nmax != nm || it.s.ndet != nd) continue;
                    BatchItem x = it;
                    const int w = it.walker_end - it.walker_begin;
                    x.walker_begin = static_cast<int>(t);
                    t += w;
                    x.walker_end = static_cast<int>(t);
                    sub.push_back(x);
                }
                BatchGroup::Reinv r{sub[0].s, nullptr, static_cast<int>(sub.size()), static_cast<int>(t)};
                if ((rc = dev_alloc(&r.items, sub.size()))) return rc;
                g.reinv_sets.push_back(r);
                VMC_HIP_TRY(hipMemcpy(r.items, sub.data(), sizeof(BatchItem) * sub.size(), hipMemcpyHostToDevice));
            }
            if ((rc = dev_alloc(&g.queue, 1))) return rc;
            VMC_HIP_TRY(hipMemset(g.queue, 0, sizeof(unsigned int)));
            // Persistent grid: as many workgroups as fit on the device at once.
            int per_cu = 0;
            const hipError_t oc =
                b->real_bytes == 8
                    ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_batch<double>, b->threads, g.lds)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_batch<float>, b->threads, g.lds);
            if (oc != hipSuccess || per_cu < 1) per_cu = 1;
            g.grid = static_cast<int>(std::min<int64_t>(total, static_cast<int64_t>(cus) * per_cu));
        } else {
            g.grid = g.total;
        }
        if (gi > 0) {
            VMC_HIP_TRY(hipStreamCreateWithFlags(&g.lane, hipStreamNonBlocking));
            VMC_HIP_TRY(hipEventCreateWithFlags(&g.join, hipEventDisableTiming));
        }
    }
    return VMC_OK;
}

// One sweep of one group on `stream`.
template <typename Real>
int batch_group_step(vmc_batch *b, const BatchGroup &g, int flags, int block_size, hipStream_t stream) {
    const int n = static_cast<int>(g.members.size());
    constexpr bool f32 = sizeof(Real) == 4;
    bool ok = true;
    // Mixed precision: a measured sweep re-inverts every walker with an FP64
    // result and takes the FP64 local energy from it (as alone).
    const bool exact = f32 && (flags & VMC_SWEEP_MEASURE) && g.items_m64;
    if (g.family == 0) {
        hipLaunchKernelGGL((k_sweep_batch<Real>), dim3(static_cast<unsigned>(g.grid)), dim3(b->threads), g.lds,
                           stream, g.items, n, g.total, g.queue, b->ctr, exact ? flags & ~VMC_SWEEP_MEASURE : flags,
                           block_size);
        for (const auto &r : exact ? g.refresh_sets : g.reinv_sets)
            reinv_batch_launch(r.shape, r.items, r.n_items, static_cast<unsigned>(r.total), stream,
                               static_cast<int>(sizeof(Real)));
        if (exact)
            hipLaunchKernelGGL(k_measure_batch64, dim3(static_cast<unsigned>(g.total)), dim3(b->threads), g.lds64,
                               stream, g.items_m64, n, block_size);
    } else if (g.family == 1) {
        ok = f32 ? rr_batch_step_f32(g.shape, g.items, n, static_cast<unsigned>(g.total), b->ctr, flags, block_size,
                                     g.reinv, stream)
                 : rr_batch_step_f64(g.shape, g.items, n, static_cast<unsigned>(g.total), b->ctr, flags, block_size,
                                     g.reinv, stream);
    } else {
        ok = f32 ? rrw_batch_step_f32(g.shape, g.items, n, static_cast<unsigned>(g.total), b->ctr, flags,
                                      block_size, g.reinv, stream, g.items_refresh, g.items_m64)
                 : rrw_batch_step_f64(g.shape, g.items, n, static_cast<unsigned>(g.total), b->ctr, flags,
                                      block_size, g.reinv, stream, nullptr, nullptr);
    }
    if (!ok) return fail(VMC_ERR_INVALID, "no register-resident instantiation for a batch group");
    return VMC_OK;
}

// n_sweeps sweeps of every group.  With overlap, the groups after the first
// run on their own streams, forked from and joined back into `stream` around
// each sweep (also under graph capture, where they become parallel branches).
template <typename Real>
int batch_launch(vmc_batch *b, int n_sweeps, int flags, int block_size, hipStream_t stream) {
    const bool fan = b->overlap && b->groups.size() > 1;
    unsigned int *queue = nullptr;
    for (const auto &g : b->groups)
        if (g.family == 0) queue = g.queue;
    for (int i = 0; i < n_sweeps; ++i) {
        if (fan) VMC_HIP_TRY(hipEventRecord(b->fork, stream));
        for (size_t gi = 0; gi < b->groups.size(); ++gi) {
            const BatchGroup &g = b->groups[gi];
            const hipStream_t s = (fan && gi > 0) ? g.lane : stream;
            if (s != stream) VMC_HIP_TRY(hipStreamWaitEvent(s, b->fork, 0));
            const int rc = batch_group_step<Real>(b, g, flags, block_size, s);
            if (rc) return rc;
            if (s != stream) VMC_HIP_TRY(hipEventRecord(g.join, s));
        }
        if (fan)
            for (size_t gi = 1; gi < b->groups.size(); ++gi)
                VMC_HIP_TRY(hipStreamWaitEvent(stream, b->groups[gi].join, 0));
        hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(1), 0, stream, queue, b->ctr);
    }
    VMC_HIP_TRY(hipGetLastError());
    return VMC_OK;
}

int batch_sweep_any(vmc_batch *b, int n_sweeps, int flags, int block_size, hipStream_t stream) {
    return b->real_bytes == 8 ? batch_launch<double>(b, n_sweeps, flags, block_size, stream)
                              : batch_launch<float>(b, n_sweeps, flags, block_size, stream);
}

}  // namespace

extern "C" {

int vmc_batch_create(const vmc_system *systems, const int64_t *n_walkers, const uint64_t *seeds,
                     const double *step0, const int32_t *inv_placement, int32_t n_systems,
                     int32_t walker_threads, vmc_batch **out) {
    if (!out || !systems || !n_walkers || !seeds || !step0) return fail(VMC_ERR_INVALID, "NULL argument");
    if (walker_threads != 64 && walker_threads != 128 && walker_threads != 256)
        return fail(VMC_ERR_INVALID, "walker_threads must be 64, 128 or 256");
    *out = nullptr;
    if (n_systems < 1 || n_systems > 4096) return fail(VMC_ERR_INVALID, "n_systems out of range");
    for (int i = 1; i < n_systems; ++i)
        if ((systems[i].precision != VMC_PREC_FP64) != (systems[0].precision != VMC_PREC_FP64))
            return fail(VMC_ERR_INVALID, "a batch runs one working precision");
    auto *b = new vmc_batch();
    b->threads = walker_threads;
    int64_t total = 0;
    for (int i = 0; i < n_systems; ++i) {
        vmc_ensemble *e = nullptr;
        int rc = vmc_ensemble_create(&systems[i], n_walkers[i], seeds[i], step0[i], &e);
        if (rc) {
            delete b;
            return rc;
        }
        e->in_batch = true;
        b->members.push_back(e);
        b->real_bytes = e->tables.real_bytes;
        const int place = inv_placement ? inv_placement[i] : 0;  // 0 auto, 1 LDS, 2 HBM
        if (place < 0 || place > 2) {
            delete b;
            return fail(VMC_ERR_INVALID, "system %d: inverse placement must be 0, 1 or 2", i);
        }
        if (place == 1 && lds_for(e->tables.dev, e->tables.real_bytes, true) > kMaxLdsBytes) {
            delete b;
            return fail(VMC_ERR_INVALID, "system %d: inverse does not fit in LDS", i);
        }
        b->placement.push_back(place);
        total += n_walkers[i];
    }
    if (total > (int64_t(1) << 30)) {
        delete b;
        return fail(VMC_ERR_INVALID, "too many walkers in one batch");
    }
    b->total = total;
    int rc;
    if ((rc = dev_alloc(&b->ctr, 1))) {
        delete b;
        return rc;
    }
    if (hipMemset(b->ctr, 0, sizeof(unsigned long long)) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return fail(VMC_ERR_DEVICE, "batch initialisation failed");
    }
    if ((rc = batch_build_groups(b))) {
        delete b;
        return rc;
    }
    *out = b;
    return VMC_OK;
}

int vmc_batch_set_kernel(vmc_batch *b, int32_t system, int32_t kernel) {
    if (!b) return fail(VMC_ERR_INVALID, "batch is NULL");
    if (kernel != VMC_KERNEL_AUTO && kernel != VMC_KERNEL_GENERIC && kernel != VMC_KERNEL_REGISTER)
        return fail(VMC_ERR_INVALID, "unknown kernel %d", kernel);
    if (system < -1 || system >= static_cast<int32_t>(b->members.size()))
        return fail(VMC_ERR_INVALID, "system %d out of range", system);
    for (size_t i = 0; i < b->members.size(); ++i) {
        if (system >= 0 && static_cast<int32_t>(i) != system) continue;
        vmc_ensemble *e = b->members[i];
        if (kernel == VMC_KERNEL_REGISTER && batch_family(e) == 0 &&
            !rrw_supported_sizes(e->tables.dev.ndet, e->tables.dev.n_up, e->tables.dev.n_dn, e->tables.dev.ng,
                                 e->tables.real_bytes) &&
            !rr_supported(e->tables.dev))
            return fail(VMC_ERR_INVALID, "system %zu has no register-resident instantiation", i);
    }
    VMC_HIP_TRY(hipDeviceSynchronize());
    for (size_t i = 0; i < b->members.size(); ++i)
        if (system < 0 || static_cast<int32_t>(i) == system) b->members[i]->kernel = kernel;
    return batch_build_groups(b);
}

int vmc_batch_set_overlap(vmc_batch *b, int32_t overlap) {
    if (!b) return fail(VMC_ERR_INVALID, "batch is NULL");
    VMC_HIP_TRY(hipDeviceSynchronize());
    b->drop_graph();
    b->overlap = overlap != 0;
    return VMC_OK;
}

int vmc_batch_destroy(vmc_batch *b) {
    delete b;
    return VMC_OK;
}

int vmc_batch_init_positions(vmc_batch *b, void *stream) {
    if (!b) return fail(VMC_ERR_INVALID, "batch is NULL");
    for (size_t i = 0; i < b->members.size(); ++i) {
        const int rc = vmc_ensemble_init_positions(b->members[i], stream);
        if (rc) return rc;
    }
    return VMC_OK;
}

int vmc_batch_sweep(vmc_batch *b, int32_t n_sweeps, int32_t flags, int32_t block_size, void *stream) {
    if (!b) return fail(VMC_ERR_INVALID, "batch is NULL");
    if (n_sweeps < 0 || block_size < 1) return fail(VMC_ERR_INVALID, "bad sweep arguments");
    if (flags & ~(VMC_SWEEP_MEASURE | VMC_SWEEP_ADAPT)) return fail(VMC_ERR_INVALID, "unknown sweep flags");
    for (auto *m : b->members)
        if (!m->positions_ready) return fail(VMC_ERR_STATE, "positions not initialised");
    const int rc = batch_sweep_any(b, n_sweeps, flags, block_size, static_cast<hipStream_t>(stream));
    if (rc) return rc;
    for (auto *m : b->members) m->sweeps_done += static_cast<uint64_t>(n_sweeps);
    return VMC_OK;
}

int vmc_batch_capture(vmc_batch *b, int32_t n_sweeps, int32_t flags, int32_t block_size) {
    if (!b) return fail(VMC_ERR_INVALID, "batch is NULL");
    if (n_sweeps < 1 || block_size < 1) return fail(VMC_ERR_INVALID, "bad capture arguments");
    if (flags & ~(VMC_SWEEP_MEASURE | VMC_SWEEP_ADAPT)) return fail(VMC_ERR_INVALID, "unknown sweep flags");
    if (b->exec) {
        VMC_HIP_TRY(hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    if (b->graph) {
        VMC_HIP_TRY(hipGraphDestroy(b->graph));
        b->graph = nullptr;
    }
    if (!b->cap_stream) VMC_HIP_TRY(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    VMC_HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_sweep_any(b, n_sweeps, flags, block_size, b->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    VMC_HIP_TRY(ec);
    b->graph = g;
    VMC_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_sweeps = n_sweeps;
    return VMC_OK;
}

int vmc_batch_replay(vmc_batch *b, void *stream) {
    if (!b || !b->exec) return fail(VMC_ERR_STATE, "no captured batch graph");
    for (auto *m : b->members)
        if (!m->positions_ready) return fail(VMC_ERR_STATE, "positions not initialised");
    VMC_HIP_TRY(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)));
    for (auto *m : b->members) m->sweeps_done += static_cast<uint64_t>(b->graph_sweeps);
    return VMC_OK;
}

int vmc_batch_member(vmc_batch *b, int32_t system, vmc_ensemble **out) {
    if (!b || !out) return fail(VMC_ERR_INVALID, "NULL argument");
    if (system < 0 || system >= static_cast<int32_t>(b->members.size()))
        return fail(VMC_ERR_INVALID, "system %d out of range", system);
    *out = b->members[static_cast<size_t>(system)];
    return VMC_OK;
}

int vmc_batch_info(const vmc_batch *b, int32_t *n_systems, int64_t *total_walkers, int32_t *grid,
                   int64_t *lds_bytes) {
    if (!b) return fail(VMC_ERR_INVALID, "batch is NULL");
    int64_t wg = 0, lds = 0;
    for (const auto &g : b->groups) {
        wg += g.grid;
        lds = std::max<int64_t>(lds, static_cast<int64_t>(g.lds));
    }
    if (n_systems) *n_systems = static_cast<int32_t>(b->members.size());
    if (total_walkers) *total_walkers = b->total;
    if (grid) *grid = static_cast<int32_t>(wg);
    if (lds_bytes) *lds_bytes = lds;
    return VMC_OK;
}

int vmc_batch_groups(const vmc_batch *b, int32_t *n_groups, int32_t *family, int32_t *group_of, int32_t cap) {
    if (!b) return fail(VMC_ERR_INVALID, "batch is NULL");
    if (n_groups) *n_groups = static_cast<int32_t>(b->groups.size());
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        if (family && static_cast<int32_t>(gi) < cap) family[gi] = b->groups[gi].family;
        if (group_of)
            for (int m : b->groups[gi].members)
                if (m < cap) group_of[m] = static_cast<int32_t>(gi);
    }
    return VMC_OK;
}

}  // extern "C"
