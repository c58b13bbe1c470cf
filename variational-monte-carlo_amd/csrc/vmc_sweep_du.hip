// vmc_sweep_du.hip -- FP64 sweep with delayed (Woodbury) inverse updates for
// determinants of up to 160 electrons, the FP64 sizes above the
// register-resident kernels (123 and 147 electrons per spin among them).
//
// One 256-thread workgroup per walker.  The transposed inverse invT_0 of each
// determinant stays in the walker's HBM block and changes only at a flush.
// Between flushes the accepted moves of a determinant form a window of at
// most KD entries: the row change delta_a = phi(r'_a) - phi(r_a) of mover
// p_a and a copy P_a of the row invT_0[p_a], both in LDS.  With
// C[a][b] = [a == b] + delta_a . P_b, Woodbury gives the current inverse
//   invT[k][j] = invT_0[k][j] - sum_ab P_a[j] Cinv[a][b] (delta_b . invT_0[k])
// (the reference's rank-1 update, slater_plane_wave.cpp:367-423, applied KD
// at a time).  A move of electron q needs only row q.  With Q = invT_0[q],
//   R = phi'.Q - sum_ab (phi'.P_a) Cinv[a][b] (delta_b . Q)
// (the ratio of slater_plane_wave.cpp:351-365), so each move forms one 16 x 16
// block of dot products on the FP64 matrix cores -- rows phi', phi(r_q),
// delta_0..; columns Q, P_0.. -- and an accepted move borders Cinv with its
// Schur complement.  A flush applies the window to invT_0 on the matrix
// cores, 16 rows of invT_0 per step:
//   Z = -(Cinv Delta) invT_0[rows]^T, then invT_0[rows] += Z^T P,
// each row read once and written once.  Per proposed move the kernel reads
// one row of invT_0 from memory instead of updating all n rows per accepted
// move; the n^2 update is paid once per KD accepted moves, as matrix products.
//
// The chain is the generic kernel's (vmc_hip.hip sweep_walker): the same
// Philox draws, sweep order, FP64 accept test and Jastrow bookkeeping, so the
// FP64 trajectory follows the oracle twin move for move.
#include "../../include/vmc_hip.h"
#include "vmc_device.hpp"
#include "vmc_rr_launch.hpp"
#include "vmc_walker.hpp"

#include <hip/hip_runtime.h>

namespace vmc {

typedef double du_v4d __attribute__((ext_vector_type(4)));

constexpr int kDuThreads = 256;
constexpr int kDuWaves = kDuThreads / kWave;
constexpr int kDuWindow = 8;  // accepted moves per determinant between flushes
// A window is also applied once the product of its Schur complements (the
// determinant's growth since the last flush) exceeds kDuGrowth: the flush then
// forms a small inverse as the difference of larger terms, and its rounding
// grows with that ratio (walkers leaving a near-singular start accept ratios
// of 5 to 12; in equilibrium this adds about 3 % to the flushes).
constexpr double kDuGrowth = 4.0;

// P: row length padded to a multiple of 16 (n <= P); at most 2P electrons.
template <int P, int KD>
struct DuShared {
    static constexpr int NE = 2 * P;
    double px[NE], py[NE], pz[NE], U[NE], uj[NE];
    double draw[4][NE];                  // proposal (dx, dy, dz) and accept uniform per electron
    alignas(16) double phin[P];          // orbital row at the proposal
    alignas(16) double phio[P];          // orbital row at the mover's position
    alignas(16) double qrow[P];          // invT_0 row of the mover
    alignas(16) double delta[2][KD][P];  // window row changes (a flush overwrites them)
    alignas(16) double prow[2][KD][P];   // window copies of invT_0 rows
    double cinv[2][16][16];              // inverse of the window's C, per determinant
    double gpart[kDuWaves][48];          // per-wave G rows 0, 1 and column 0
    double jpart[kDuWaves];
    double dec[4];                       // decision, ratio, new Jastrow row sum, Schur complement
};

__device__ __forceinline__ double du_readlane(double v, int l) {
    union { double d; int i[2]; } u;
    u.d = v;
    u.i[0] = __builtin_amdgcn_readlane(u.i[0], l);
    u.i[1] = __builtin_amdgcn_readlane(u.i[1], l);
    return u.d;
}

// Applies the m-entry window of determinant h to its inverse inv (n x n, row
// k = invT_0[k][.]) and leaves the window empty.  Called by every thread.
template <int P, int KD>
__device__ void du_flush(DuShared<P, KD> &sh, double *inv, int n, int h, int m) {
    const int t = threadIdx.x, wave = t >> 6, lane = t & (kWave - 1);
    // delta_a <- -(Cinv Delta)_a in place, each thread one column j (rows a >= m
    // become zero).
    if (t < P) {
        double d[KD];
#pragma unroll
        for (int i = 0; i < KD; ++i) d[i] = i < m ? sh.delta[h][i][t] : 0.0;
#pragma unroll
        for (int a = 0; a < KD; ++a) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < KD; ++i)
                if (a < m && i < m) acc += sh.cinv[h][a][i] * d[i];
            sh.delta[h][a][t] = -acc;
        }
    }
    __syncthreads();
    // Row tiles of 16 rows per wave: Z = -(Cinv Delta) invT_0[rows]^T on the
    // matrix cores, then invT_0[rows] += Z^T P.  The output registers of the
    // first product are the A operands of the second (register q of lane l
    // holds Z[4q + (l >> 4)][k0 + (l & 15)]).
    const int r = lane & 15, g = lane >> 4;
    constexpr int T = P / 16;
    constexpr int SJ = P / 4;
    const double *arow = sh.delta[h][r < KD ? r : 0];
    for (int kt = wave; kt < T; kt += kDuWaves) {
        const int k0 = 16 * kt;
        const int kr = k0 + r < n ? k0 + r : n - 1;
        const double *brow = inv + (size_t)kr * n;
        du_v4d z = {0.0, 0.0, 0.0, 0.0};
        const int jb = g * SJ;
#pragma unroll 4
        for (int s = 0; s < SJ; ++s) {
            const int j = jb + s;
            const double a = r < KD ? arow[j] : 0.0;
            const double b = brow[j < n ? j : n - 1];
            z = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, z, 0, 0, 0);
        }
        for (int jt = 0; jt < T; ++jt) {
            const int jj = 16 * jt + r;
            du_v4d c;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = k0 + 4 * q + g;
                c[q] = (k < n && jj < n) ? inv[(size_t)k * n + jj] : 0.0;
            }
#pragma unroll
            for (int s = 0; s < (KD + 3) / 4; ++s) {
                const int a = 4 * s + g;
                const double b = sh.prow[h][a < KD ? a : 0][jj];
                c = __builtin_amdgcn_mfma_f64_16x16x4f64(z[s], b, c, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = k0 + 4 * q + g;
                if (k < n && jj < n) inv[(size_t)k * n + jj] = c[q];
            }
        }
    }
    __syncthreads();
}

template <int P, int KD>
__device__ void du_sweep_walker(const DevSys &s, const DevState &st, int w, int n_sweeps,
                                unsigned long long sweep0, int flags, const SweepGiven &given) {
    __shared__ DuShared<P, KD> sh;
    const int t = threadIdx.x, wave = t >> 6, lane = t & (kWave - 1);
    const int N = s.N;
    double *invw = reinterpret_cast<double *>(st.inv) + (size_t)w * s.ninv;
    double *gp = st.pos + (size_t)w * 3 * N;
    WalkerView<double> v{};
    v.inv = invw;
    v.px = sh.px;
    v.py = sh.py;
    v.pz = sh.pz;
    v.U = sh.U;
    v.uj = sh.uj;
    for (int j = t; j < P; j += kDuThreads) {
        sh.phin[j] = 0.0;
        sh.phio[j] = 0.0;
        sh.qrow[j] = 0.0;
    }
    for (int e = t; e < 2 * KD * P; e += kDuThreads) {
        (&sh.delta[0][0][0])[e] = 0.0;
        (&sh.prow[0][0][0])[e] = 0.0;
    }
    for (int i = t; i < N; i += kDuThreads) {
        sh.px[i] = gp[i];
        sh.py[i] = gp[N + i];
        sh.pz[i] = gp[2 * N + i];
    }
    __syncthreads();
    jastrow_rows<double>(s, v, t);

    double *ws = st.stats + (size_t)w * kStatCount;
    double stat[kStatCount];
#pragma unroll
    for (int k = 0; k < kStatCount; ++k) stat[k] = ws[k];
    double sigma = st.sigma[w];
    int since = st.since[w];
    if (s.reinv_k > 0 && since >= s.reinv_k) since = 0;  // re-inverted after the last sweep
    double ld0 = st.logdet[2 * w + 0], ld1 = st.logdet[2 * w + 1];
    const double b = s.b;

    for (int sw = 0; sw < n_sweeps; ++sw) {
        const uint32_t sweep = (uint32_t)(sweep0 + (unsigned long long)sw);
        // This sweep's draws (vmc_hip.hip sweep_walker): Philox tag MOVE and
        // ACCP, counter {electron, walker, sweep}; or the caller's moves.
        for (int q = t; q < N; q += kDuThreads) {
            if (given.prop) {
                const size_t g = (size_t)w * N + q;
                sh.draw[0][q] = given.prop[3 * g];
                sh.draw[1][q] = given.prop[3 * g + 1];
                sh.draw[2][q] = given.prop[3 * g + 2];
                sh.draw[3][q] = given.u[g];
                continue;
            }
            const uint4 r = philox4x32_10(make_uint4((uint32_t)q, (uint32_t)w, sweep, kTagMove), s.key0, s.key1);
            const uint4 ra = philox4x32_10(make_uint4((uint32_t)q, (uint32_t)w, sweep, kTagAccept), s.key0, s.key1);
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
            sh.draw[0][q] = dx;
            sh.draw[1][q] = dy;
            sh.draw[2][q] = dz;
            sh.draw[3][q] = u01(ra.x);
        }
        __syncthreads();
        int acc = 0;
        int m0 = 0, m1 = 0;           // window entries per determinant
        double g0 = 1.0, g1 = 1.0;    // |product of the window's Schur complements|
        for (int tt = 0; tt < N; ++tt) {
            const int p = sweep_visit(s.n_up, s.n_dn, tt);
            const int h = spin_of(s, p);
            const int n = s.n[h];
            const int q = p - s.first[h];
            const int mh = h ? m1 : m0;
            const double xo = sh.px[p], yo = sh.py[p], zo = sh.pz[p];
            double x, y, z;
            if (given.prop) {
                x = sh.draw[0][p];
                y = sh.draw[1][p];
                z = sh.draw[2][p];
            } else {
                x = xo + sigma * sh.draw[0][p];
                y = yo + sigma * sh.draw[1][p];
                z = zo + sigma * sh.draw[2][p];
                x -= s.L * floor(x * s.inv_L);
                y -= s.L * floor(y * s.inv_L);
                z -= s.L * floor(z * s.inv_L);
            }
            // The orbital rows at the proposal and at the mover's position, the
            // mover's invT_0 row, and the new Jastrow pair terms.
            double *invh = invw + s.inv_off[h];
            if (t < n) {
                sh.phin[t] = orbital<double>(s, h, t, x * s.inv_L, y * s.inv_L, z * s.inv_L);
                sh.phio[t] = orbital<double>(s, h, t, xo * s.inv_L, yo * s.inv_L, zo * s.inv_L);
                sh.qrow[t] = invh[(size_t)q * n + t];
            }
            double jp = 0.0;
            for (int e = t; e < N; e += kDuThreads) {
                double u = 0.0;
                if (e != p) {
                    const double r = pair_dist<double>(s, x, y, z, v, e, nullptr, nullptr, nullptr);
                    u = pade_u<double>(r, pair_a<double>(s, p, e), b);
                }
                sh.uj[e] = u;
                jp += u;
            }
            jp = wave_sum(jp);
            if (lane == 0) sh.jpart[wave] = jp;
            __syncthreads();
            // G = L R^T over the wave's quarter of the row: L rows phi', phi(r_q),
            // delta_0..; R rows Q, P_0..  (rows past the window hold finite
            // values whose products are not read).
            {
                const int r = lane & 15, g = lane >> 4;
                const double *arow = r == 0 ? sh.phin : r == 1 ? sh.phio : sh.delta[h][r - 2 < KD ? r - 2 : 0];
                const double *brow = r == 0 ? sh.qrow : sh.prow[h][r - 1 < KD ? r - 1 : 0];
                constexpr int S = P / 16;
                const int j0 = wave * (P / 4) + g * S;
                du_v4d gacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k = 0; k < S; ++k)
                    gacc = __builtin_amdgcn_mfma_f64_16x16x4f64(arow[j0 + k], brow[j0 + k], gacc, 0, 0, 0);
                // register q of lane l holds G[4q + (l >> 4)][l & 15]
                if (lane < 32) sh.gpart[wave][lane] = gacc[0];
                if (r == 0) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) sh.gpart[wave][32 + 4 * k + g] = gacc[k];
                }
            }
            __syncthreads();
            if (wave == 0) {
                // val: lanes 0..15 G[0][l] (phi'.Q, phi'.P_a), 16..31 G[1][l-16]
                // (phi(r_q).Q, phi(r_q).P_a), 32..47 G[l-32][0] (delta_a.Q at 34 + a).
                double val = 0.0;
                if (lane < 48) {
#pragma unroll
                    for (int wv = 0; wv < kDuWaves; ++wv) val += sh.gpart[wv][lane];
                }
                // z = Cinv c on lanes a < mh, then R = phi'.Q - sum_a (phi'.P_a) z_a.
                double za = 0.0;
                for (int k = 0; k < mh; ++k) {
                    const double ck = du_readlane(val, 34 + k);
                    if (lane < mh) za += sh.cinv[h][lane][k] * ck;
                }
                double R = du_readlane(val, 0);
                for (int a = 0; a < mh; ++a) R -= du_readlane(val, 1 + a) * du_readlane(za, a);
                double jn = 0.0;
#pragma unroll
                for (int wv = 0; wv < kDuWaves; ++wv) jn += sh.jpart[wv];
                const double dj = jn - sh.U[p];
                const double prob = R * R * exp(2.0 * dj);
                const int accept = wave_uniform((int)(sh.draw[3][p] < prob));
                double sc = 0.0;
                if (accept) {
                    // Border Cinv with the new entry mh: column c (delta_a.Q),
                    // row rr_a = (phi' - phi(r_q)).P_a, corner 1 + (phi' - phi(r_q)).Q,
                    // Schur complement sc = corner - rr.z (= R in exact arithmetic).
                    const double corner = 1.0 + du_readlane(val, 0) - du_readlane(val, 16);
                    double rz = 0.0, wk = 0.0;
                    for (int a = 0; a < mh; ++a) {
                        const double rra = du_readlane(val, 1 + a) - du_readlane(val, 17 + a);
                        rz += rra * du_readlane(za, a);
                        if (lane < mh) wk += rra * sh.cinv[h][a][lane];
                    }
                    sc = corner - rz;
                    const double isc = 1.0 / sc;
                    for (int a = 0; a < mh; ++a) {
                        const double zi = du_readlane(za, a);
                        if (lane < mh) sh.cinv[h][a][lane] += zi * wk * isc;
                    }
                    if (lane < mh) {
                        sh.cinv[h][lane][mh] = -za * isc;
                        sh.cinv[h][mh][lane] = -wk * isc;
                    }
                    if (lane == 0) sh.cinv[h][mh][mh] = isc;
                }
                if (lane == 0) {
                    sh.dec[0] = (double)accept;
                    sh.dec[1] = R;
                    sh.dec[2] = jn;
                    sh.dec[3] = sc;
                    if (given.trace) {
                        double *tr = given.trace + ((size_t)w * N + p) * kTraceCount;
                        tr[0] = R;
                        tr[1] = dj;
                        tr[2] = (double)accept;
                    }
                }
            }
            __syncthreads();
            if (sh.dec[0] != 0.0) {
                const double R = sh.dec[1];
                for (int j = t; j < P; j += kDuThreads) {
                    sh.delta[h][mh][j] = sh.phin[j] - sh.phio[j];
                    sh.prow[h][mh][j] = sh.qrow[j];
                }
                // Jastrow row sums (simulation.cpp:133-145 via the cached pair terms).
                for (int e = t; e < N; e += kDuThreads) {
                    if (e == p) continue;
                    const double r = pair_dist<double>(s, xo, yo, zo, v, e, nullptr, nullptr, nullptr);
                    sh.U[e] += sh.uj[e] - pade_u<double>(r, pair_a<double>(s, p, e), b);
                }
                if (t == 0) {
                    sh.U[p] = sh.dec[2];
                    sh.px[p] = x;
                    sh.py[p] = y;
                    sh.pz[p] = z;
                }
                ++acc;
                ++since;
                if (h) ld1 += log(fabs(R)); else ld0 += log(fabs(R));
                int mn = mh + 1;
                const double gn = (h ? g1 : g0) * fabs(sh.dec[3]);
                if (mn == KD || gn > kDuGrowth) {
                    __syncthreads();
                    du_flush<P, KD>(sh, invh, n, h, mn);
                    mn = 0;
                }
                if (h) {
                    m1 = mn;
                    g1 = mn ? gn : 1.0;
                } else {
                    m0 = mn;
                    g0 = mn ? gn : 1.0;
                }
            }
            __syncthreads();
        }
        // Apply the open windows: the inverse in memory is current between sweeps.
        if (m0 > 0) du_flush<P, KD>(sh, invw + s.inv_off[0], s.n[0], 0, m0);
        if (m1 > 0) du_flush<P, KD>(sh, invw + s.inv_off[1], s.n[1], 1, m1);
        stat[6] += (double)acc;
        stat[7] += (double)N;
        if (flags & VMC_SWEEP_ADAPT) {
            sigma *= exp(0.25 * ((double)acc / (double)N - 0.5));
            if (sigma > 0.5 * s.L) sigma = 0.5 * s.L;
        }
    }
    for (int i = t; i < N; i += kDuThreads) {
        gp[i] = sh.px[i];
        gp[N + i] = sh.py[i];
        gp[2 * N + i] = sh.pz[i];
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < kStatCount; ++k) ws[k] = stat[k];
        st.sigma[w] = sigma;
        st.since[w] = since;
        st.flags[w] = st.flags[w] & ~kFlagJrowValid;  // this kernel keeps no row sums
        st.logdet[2 * w + 0] = ld0;
        st.logdet[2 * w + 1] = ld1;
    }
}

template <int P, int KD>
__global__ __launch_bounds__(kDuThreads) void k_sweep_du(DevSys s, DevState st, int n_sweeps,
                                                        unsigned long long sweep0, int flags, SweepGiven given) {
    du_sweep_walker<P, KD>(s, st, blockIdx.x, n_sweeps, sweep0, flags, given);
}

bool du_supported(const DevSys &d) {
    return d.nmax <= 160 && (d.ndet == 1 || d.n_dn == d.n_up) && d.N <= 2 * 160;
}

bool du_launch(const DevSys &d, const DevState &st, unsigned grid, int n_sweeps, unsigned long long s0, int flags,
               const SweepGiven &given, hipStream_t stream) {
    if (!du_supported(d)) return false;
    if (d.nmax <= 64)
        hipLaunchKernelGGL((k_sweep_du<64, kDuWindow>), dim3(grid), dim3(kDuThreads), 0, stream, d, st, n_sweeps,
                           s0, flags, given);
    else if (d.nmax <= 128)
        hipLaunchKernelGGL((k_sweep_du<128, kDuWindow>), dim3(grid), dim3(kDuThreads), 0, stream, d, st,
                           n_sweeps, s0, flags, given);
    else
        hipLaunchKernelGGL((k_sweep_du<160, kDuWindow>), dim3(grid), dim3(kDuThreads), 0, stream, d, st,
                           n_sweeps, s0, flags, given);
    return true;
}

}  // namespace vmc
