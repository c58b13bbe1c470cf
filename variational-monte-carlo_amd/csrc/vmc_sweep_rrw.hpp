// vmc_sweep_rrw.hpp -- register-resident sweep and local energy for spins of
// 33 to 192 electrons: one to three wavefronts per spin.
//
// The workgroup of a walker holds WPS wavefronts per determinant; thread
// (wave v, lane l) owns row r = (v mod WPS) * 64 + l of spin h = v / WPS:
// that electron's position, its Jastrow row sum U and its row of the spin's
// transposed inverse in registers.  A move of electron p (spin hp) is three
// phases, A and B separated by a workgroup barrier only when a spin spans
// several wavefronts, B and C always:
//   A  every thread forms its new pair term against the proposal (DPP wave
//      sum -> LDS); the rows of spin hp form the orbital row of the proposal;
//   B  spin hp's rows form s_k = sum_m phi_m invT[k][m]; the mover publishes
//      R = s_p, its U, old position and inverse row;
//   C  every thread takes the same accept decision from the same LDS values;
//      on accept spin hp's rows apply Sherman-Morrison (one FMA per element,
//      the other spin's wavefronts skip it on a wave-uniform branch) and every
//      thread updates its U.
// Same chain as the one-wavefront kernel (the reference's
// src/simulation/simulation.cpp:80-156 with per-sweep electron order and
// Philox draws), same HBM layout, so ensembles switch kernels freely.
#pragma once

#include "vmc_sweep_rr.hpp"

namespace vmc {

template <typename Real, int NU, int ND, int WPS>
struct RRWShared {
    static constexpr int N = NU + ND;
    static constexpr int NMAX = NU > ND ? NU : ND;
    static constexpr int NWAVE = (ND > 0 ? 2 : 1) * WPS;
    double pos[3][N];
    alignas(16) double prop[4][N];  // this sweep's proposals (x, y, z) and accept uniforms
    Real pph[6][N];      // axis phasors of the proposals
    Real tab[kRRChunk * 3 * (kRRMaxPow + 1) * 2];
    // Rows padded to 16 bytes, so that a row picked at run time still moves
    // with 16-byte LDS reads.
    static constexpr int kRowPad = (NMAX + 3) / 4 * 4;
    alignas(16) Real phi[2][kRowPad];   // orbital rows of the pending moves (by spin)
    alignas(16) Real row[2][kRowPad];   // inverse row of the (up) mover (by move parity)
    alignas(16) Real row2[2][kRowPad];  // inverse row of the down mover of a paired move
    double red[NWAVE];   // per-wave partial sums
    double red4[NWAVE][4];
    double jred[2][NWAVE];   // per-wave new Jastrow pair sums (by move parity)
    double jred2[2][NWAVE];  // the same for the down proposal of a paired move
    double mv[2][8];     // (up) mover: R, U, x, y, z (by move parity)
    double mv2[2][8];    // down mover of a paired move
    double stat[kStatCount];
    double misc[4];      // sigma, log|det| up, log|det| down
};

template <typename Real, int NU, int ND, int WPS>
struct RRW {
    static constexpr int N = NU + ND;
    static constexpr int NMAX = NU > ND ? NU : ND;
    static constexpr int NWAVE = (ND > 0 ? 2 : 1) * WPS;
    static constexpr int NT = NWAVE * kWave;
    static constexpr auto orb = rr::shell_orbitals<NMAX>();
    using Sh = RRWShared<Real, NU, ND, WPS>;
    static_assert(ND == 0 || ND == NU, "polarised or equal spins");
    static_assert(NMAX <= WPS * kWave, "rows per spin exceed the spin's wavefronts");
    static constexpr int kMaxC = rr::max_component<NMAX>();
    static constexpr bool kOneBarrier = WPS == 1 && NMAX >= 57;
    // Sizes whose sweep moves an up and a down electron per pass (move_pair):
    // measured faster there; at 81 and 123 per spin in FP32 (and 33, 57 in
    // FP64) the paired pass's registers cost more than it saves.
    static constexpr bool kPaired = ND == NU && (sizeof(Real) == 4 ? (NMAX == 33 || NMAX == 57 || NMAX == 147)
                                                                  : NMAX == 81);
    static_assert(kMaxC <= 3, "orbital phasors use powers up to 3");

    // e^{2 pi i n.u} from the axis phasors a and their powers.
    __device__ static rr::Cx<Real> phasor(int nx, int ny, int nz, const rr::Cx<Real> a[3]) {
        rr::Cx<Real> a2[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) a2[k] = rr::cmul(a[k], a[k]);
        if constexpr (kMaxC <= 2) {
            return rr::orbital_phasor<Real>(nx, ny, nz, a, a2);
        } else {
            rr::Cx<Real> a3[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) a3[k] = rr::cmul(a2[k], a[k]);
            return rr::orbital_phasor3<Real>(nx, ny, nz, a, a2, a3);
        }
    }

    // Orbital r of the compile-time table, packed (nx+4) | (ny+4)<<4 |
    // (nz+4)<<8 | type<<12 (selected once per kernel).
    __device__ static int packed_orbital(int r) {
        int o = 0;
#pragma unroll
        for (int m = 0; m < NMAX; ++m)
            if (r == m) o = (orb[m].x + 4) | ((orb[m].y + 4) << 4) | ((orb[m].z + 4) << 8) | (orb[m].t << 12);
        return o;
    }

    struct Thread {
        int v, lane, h, r, nh, e;
        bool has_e;
        __device__ Thread() {
            const int t = threadIdx.x;
            v = t >> 6;
            lane = t & (kWave - 1);
            h = ND > 0 ? v / WPS : 0;
            r = (v % WPS) * kWave + lane;
            nh = h ? ND : NU;
            has_e = r < nh;
            e = h ? NU + r : r;
        }
    };

    // Workgroup sum of one value per thread (fixed order: every thread gets
    // the same double).
    __device__ static double block_total(double x, double *red) {
        const double wsum = rr::wave_total(x);
        const int v = threadIdx.x >> 6;
        if ((threadIdx.x & (kWave - 1)) == 0) red[v] = wsum;
        __syncthreads();
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < NWAVE; ++i) t += red[i];
        __syncthreads();
        return t;
    }

    __device__ static void load(const DevSys &s, const DevState &st, int w, Sh &sh, const Thread &T,
                                Real (&inv)[NMAX], double &x, double &y, double &z) {
        const Real *blk = reinterpret_cast<const Real *>(st.inv) + (size_t)w * s.ninv;
#pragma unroll
        for (int m = 0; m < NMAX; ++m)
            inv[m] = (T.has_e && m < T.nh) ? blk[(T.h ? NU * NU : 0) + T.r * T.nh + m] : (Real)0;
        load_pos(s, st, w, sh, T, x, y, z);
    }

    // The thread's electron position, also into the LDS mirror.
    __device__ static void load_pos(const DevSys &, const DevState &st, int w, Sh &sh, const Thread &T,
                                    double &x, double &y, double &z) {
        const double *gp = st.pos + (size_t)w * 3 * N;
        x = T.has_e ? gp[T.e] : 0.0;
        y = T.has_e ? gp[N + T.e] : 0.0;
        z = T.has_e ? gp[2 * N + T.e] : 0.0;
        if (T.has_e) {
            sh.pos[0][T.e] = x;
            sh.pos[1][T.e] = y;
            sh.pos[2][T.e] = z;
        }
    }

    // Jastrow row sum of the thread's electron against every other.
    __device__ static double jastrow_row(const RRConst<Real> &c, const Sh &sh, const Thread &T,
                                         double x, double y, double z) {
        double U = 0.0;
        if (!T.has_e) return U;
        for (int j = 0; j < N; ++j) {
            if (j == T.e) continue;
            const Real dx = c.mi_r(x, sh.pos[0][j]);
            const Real dy = c.mi_r(y, sh.pos[1][j]);
            const Real dz = c.mi_r(z, sh.pos[2][j]);
            const Real rr_ = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
            const Real a = ((ND > 0 && j >= NU) ? 1 : 0) == T.h ? c.a_same : c.a_anti;
            U += (double)(a * rr_ * wrcp<Real>((Real)1 + c.b * rr_));
        }
        return U;
    }

    // Pade pair value of the point (xn, yn, zn) with (x, y, z), as the moves
    // form it.
    __device__ static Real pair_at(const RRConst<Real> &c, double xn, double yn, double zn, double x, double y,
                                   double z, Real a) {
        const Real dx = c.mi_r(xn, x);
        const Real dy = c.mi_r(yn, y);
        const Real dz = c.mi_r(zn, z);
        const Real rn = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
        return a * rn * wrcp<Real>((Real)1 + c.b * rn);
    }

    // Metropolis test (simulation.cpp:125-131), exact FP64 exponential within
    // 1e-5 of the threshold; wave-uniform inputs and result.
    __device__ static int accept_test(double R, double dj, double ua) {
        const double r2 = R * R;
        const float e32 = __builtin_amdgcn_exp2f((float)(2.0 * dj * 1.4426950408889634));
        const double approx = r2 * (double)e32;
        int acc;
        if (!(approx == approx)) acc = 0;
        else if (fabs(ua - approx) > 1e-5 * approx || !(fabs(dj) < 40.0)) acc = ua < approx;
        else acc = ua < r2 * rr_exp_f64(2.0 * dj);
        return wave_uniform(acc);
    }

    // The moves of up electron i and then down electron i (the sweep order
    // of equal spins) in one pass: each spin's wavefronts form their orbital
    // row and ratios at their spin's proposal, and every thread forms its pair
    // terms with both proposals.  The down move then takes the up move's
    // outcome into its Jastrow change (the pair terms of the up electron with
    // the down electron, formed as the threads form them).  Same decisions
    // and updates as two calls of move.  Returns the accepted moves.
    __device__ static int move_pair(const RRConst<Real> &c, Sh &sh, const Thread &T, Real (&inv)[NMAX],
                                    double &x, double &y, double &z, double &U, int ob, int i, int par,
                                    LogDetTrack &ldt, double *rec) {
        static_assert(ND == NU, "paired moves need equal spins");
        const int pu = i, pd = NU + i;
        const double xu = sh.prop[0][pu], yu = sh.prop[1][pu], zu = sh.prop[2][pu];
        const double xd = sh.prop[0][pd], yd = sh.prop[1][pd], zd = sh.prop[2][pd];
        // Phase A.
        if (T.has_e) {
            const int pp = T.h ? pd : pu;
            rr::Cx<Real> a[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a[k].re = sh.pph[2 * k][pp];
                a[k].im = sh.pph[2 * k + 1][pp];
            }
            const rr::Cx<Real> q = phasor(orb_nx(ob), orb_ny(ob), orb_nz(ob), a);
            sh.phi[T.h][T.r] = orb_t(ob) ? q.im : q.re;
        }
        Real un_u = (Real)0, un_d = (Real)0;
        if (T.has_e && T.e != pu) un_u = pair_at(c, xu, yu, zu, x, y, z, T.h == 0 ? c.a_same : c.a_anti);
        if (T.has_e && T.e != pd) un_d = pair_at(c, xd, yd, zd, x, y, z, T.h == 1 ? c.a_same : c.a_anti);
        {
            const Real wu = rr::wave_total(un_u), wd = rr::wave_total(un_d);
            if (T.lane == 0) {
                sh.jred[par][T.v] = (double)wu;
                sh.jred2[par][T.v] = (double)wd;
            }
        }
        if constexpr (kOneBarrier) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
        // Phase B: every thread's s_k with its spin's orbital row; the two
        // movers publish their values and rows.
        Real sk;
        {
            Real s0 = 0, s1 = 0;
#pragma unroll
            for (int m = 0; m < NMAX; m += 2) {
                s0 += sh.phi[T.h][m] * inv[m];
                if (m + 1 < NMAX) s1 += sh.phi[T.h][m + 1] * inv[m + 1];
            }
            sk = s0 + s1;
        }
        if (T.r == i) {
            double *mvp = T.h ? sh.mv2[par] : sh.mv[par];
            Real *rowp = T.h ? sh.row2[par] : sh.row[par];
            mvp[0] = (double)sk;
            mvp[1] = U;
            mvp[2] = x;
            mvp[3] = y;
            mvp[4] = z;
#pragma unroll
            for (int m = 0; m < NMAX; ++m) rowp[m] = inv[m];
        }
        double jn_u = 0.0, jn_d = 0.0;
        if constexpr (!kOneBarrier) {
#pragma unroll
            for (int v = 0; v < NWAVE; ++v) {
                jn_u += sh.jred[par][v];
                jn_d += sh.jred2[par][v];
            }
        }
        __syncthreads();
        if constexpr (kOneBarrier) {
#pragma unroll
            for (int v = 0; v < NWAVE; ++v) {
                jn_u += sh.jred[par][v];
                jn_d += sh.jred2[par][v];
            }
        }
        // Phase C, up move.
        const double Ru = sh.mv[par][0];
        const double uxo = sh.mv[par][2], uyo = sh.mv[par][3], uzo = sh.mv[par][4];
        const double dj_u = jn_u - sh.mv[par][1];
        const int acc_u = accept_test(Ru, dj_u, sh.prop[3][pu]);
        double Ud = sh.mv2[par][1];
        if (acc_u) {
            ldt.add(0, Ru);
            if (T.h == 0) {
                const Real inv_ratio = wrcp<Real>((Real)Ru);
                Real f = (T.r == i) ? (Real)1 - inv_ratio : sk * inv_ratio;
                f = T.has_e ? f : (Real)0;
#pragma unroll
                for (int m = 0; m < NMAX; ++m) inv[m] -= sh.row[par][m] * f;
            }
            if (T.has_e && T.e != pu)
                U += (double)un_u - (double)pair_at(c, uxo, uyo, uzo, x, y, z, T.h == 0 ? c.a_same : c.a_anti);
            if (T.has_e && T.e == pu) {
                U = jn_u;
                x = xu;
                y = yu;
                z = zu;
                sh.pos[0][pu] = xu;
                sh.pos[1][pu] = yu;
                sh.pos[2][pu] = zu;
            }
            // The down mover's row sum after this update, and the down
            // proposal's pair with the up electron where it now is.
            const double dxo = sh.mv2[par][2], dyo = sh.mv2[par][3], dzo = sh.mv2[par][4];
            Ud += (double)pair_at(c, xu, yu, zu, dxo, dyo, dzo, c.a_anti) -
                  (double)pair_at(c, uxo, uyo, uzo, dxo, dyo, dzo, c.a_anti);
            const Real ud_new = pair_at(c, xd, yd, zd, xu, yu, zu, c.a_anti);
            const Real ud_old = pair_at(c, xd, yd, zd, uxo, uyo, uzo, c.a_anti);
            jn_d += (double)ud_new - (double)ud_old;
            if (T.has_e && T.e == pu) un_d = ud_new;
        }
        // Phase C, down move.
        const double Rd = sh.mv2[par][0];
        const double dj_d = jn_d - Ud;
        const int acc_d = accept_test(Rd, dj_d, sh.prop[3][pd]);
        if (rec && threadIdx.x == 0) {
            double *ru = rec + kTraceCount * pu, *rd = rec + kTraceCount * pd;
            ru[0] = Ru;
            ru[1] = dj_u;
            ru[2] = (double)acc_u;
            rd[0] = Rd;
            rd[1] = dj_d;
            rd[2] = (double)acc_d;
        }
        if (acc_d) {
            ldt.add(1, Rd);
            if (T.h == 1) {
                const Real inv_ratio = wrcp<Real>((Real)Rd);
                Real f = (T.r == i) ? (Real)1 - inv_ratio : sk * inv_ratio;
                f = T.has_e ? f : (Real)0;
#pragma unroll
                for (int m = 0; m < NMAX; ++m) inv[m] -= sh.row2[par][m] * f;
            }
            const double dxo = sh.mv2[par][2], dyo = sh.mv2[par][3], dzo = sh.mv2[par][4];
            if (T.has_e && T.e != pd)
                U += (double)un_d - (double)pair_at(c, dxo, dyo, dzo, x, y, z, T.h == 1 ? c.a_same : c.a_anti);
            if (T.has_e && T.e == pd) {
                U = jn_d;
                x = xd;
                y = yd;
                z = zd;
                sh.pos[0][pd] = xd;
                sh.pos[1][pd] = yd;
                sh.pos[2][pd] = zd;
            }
        }
        return acc_u + acc_d;
    }

    // One move of electron p, the sweep's move number par (mod 2) selecting
    // the LDS buffers phase C reads.  Returns 1 if accepted (uniform).
    __device__ static int move(const RRConst<Real> &c, Sh &sh, const Thread &T, Real (&inv)[NMAX],
                               double &x, double &y, double &z, double &U, int ob, int p, int par,
                               LogDetTrack &ldt, double *rec) {
        const int hp = (ND > 0 && p >= NU) ? 1 : 0;
        const int rp = hp ? p - NU : p;
        const int NP = hp ? ND : NU;
        // Phase A.
        const double xn = sh.prop[0][p], yn = sh.prop[1][p], zn = sh.prop[2][p];
        if (T.h == hp && T.has_e) {
            rr::Cx<Real> a[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a[k].re = sh.pph[2 * k][p];
                a[k].im = sh.pph[2 * k + 1][p];
            }
            const rr::Cx<Real> q = phasor(orb_nx(ob), orb_ny(ob), orb_nz(ob), a);
            sh.phi[hp][T.r] = orb_t(ob) ? q.im : q.re;
        }
        Real un = (Real)0;
        if (T.has_e && T.e != p) {
            const Real dx = c.mi_r(xn, x);
            const Real dy = c.mi_r(yn, y);
            const Real dz = c.mi_r(zn, z);
            const Real rn = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
            const Real a = (T.h == hp) ? c.a_same : c.a_anti;
            un = a * rn * wrcp<Real>((Real)1 + c.b * rn);
        }
        {
            const Real ws = rr::wave_total(un);
            if (T.lane == 0) sh.jred[par][T.v] = (double)ws;
        }
        // With one wavefront per spin the orbital row is written and read by
        // the mover's wavefront alone, in issue order: phases A and B need no
        // workgroup barrier between them (used from 57 per spin; at 33 the
        // pair sums read after the one barrier cost more than the barrier).  The mover's values and the pair
        // sums that phase C reads alternate between two buffers by move
        // parity, so a wavefront already in the next move's phase A or B
        // cannot overwrite what a slower one still reads.
        if constexpr (kOneBarrier) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
        // Phase B.
        Real sk = (Real)0;
        if (T.h == hp) {
            Real s0 = 0, s1 = 0;
#pragma unroll
            for (int m = 0; m < NMAX; m += 2) {
                s0 += sh.phi[hp][m] * inv[m];
                if (m + 1 < NMAX) s1 += sh.phi[hp][m + 1] * inv[m + 1];
            }
            sk = s0 + s1;
            if (T.r == rp) {
                sh.mv[par][0] = (double)sk;
                sh.mv[par][1] = U;
                sh.mv[par][2] = x;
                sh.mv[par][3] = y;
                sh.mv[par][4] = z;
#pragma unroll
                for (int m = 0; m < NMAX; ++m) sh.row[par][m] = inv[m];
            }
        }
        // The pair sums: complete since the first barrier when there is one
        // (read before the second, off phase C's path), else after this one.
        double jn = 0.0;
        if constexpr (!kOneBarrier) {
#pragma unroll
            for (int i = 0; i < NWAVE; ++i) jn += sh.jred[par][i];
        }
        __syncthreads();
        if constexpr (kOneBarrier) {
#pragma unroll
            for (int i = 0; i < NWAVE; ++i) jn += sh.jred[par][i];
        }
        // Phase C: the same decision in every thread (simulation.cpp:125-131),
        // exact FP64 exponential within 1e-5 of the threshold.
        const double R = sh.mv[par][0];
        const double dj = jn - sh.mv[par][1];
        const double ua = sh.prop[3][p];
        const double r2 = R * R;
        const float e32 = __builtin_amdgcn_exp2f((float)(2.0 * dj * 1.4426950408889634));
        const double approx = r2 * (double)e32;
        int acc;
        if (!(approx == approx)) acc = 0;
        else if (fabs(ua - approx) > 1e-5 * approx || !(fabs(dj) < 40.0)) acc = ua < approx;
        else acc = ua < r2 * rr_exp_f64(2.0 * dj);
        acc = wave_uniform(acc);
        if (rec && threadIdx.x == 0) {
            double *q = rec + kTraceCount * p;
            q[0] = R;
            q[1] = dj;
            q[2] = (double)acc;
        }
        if (!acc) return 0;
        ldt.add(hp, R);
        if (T.h == hp) {
            const Real inv_ratio = wrcp<Real>((Real)R);
            Real f = (T.r == rp) ? (Real)1 - inv_ratio : sk * inv_ratio;
            f = T.has_e ? f : (Real)0;
#pragma unroll
            for (int m = 0; m < NMAX; ++m) inv[m] -= sh.row[par][m] * f;
        }
        if (T.has_e && T.e != p) {
            const Real dx = c.mi_r(sh.mv[par][2], x);
            const Real dy = c.mi_r(sh.mv[par][3], y);
            const Real dz = c.mi_r(sh.mv[par][4], z);
            const Real ro = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
            const Real a = (T.h == hp) ? c.a_same : c.a_anti;
            U += (double)un - (double)(a * ro * wrcp<Real>((Real)1 + c.b * ro));
        }
        if (T.has_e && T.e == p) {
            U = jn;
            x = xn;
            y = yn;
            z = zn;
            sh.pos[0][p] = xn;
            sh.pos[1][p] = yn;
            sh.pos[2][p] = zn;
        }
        (void)NP;
        return 1;
    }
};

// Proposal q (in [0, L)) of electron e, its axis phasors and accept uniform
// ua into the electron's slots.
template <typename Real>
__device__ __forceinline__ void rrw_stage_proposal(const DevSys &s, const double (&q)[3], double ua, int e, int N,
                                                   double *prop, Real *pph) {
    for (int k = 0; k < 3; ++k) {
        prop[k * N + e] = q[k];
        const rr::Cx<Real> ph = rr::axis_phasor<Real>(q[k], s.inv_L);
        pph[(2 * k) * N + e] = ph.re;
        pph[(2 * k + 1) * N + e] = ph.im;
    }
    prop[3 * N + e] = ua;
}

// Philox draws of one electron for this sweep into its proposal slots.
template <typename Real>
__device__ __forceinline__ void rrw_proposal(const DevSys &s, double sigma, double x0, double y0,
                                             double z0, int e, int w, uint32_t sweep, int N,
                                             double *prop, Real *pph) {
    const uint4 rm = philox4x32_10(make_uint4((uint32_t)e, (uint32_t)w, sweep, kTagMove), s.key0, s.key1);
    const uint4 ra = philox4x32_10(make_uint4((uint32_t)e, (uint32_t)w, sweep, kTagAccept), s.key0, s.key1);
    double dx, dy, dz;
    if (s.move_kind == 1) {
        const double rho0 = sqrt(-2.0 * log(u01(rm.x)));
        const double rho1 = sqrt(-2.0 * log(u01(rm.z)));
        double s0, c0, s1, c1;
        sincospi(2.0 * u01(rm.y), &s0, &c0);
        sincospi(2.0 * u01(rm.w), &s1, &c1);
        dx = rho0 * c0;
        dy = rho0 * s0;
        dz = rho1 * c1;
    } else {
        dx = 2.0 * u01(rm.x) - 1.0;
        dy = 2.0 * u01(rm.y) - 1.0;
        dz = 2.0 * u01(rm.z) - 1.0;
    }
    const double x = x0 + sigma * dx, y = y0 + sigma * dy, z = z0 + sigma * dz;
    const double q[3] = {x - s.L * floor(x * s.inv_L), y - s.L * floor(y * s.inv_L),
                         z - s.L * floor(z * s.inv_L)};
    rrw_stage_proposal<Real>(s, q, u01(ra.x), e, N, prop, pph);
}

// kGiven: the moves come from the caller (SweepGiven) instead of Philox and
// each move's ratio, Jastrow change and decision are recorded.
template <typename Real, int NU, int ND, int WPS, bool kGiven = false>
__device__ __forceinline__ void sweep_rrw_body(const DevSys &s, const DevState &st, int w,
                                               unsigned long long sweep0, int flags,
                                               SweepGiven given = SweepGiven{}) {
    using K = RRW<Real, NU, ND, WPS>;
    constexpr int N = K::N;
    constexpr int NMAX = K::NMAX;
    __shared__ typename K::Sh sh;
    const RRConst<Real> cst(s);
    const typename K::Thread T;
    Real inv[NMAX];
    double x, y, z;
    K::load(s, st, w, sh, T, inv, x, y, z);
    if (threadIdx.x < kStatCount) sh.stat[threadIdx.x] = st.stats[(size_t)w * kStatCount + threadIdx.x];
    if (threadIdx.x == 0) {
        sh.misc[0] = st.sigma[w];
        sh.misc[1] = st.logdet[2 * w + 0];
        sh.misc[2] = st.logdet[2 * w + 1];
    }
    __syncthreads();
    const int wflags = st.flags[w];
    double U;
    if (!(wflags & kFlagJrowValid) || ((uint32_t)sweep0 % kJrowRefresh) == 0)
        U = K::jastrow_row(cst, sh, T, x, y, z);
    else
        U = T.has_e ? st.jrow[(size_t)w * N + T.e] : 0.0;
    const int ob = K::packed_orbital(T.r);
    double *rec = nullptr;
    if constexpr (kGiven) {
        if (T.has_e) {
            const size_t q = (size_t)w * N + T.e;
            const double p3[3] = {given.prop[3 * q], given.prop[3 * q + 1], given.prop[3 * q + 2]};
            rrw_stage_proposal<Real>(s, p3, given.u[q], T.e, N, &sh.prop[0][0], &sh.pph[0][0]);
        }
        rec = given.trace + (size_t)w * N * kTraceCount;
    } else {
        if (T.has_e)
            rrw_proposal<Real>(s, sh.misc[0], x, y, z, T.e, w, (uint32_t)sweep0, N, &sh.prop[0][0], &sh.pph[0][0]);
    }
    __syncthreads();
    int acc = 0;
    LogDetTrack ldt;
    if constexpr (K::kPaired) {
#pragma unroll 1
        for (int i = 0; i < NU; ++i) acc += K::move_pair(cst, sh, T, inv, x, y, z, U, ob, i, i & 1, ldt, rec);
    } else if constexpr (ND > 0 && sizeof(Real) == 8) {
        // The sweep order up i, down i as two calls with constant spin and
        // buffer parity (the move's addressing then folds; in FP32 at 81 and
        // 123 per spin the doubled loop body measured slower).
#pragma unroll 1
        for (int i = 0; i < NU; ++i) {
            acc += K::move(cst, sh, T, inv, x, y, z, U, ob, i, 0, ldt, rec);
            acc += K::move(cst, sh, T, inv, x, y, z, U, ob, NU + i, 1, ldt, rec);
        }
    } else {
#pragma unroll 1
        for (int t = 0; t < N; ++t)
            acc += K::move(cst, sh, T, inv, x, y, z, U, ob, sweep_visit(NU, ND, t), t & 1, ldt, rec);
    }
    __syncthreads();
    if (flags & VMC_SWEEP_ADAPT) {
        const double f = exp(0.25 * ((double)acc / (double)N - 0.5));
        if (threadIdx.x == 0) {
            const double sigma = sh.misc[0] * f;
            sh.misc[0] = sigma > 0.5 * s.L ? 0.5 * s.L : sigma;
        }
    }
    Real *blk = reinterpret_cast<Real *>(st.inv) + (size_t)w * s.ninv;
    if (T.has_e) {
#pragma unroll
        for (int m = 0; m < NMAX; ++m)
            if (m < T.nh) blk[(T.h ? NU * NU : 0) + T.r * T.nh + m] = inv[m];
        double *gq = st.pos + (size_t)w * 3 * N;
        gq[T.e] = x;
        gq[N + T.e] = y;
        gq[2 * N + T.e] = z;
        st.jrow[(size_t)w * N + T.e] = U;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sh.stat[6] += (double)acc;
        sh.stat[7] += (double)N;
        st.sigma[w] = sh.misc[0];
        // (a count at K was re-inverted after the last sweep)
        st.since[w] = (s.reinv_k > 0 && st.since[w] >= s.reinv_k ? 0 : st.since[w]) + acc;
        st.flags[w] = wflags | kFlagJrowValid;
        st.logdet[2 * w + 0] = ldt.fold(0, sh.misc[1]);
        if (ND > 0) st.logdet[2 * w + 1] = ldt.fold(1, sh.misc[2]);
    }
    __syncthreads();
    if (threadIdx.x < kStatCount) st.stats[(size_t)w * kStatCount + threadIdx.x] = sh.stat[threadIdx.x];
}

// Local energy after a sweep (energy_tracking.cpp:253-382 with the Jastrow
// and Slater derivatives of jastrow_pade.cpp:57-138, slater_plane_wave.cpp:
// 425-499): per thread its electron's pair terms and Slater terms; the
// structure factor from per-electron axis-power tables staged in LDS, the
// workgroup's threads owning G vectors.
// With `parts` the walker's kinetic, V_real, V_recip, E_L, log|det| and J go
// there ([w][kPartCount]) and the statistics are left alone.
template <typename Real, int NU, int ND, int WPS>
__device__ __forceinline__ void measure_rrw_body(const DevSys &s, const DevState &st, int w,
                                                 int block_size, double *parts = nullptr) {
    using K = RRW<Real, NU, ND, WPS>;
    constexpr int N = K::N;
    constexpr int NMAX = K::NMAX;
    constexpr int NT = K::NT;
    __shared__ typename K::Sh sh;
    const RRConst<Real> c(s);
    const typename K::Thread T;
    // Above 96 per spin the inverse row is read where the Slater loop needs
    // it (cached global loads) instead of being held in registers.
    constexpr bool kRowsInMemory = NMAX > 96;
    Real inv[kRowsInMemory ? 1 : NMAX];
    double x, y, z;
    if constexpr (kRowsInMemory)
        K::load_pos(s, st, w, sh, T, x, y, z);
    else
        K::load(s, st, w, sh, T, inv, x, y, z);
    const Real *row = reinterpret_cast<const Real *>(st.inv) + (size_t)w * s.ninv + (T.h ? NU * NU : 0) +
                      (T.has_e ? T.r * T.nh : 0);
    const auto inv_at = [&](int m) -> Real {
        if constexpr (kRowsInMemory) return row[m];
        else return inv[m];
    };
    __syncthreads();
    // Pair terms: Jastrow gradient / Laplacian and real-space Ewald, each
    // unordered pair once.  In step k (1 <= k <= N/2) electron e pairs with
    // (e + k) mod N and keeps its side; the partner's side (gradient negated,
    // same Laplacian term) goes through LDS -- the sweep's proposal arrays,
    // unused here -- kXch steps at a time between two barriers.  For even N
    // the last step's pairs are met from both ends: each keeps its own side
    // and their energy counts half.  Working-precision sums per chunk, FP64
    // across chunks.
    constexpr int kSteps = N / 2;
    constexpr bool kEven = N % 2 == 0;
    constexpr int kXch = (int)((sizeof(sh.prop) + sizeof(sh.pph)) / (4 * N * sizeof(Real)));
    static_assert(kXch >= 1, "exchange buffer");
    static_assert(offsetof(typename K::Sh, pph) == offsetof(typename K::Sh, prop) + sizeof(sh.prop),
                  "prop and pph are contiguous");
    Real *xch = reinterpret_cast<Real *>(&sh.prop[0][0]);  // [kXch][4][N]
    double gx = 0, gy = 0, gz = 0, lap = 0, vreal = 0;
    {
        const Real m2b = (Real)-2 * c.b;
        for (int k0 = 1; k0 <= kSteps; k0 += kXch) {
            Real gx_c = 0, gy_c = 0, gz_c = 0, l_c = 0, v_c = 0;
#pragma unroll
            for (int kk = 0; kk < kXch; ++kk) {
                const int k = k0 + kk;
                if (k > kSteps || !T.has_e) break;
                const bool last_even = kEven && k == kSteps;
                int j = T.e + k;
                j = j >= N ? j - N : j;
                const Real dx = c.mi_r(x, sh.pos[0][j]);
                const Real dy = c.mi_r(y, sh.pos[1][j]);
                const Real dz = c.mi_r(z, sh.pos[2][j]);
                Real r, inv_r, den;
                c.pair_geom(dx * dx + dy * dy + dz * dz, r, inv_r, den);
                const Real a = ((ND > 0 && j >= NU) ? 1 : 0) == T.h ? c.a_same : c.a_anti;
                const Real v = c.ewald_ri(r, inv_r);
                v_c += last_even ? (Real)0.5 * v : v;
                Real pgx = 0, pgy = 0, pgz = 0, pl = 0;
                if (r >= (Real)1e-12) {
                    const Real d1 = a * den * den;
                    const Real gf = d1 * inv_r;
                    const Real l = d1 * (m2b * den + (Real)2 * inv_r);
                    gx_c += gf * dx;
                    gy_c += gf * dy;
                    gz_c += gf * dz;
                    l_c += l;
                    pgx = -gf * dx;
                    pgy = -gf * dy;
                    pgz = -gf * dz;
                    pl = l;
                }
                if (!last_even) {
                    Real *slot = xch + (size_t)kk * 4 * N;
                    slot[j] = pgx;
                    slot[N + j] = pgy;
                    slot[2 * N + j] = pgz;
                    slot[3 * N + j] = pl;
                }
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < kXch; ++kk) {
                const int k = k0 + kk;
                if (k > kSteps || !T.has_e) break;
                if (kEven && k == kSteps) break;
                const Real *slot = xch + (size_t)kk * 4 * N;
                gx_c += slot[T.e];
                gy_c += slot[N + T.e];
                gz_c += slot[2 * N + T.e];
                l_c += slot[3 * N + T.e];
            }
            __syncthreads();
            gx += (double)gx_c;
            gy += (double)gy_c;
            gz += (double)gz_c;
            lap += (double)l_c;
            vreal += (double)v_c;
        }
    }
    // Slater gradient and Laplacian of ln|D_h| at the thread's electron.
    double sx = 0, sy = 0, sz = 0, sl = 0;
    if (T.has_e) {
        rr::Cx<Real> a[3];
        a[0] = rr::axis_phasor<Real>(x, s.inv_L);
        a[1] = rr::axis_phasor<Real>(y, s.inv_L);
        a[2] = rr::axis_phasor<Real>(z, s.inv_L);
        const Real ku = (Real)s.kunit;
#pragma unroll
        for (int m = 0; m < NMAX; ++m) {
            if (K::orb[m].t == 1) continue;
            const rr::Cx<Real> ph = K::phasor(K::orb[m].x, K::orb[m].y, K::orb[m].z, a);
            const Real k2 = ku * ku * (Real)(K::orb[m].x * K::orb[m].x + K::orb[m].y * K::orb[m].y + K::orb[m].z * K::orb[m].z);
            const Real wc = inv_at(m);
            Real gsc = -wc * ph.im;
            Real lpc = -wc * ph.re;
            if (m + 1 < NMAX && K::orb[m + 1].t == 1) {
                const Real ws = inv_at(m + 1);
                gsc += ws * ph.re;
                lpc -= ws * ph.im;
            }
            sx += (double)(gsc * ku * (Real)K::orb[m].x);
            sy += (double)(gsc * ku * (Real)K::orb[m].y);
            sz += (double)(gsc * ku * (Real)K::orb[m].z);
            sl += (double)(lpc * k2);
        }
    }
    const double slap = sl - (sx * sx + sy * sy + sz * sz);
    const double tx = sx + gx, ty = sy + gy, tz = sz + gz;
    const double kin_t = T.has_e ? (slap + lap + tx * tx + ty * ty + tz * tz) : 0.0;
    const double kinetic = -0.5 * K::block_total(kin_t, sh.red);
    const double vr = K::block_total(T.has_e ? vreal : 0.0, sh.red);
    // Structure factor S(G) = sum_j e^{i G.r_j}: threads own G vectors; the
    // tables are staged in the run prop | pph | tab.
    using Sh = typename K::Sh;
    static_assert(offsetof(Sh, pph) == offsetof(Sh, prop) + sizeof(sh.prop) &&
                      offsetof(Sh, tab) == offsetof(Sh, pph) + sizeof(sh.pph) &&
                      sizeof(sh.prop) + sizeof(sh.pph) + sizeof(sh.tab) >= sizeof(Real) * sf::table_reals<Real>(),
                  "structure-factor tables exceed the run prop | pph | tab");
    // Column segments (a third of the LDS reads per G) keep only as many
    // threads busy as there are segments (about 80 for |n|^2 <= 25): from
    // 256 threads one G vector per thread (at most 768 G) keeps more of
    // them busy, which measured faster there.
    const auto pos_of = [](int ax, int e) { return sh.pos[ax][e]; };
    auto *sf_tab = reinterpret_cast<sf::Pair<Real> *>(&sh.prop[0][0]);
    double acc;
    if constexpr (NT >= 256)
        acc = sf::energy<Real, N, (12 * kWave + NT - 1) / NT, NT>(s, sf_tab, (int)threadIdx.x, pos_of);
    else
        acc = sf::energy_cols<Real, N, NT>(s, sf_tab, (int)threadIdx.x, pos_of);
    const double vrecip = __dmul_rn(s.pref, K::block_total(acc, sh.red));  // (uncontracted, as in every kernel)
    if (parts) {
        const double J = 0.5 * K::block_total(K::jastrow_row(c, sh, T, x, y, z), sh.red);
        if (threadIdx.x == 0) {
            double *q = parts + (size_t)w * kPartCount;
            q[0] = kinetic;
            q[1] = vr;
            q[2] = vrecip;
            q[3] = kinetic + vr + vrecip + s.e_const;
            q[4] = st.logdet[2 * w + 0];
            q[5] = ND > 0 ? st.logdet[2 * w + 1] : 0.0;
            q[6] = J;
        }
        return;
    }
    if (threadIdx.x == 0) {
        double *t = st.stats + (size_t)w * kStatCount;
        const double en = kinetic + vr + vrecip + s.e_const;
        const double pot = vr + vrecip + s.e_const;
        t[0] += 1.0;
        t[1] += en;
        t[2] += en * en;
        t[3] += kinetic;
        t[4] += pot;
        t[5] = en;
        t[13] = kinetic;
        t[14] = pot;
        t[8] += en;
        t[9] += 1.0;
        if (t[9] >= (double)block_size) {
            const double bm = t[8] / (double)block_size;
            t[10] += 1.0;
            const double delta = bm - t[11];
            t[11] += delta / t[10];
            t[12] += delta * (bm - t[11]);
            t[8] = 0.0;
            t[9] = 0.0;
        }
    }
}

// Waves per SIMD the sweep kernel is compiled for.  57 + 57 and 81 + 81 in
// FP32 take four (at most 128 VGPRs, 16 and 32 B of spill per lane): eight
// and four walkers per CU, so 2048 and 1024 walkers are resident at once.
// 123 + 123 takes three (168 VGPRs, 32 B of spill); 147 + 147 stays at two
// (three would spill 144 B of its inverse row per lane and run at half speed).
template <typename Real, int NU, int ND, int WPS>
constexpr int rrw_sweep_waves() {
    return (sizeof(Real) == 4 && ((NU == 57 && ND == 57) || (NU == 81 && ND == 81))) ? 4
           : (NU == 123 && ND == 123) || (sizeof(Real) == 8 && NU == 57 && ND == 57) ? 3 : 1;
}

template <typename Real, int NU, int ND, int WPS>
__global__ __launch_bounds__((ND > 0 ? 2 : 1) * WPS * 64, (rrw_sweep_waves<Real, NU, ND, WPS>())) void k_sweep_rrw(DevSys s, DevState st,
                                                                          unsigned long long sweep0,
                                                                          int flags) {
    sweep_rrw_body<Real, NU, ND, WPS>(s, st, blockIdx.x, sweep0, flags);
}

template <typename Real, int NU, int ND, int WPS>
constexpr int rrw_measure_waves() {
    if constexpr (sizeof(Real) == 4 && ((NU == 33 && ND == 33) || (NU == 147 && ND == 147))) return 5;
    return (sizeof(Real) == 4 && ((NU == 81 && ND == 81) || (NU == 123 && ND == 123))) ? 4 : 1;
}

template <typename Real, int NU, int ND, int WPS>
__global__ __launch_bounds__((ND > 0 ? 2 : 1) * WPS * 64, (rrw_measure_waves<Real, NU, ND, WPS>())) void k_measure_rrw(DevSys s, DevState st,
                                                                            int block_size) {
    measure_rrw_body<Real, NU, ND, WPS>(s, st, blockIdx.x, block_size);
}

// The sweep with the caller's moves, each recorded (parity checks of the
// timed move code).
template <typename Real, int NU, int ND, int WPS>
__global__ __launch_bounds__((ND > 0 ? 2 : 1) * WPS * 64, (rrw_sweep_waves<Real, NU, ND, WPS>())) void k_sweep_rrw_given(
    DevSys s, DevState st, int flags, SweepGiven given) {
    sweep_rrw_body<Real, NU, ND, WPS, true>(s, st, blockIdx.x, 0, flags, given);
}

// The local energy's parts per walker, statistics untouched.
template <typename Real, int NU, int ND, int WPS>
__global__ __launch_bounds__((ND > 0 ? 2 : 1) * WPS * 64) void k_measure_rrw_parts(DevSys s, DevState st,
                                                                                  double *parts) {
    measure_rrw_body<Real, NU, ND, WPS>(s, st, blockIdx.x, 1, parts);
}

// Batch forms over a group of systems of one size (as k_sweep_rr_b).
template <typename Real, int NU, int ND, int WPS>
__global__ __launch_bounds__((ND > 0 ? 2 : 1) * WPS * 64, (rrw_sweep_waves<Real, NU, ND, WPS>())) void k_sweep_rrw_b(
    const BatchItem *__restrict__ items, int n_items, const unsigned long long *__restrict__ sweep_ctr,
    int flags) {
    const BatchItem &it = items[batch_find(items, n_items, blockIdx.x)];
    sweep_rrw_body<Real, NU, ND, WPS>(it.s, it.st, blockIdx.x - it.walker_begin, *sweep_ctr, flags);
}

template <typename Real, int NU, int ND, int WPS>
__global__ __launch_bounds__((ND > 0 ? 2 : 1) * WPS * 64, (rrw_measure_waves<Real, NU, ND, WPS>())) void k_measure_rrw_b(
    const BatchItem *__restrict__ items, int n_items, int block_size) {
    const BatchItem &it = items[batch_find(items, n_items, blockIdx.x)];
    measure_rrw_body<Real, NU, ND, WPS>(it.s, it.st, blockIdx.x - it.walker_begin, block_size);
}

}  // namespace vmc
