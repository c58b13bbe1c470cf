// vmc_walker.hpp -- per-walker device routines shared by the sweep, init and
// frozen-configuration kernels.  One wavefront runs one walker; every
// routine is called by all 64 lanes with wave-uniform arguments.
#pragma once

#include "vmc_device.hpp"

namespace vmc {

// Per-walker working set.  Positions, Jastrow row sums and draws live in
// LDS; the transposed inverses live in LDS when they fit, otherwise in the
// walker's HBM block (inv points there).
template <typename Real>
struct WalkerView {
    Real *inv;       // [ninv] invT per determinant, invT[k][j] = (D^-1)[j][k]
    double *px, *py, *pz;  // [N]
    double *U;       // [N] Jastrow row sums U_i = sum_{j != i} u(r_ij)
    double *uj;      // [N] pair values of the pending move
    double *draw;    // [4][N] proposal (dx, dy, dz) and accept uniform
    Real *row;       // [nmax] orbital row of the pending move
    Real *col;       // [nmax] copy of inverse row p
    double *gjcol;   // [nmax] Gauss-Jordan column scratch
    int *perm;       // [nmax] Gauss-Jordan pivot rows
    double *A;       // [nmax^2] FP64 re-inversion workspace (HBM)
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// LDS bytes of one walker; the inverse block is included when inv_in_lds.
__host__ __device__ inline size_t walker_lds_bytes(int N, int nmax, int ninv, int real_bytes,
                                                   bool inv_in_lds) {
    size_t b = 0;
    if (inv_in_lds) b += align16((size_t)ninv * real_bytes);
    b += align16((size_t)N * 8) * 4;        // px py pz U
    b += align16((size_t)N * 8);            // uj
    b += align16((size_t)N * 8) * 4;        // draws
    b += align16((size_t)nmax * real_bytes) * 2;  // row col
    b += align16((size_t)nmax * 8);         // gjcol
    b += align16((size_t)nmax * 4);         // perm
    return b;
}

template <typename Real>
__device__ inline WalkerView<Real> carve(unsigned char *smem, const DevSys &s, bool inv_in_lds,
                                         Real *inv_hbm, double *A) {
    WalkerView<Real> v;
    size_t off = 0;
    if (inv_in_lds) {
        v.inv = reinterpret_cast<Real *>(smem + off);
        off += align16((size_t)s.ninv * sizeof(Real));
    } else {
        v.inv = inv_hbm;
    }
    const size_t nb = align16((size_t)s.N * 8);
    v.px = reinterpret_cast<double *>(smem + off); off += nb;
    v.py = reinterpret_cast<double *>(smem + off); off += nb;
    v.pz = reinterpret_cast<double *>(smem + off); off += nb;
    v.U = reinterpret_cast<double *>(smem + off); off += nb;
    v.uj = reinterpret_cast<double *>(smem + off); off += nb;
    v.draw = reinterpret_cast<double *>(smem + off); off += 4 * nb;
    const size_t rb = align16((size_t)s.nmax * sizeof(Real));
    v.row = reinterpret_cast<Real *>(smem + off); off += rb;
    v.col = reinterpret_cast<Real *>(smem + off); off += rb;
    v.gjcol = reinterpret_cast<double *>(smem + off); off += align16((size_t)s.nmax * 8);
    v.perm = reinterpret_cast<int *>(smem + off);
    v.A = A;
    return v;
}

__device__ inline int spin_of(const DevSys &s, int i) { return (s.ndet == 2 && i >= s.n_up) ? 1 : 0; }

template <typename Real>
__device__ inline Real pair_a(const DevSys &s, int i, int j) {
    return (Real)(spin_of(s, i) == spin_of(s, j) ? s.a_same : s.a_anti);
}

// Minimum-image distance of (x, y, z) to electron j.
template <typename Real>
__device__ inline Real pair_dist(const DevSys &s, double x, double y, double z,
                                 const WalkerView<Real> &v, int j, Real *dx, Real *dy, Real *dz) {
    const Real ex = (Real)min_image(x - v.px[j], s);
    const Real ey = (Real)min_image(y - v.py[j], s);
    const Real ez = (Real)min_image(z - v.pz[j], s);
    if (dx) { *dx = ex; *dy = ey; *dz = ez; }
    return sqrt(ex * ex + ey * ey + ez * ez);
}

// Orbital j of determinant d at fractional position u: cos or sin of 2 pi n.u.
template <typename Real>
__device__ inline Real orbital(const DevSys &s, int d, int j, double ux, double uy, double uz) {
    const int4 o = s.orb[d * s.nmax + j];
    Real sn, cs;
    sincos_2pi<Real>(frac_phase(o.x, o.y, o.z, ux, uy, uz), &sn, &cs);
    return o.w ? sn : cs;
}

// ----------------------------------------------------- FP64 re-inversion
// Builds A = D^T (A[j][i] = phi_j(r_i)) of determinant d in FP64 and inverts
// it in place by Gauss-Jordan with partial (row) pivoting, so A ends as
// (D^T)^-1 = invT.  Returns log|det D| = sum log|pivot|, or -inf when a
// pivot falls below the reference's 1e-12 tolerance (matrix.hpp:35-39).
// Replaces log_abs_det's LU + n solves (slater_plane_wave.cpp:202-318).
// With `drift` non-null, the largest |I - A invT|, the drift of the inverse
// being replaced, is computed first (FP64, before the Gauss-Jordan) and
// returned there.
template <typename Real>
__device__ double reinvert(const DevSys &s, int d, const WalkerView<Real> &v, int lane,
                           double *drift = nullptr) {
    const int n = s.n[d];
    const int first = s.first[d];
    double *A = v.A;
    for (int e = lane; e < n * n; e += kWave) {
        const int j = e / n, i = e - (e / n) * n;
        const int4 o = s.orb[d * s.nmax + j];
        const double ux = v.px[first + i] * s.inv_L;
        const double uy = v.py[first + i] * s.inv_L;
        const double uz = v.pz[first + i] * s.inv_L;
        double sn, cs;
        sincos_2pi<double>(frac_phase(o.x, o.y, o.z, ux, uy, uz), &sn, &cs);
        A[e] = o.w ? sn : cs;
    }
    wave_sync();
    if (drift) {
        const Real *Y = v.inv + s.inv_off[d];
        double dmax = 0.0;
        for (int e = lane; e < n * n; e += kWave) {
            const int m = e / n, mp = e - (e / n) * n;
            double z = 0.0;
            for (int i = 0; i < n; ++i) z += A[m * n + i] * (double)Y[i * n + mp];
            const double r = fabs((m == mp ? 1.0 : 0.0) - z);
            dmax = (r > dmax || r != r) ? r : dmax;
        }
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const double o = __shfl_xor(dmax, off, kWave);
            dmax = (o > dmax || o != o) ? o : dmax;
        }
        *drift = wave_uniform(dmax);
    }
    double logdet = 0.0;
    bool singular = false;
    for (int c = 0; c < n; ++c) {
        double best = -1.0;
        int bi = n;
        for (int r = c + lane; r < n; r += kWave) {
            const double a = fabs(A[r * n + c]);
            if (a > best) { best = a; bi = r; }
        }
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, kWave);
            const int oi = __shfl_xor(bi, off, kWave);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        best = wave_uniform(best);
        bi = wave_uniform(bi);
        if (!(best >= 1e-12)) { singular = true; break; }
        if (lane == 0) v.perm[c] = bi;
        if (bi != c) {
            for (int j = lane; j < n; j += kWave) {
                const double t = A[c * n + j];
                A[c * n + j] = A[bi * n + j];
                A[bi * n + j] = t;
            }
        }
        wave_sync();
        const double piv = A[c * n + c];
        logdet += log(fabs(piv));
        const double pinv = 1.0 / piv;
        for (int r = lane; r < n; r += kWave) v.gjcol[r] = A[r * n + c];
        wave_sync();
        for (int j = lane; j < n; j += kWave) A[c * n + j] = (j == c ? 1.0 : A[c * n + j]) * pinv;
        wave_sync();
        for (int e = lane; e < n * n; e += kWave) {
            const int r = e / n, j = e - (e / n) * n;
            if (r == c) continue;
            const double f = v.gjcol[r];
            A[e] = (j == c ? 0.0 : A[e]) - f * A[c * n + j];
        }
        wave_sync();
    }
    if (singular) return -INFINITY;
    for (int c = n - 1; c >= 0; --c) {
        const int q = v.perm[c];
        if (q != c) {
            for (int r = lane; r < n; r += kWave) {
                const double t = A[r * n + c];
                A[r * n + c] = A[r * n + q];
                A[r * n + q] = t;
            }
        }
        wave_sync();
    }
    Real *invs = v.inv + s.inv_off[d];
    for (int e = lane; e < n * n; e += kWave) invs[e] = (Real)A[e];
    wave_sync();
    return logdet;
}

// U_i for every electron (jastrow_pade.cpp:7-55 row sums; J = sum_i U_i / 2).
template <typename Real>
__device__ inline void jastrow_rows(const DevSys &s, const WalkerView<Real> &v, int lane) {
    const Real b = (Real)s.b;
    for (int i = lane; i < s.N; i += (int)blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < s.N; ++j) {
            if (j == i) continue;
            const Real r = pair_dist<Real>(s, v.px[i], v.py[i], v.pz[i], v, j, nullptr, nullptr,
                                           nullptr);
            acc += (double)pade_u<Real>(r, pair_a<Real>(s, i, j), b);
        }
        v.U[i] = acc;
    }
    wave_sync();
}

struct Measured {
    double kinetic, vreal, vrecip, e_local;
};

// One measured local energy into a walker's FP64 statistics t[kStatCount]
// and its blocking state: block means of block_size samples, Welford
// over the block means (blocking_analysis.cpp:17-39).
__device__ inline void accumulate_measure(double *t, const Measured &m, double e_const, int block_size) {
    const double en = m.e_local;
    const double pot = m.vreal + m.vrecip + e_const;
    t[0] += 1.0;
    t[1] += en;
    t[2] += en * en;
    t[3] += m.kinetic;
    t[4] += pot;
    t[5] = en;
    t[13] = m.kinetic;
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

// Local energy of the current configuration (energy_tracking.cpp:253-382):
// full Jastrow and Slater derivatives (jastrow_pade.cpp:57-138,
// slater_plane_wave.cpp:425-499), real-space Ewald over all pairs and the
// structure factor rebuilt from positions.  Optional outputs: grad/lap of
// ln psi per electron ([N][4]) and S_G ([ng][2]).
template <typename Real>
__device__ Measured measure(const DevSys &s, const WalkerView<Real> &v, int lane, double *gl_out,
                            double *sk_out) {
    const int N = s.N;
    const Real b = (Real)s.b;
    const Real ku = (Real)s.kunit;
    double kin = 0.0;
    for (int i = lane; i < N; i += (int)blockDim.x) {
        double gx = 0, gy = 0, gz = 0, lp = 0;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            Real dx, dy, dz;
            const Real r = pair_dist<Real>(s, v.px[i], v.py[i], v.pz[i], v, j, &dx, &dy, &dz);
            if (r < (Real)1e-12) continue;
            const Real a = pair_a<Real>(s, i, j);
            const Real den = (Real)1 / ((Real)1 + b * r);
            const Real den2 = den * den;
            const Real d1 = a * den2;
            const Real d2 = (Real)-2 * a * b * den2 * den;
            const Real inv_r = (Real)1 / r;
            const Real gf = d1 * inv_r;
            gx += (double)(gf * dx);
            gy += (double)(gf * dy);
            gz += (double)(gf * dz);
            lp += (double)(d2 + (Real)2 * d1 * inv_r);
        }
        const int d = spin_of(s, i);
        const int n = s.n[d];
        const int il = i - s.first[d];
        const Real *invs = v.inv + s.inv_off[d];
        const double ux = v.px[i] * s.inv_L, uy = v.py[i] * s.inv_L, uz = v.pz[i] * s.inv_L;
        double sx = 0, sy = 0, sz = 0, sl = 0;
        for (int m = 0; m < n; ++m) {
            const int4 o = s.orb[d * s.nmax + m];
            Real sn, cs;
            sincos_2pi<Real>(frac_phase(o.x, o.y, o.z, ux, uy, uz), &sn, &cs);
            const Real wgt = invs[il * n + m];
            const Real phi = o.w ? sn : cs;
            const Real gfac = o.w ? cs : -sn;
            const Real wg = wgt * gfac * ku;
            sx += (double)(wg * (Real)o.x);
            sy += (double)(wg * (Real)o.y);
            sz += (double)(wg * (Real)o.z);
            sl -= (double)(wgt * phi) * (double)(ku * ku) * (double)(o.x * o.x + o.y * o.y + o.z * o.z);
        }
        const double slap = sl - (sx * sx + sy * sy + sz * sz);
        const double tx = sx + gx, ty = sy + gy, tz = sz + gz, tl = slap + lp;
        kin += tl + tx * tx + ty * ty + tz * tz;
        if (gl_out) {
            gl_out[4 * i + 0] = tx;
            gl_out[4 * i + 1] = ty;
            gl_out[4 * i + 2] = tz;
            gl_out[4 * i + 3] = tl;
        }
    }
    Measured m;
    m.kinetic = -0.5 * block_sum(kin);
    double vr = 0.0;
    for (int i = lane; i < N; i += (int)blockDim.x) {
        for (int j = i + 1; j < N; ++j) {
            const Real r = pair_dist<Real>(s, v.px[i], v.py[i], v.pz[i], v, j, nullptr, nullptr,
                                           nullptr);
            vr += (double)ewald_pair<Real>(r, s);
        }
    }
    m.vreal = block_sum(vr);
    double vq = 0.0;
    for (int g = lane; g < s.ng; g += (int)blockDim.x) {
        const int4 G = s.gvec[g];
        double sre = 0.0, sim = 0.0;
        for (int i = 0; i < N; ++i) {
            Real sn, cs;
            sincos_2pi<Real>(frac_phase(G.x, G.y, G.z, v.px[i] * s.inv_L, v.py[i] * s.inv_L,
                                        v.pz[i] * s.inv_L),
                             &sn, &cs);
            sre += (double)cs;
            sim += (double)sn;
        }
        vq += s.gw[g] * (sre * sre + sim * sim);
        if (sk_out) {
            sk_out[2 * g + 0] = sre;
            sk_out[2 * g + 1] = sim;
        }
    }
    // (an uncontracted product: the sums callers form with it round the same
    // way in every kernel this is inlined into, batch and alone)
    m.vrecip = __dmul_rn(s.pref, block_sum(vq));
    m.e_local = m.kinetic + m.vreal + m.vrecip + s.e_const;
    return m;
}

struct Proposal {
    double ratio;   // determinant ratio R
    double jnew;    // sum_j u(|r' - r_j|)
    double dj;      // Jnew - U_p
};

// Trial move of electron p to (x, y, z): orbital row, ratio against invT row
// p (slater_plane_wave.cpp:172-200,320-365) and the Jastrow change
// (jastrow_pade.cpp:140-200, here as the new-pair sum minus the cached
// row sum U_p).  Leaves the row in v.row and the new pair terms in v.uj.
template <typename Real>
__device__ inline Proposal propose(const DevSys &s, const WalkerView<Real> &v, int p, double x,
                                   double y, double z, int lane) {
    const int d = spin_of(s, p);
    const int n = s.n[d];
    const int pl = p - s.first[d];
    const Real *invs = v.inv + s.inv_off[d];
    const double ux = x * s.inv_L, uy = y * s.inv_L, uz = z * s.inv_L;
    double rp = 0.0;
    for (int j = lane; j < n; j += (int)blockDim.x) {
        const Real phi = orbital<Real>(s, d, j, ux, uy, uz);
        v.row[j] = phi;
        rp += (double)phi * (double)invs[pl * n + j];
    }
    const Real b = (Real)s.b;
    double jp = 0.0;
    for (int j = lane; j < s.N; j += (int)blockDim.x) {
        double u = 0.0;
        if (j != p) {
            const Real r = pair_dist<Real>(s, x, y, z, v, j, nullptr, nullptr, nullptr);
            u = (double)pade_u<Real>(r, pair_a<Real>(s, p, j), b);
        }
        v.uj[j] = u;
        jp += u;
    }
    Proposal pr;
    pr.ratio = block_sum(rp);
    pr.jnew = block_sum(jp);
    pr.dj = pr.jnew - v.U[p];
    wave_sync();
    return pr;
}

// Commits a proposed move (simulation.cpp:133-145): Jastrow row sums,
// Sherman-Morrison on invT (slater_plane_wave.cpp:367-423) and the position.
template <typename Real>
__device__ inline void accept_move(const DevSys &s, const WalkerView<Real> &v, int p, double x,
                                   double y, double z, const Proposal &pr, int lane) {
    const int d = spin_of(s, p);
    const int n = s.n[d];
    const int pl = p - s.first[d];
    Real *invs = v.inv + s.inv_off[d];
    const Real b = (Real)s.b;
    const double ox = v.px[p], oy = v.py[p], oz = v.pz[p];
    for (int j = lane; j < s.N; j += (int)blockDim.x) {
        if (j == p) continue;
        const Real r = pair_dist<Real>(s, ox, oy, oz, v, j, nullptr, nullptr, nullptr);
        v.U[j] += v.uj[j] - (double)pade_u<Real>(r, pair_a<Real>(s, p, j), b);
    }
    for (int j = lane; j < n; j += (int)blockDim.x) v.col[j] = invs[pl * n + j];
    wave_sync();
    const Real inv_ratio = (Real)(1.0 / pr.ratio);
    for (int k = lane; k < n; k += (int)blockDim.x) {
        if (k == pl) continue;
        double sk = 0.0;
        for (int m = 0; m < n; ++m) sk += (double)v.row[m] * (double)invs[k * n + m];
        const Real f = (Real)sk * inv_ratio;
        for (int j = 0; j < n; ++j) invs[k * n + j] -= v.col[j] * f;
    }
    wave_sync();
    for (int j = lane; j < n; j += (int)blockDim.x) invs[pl * n + j] = v.col[j] * inv_ratio;
    if (lane == 0) {
        v.U[p] = pr.jnew;
        v.px[p] = x;
        v.py[p] = y;
        v.pz[p] = z;
    }
    wave_sync();
}

}  // namespace vmc
