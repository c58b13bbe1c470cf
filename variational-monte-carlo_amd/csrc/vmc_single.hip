// vmc_single.hip -- one configuration's determinant, Jastrow and Ewald
// objects on the device, driven by the reference's per-move call sequence.
//
// These back the host facade's SlaterPlaneWave, JastrowPade and EnergyTracker
// classes (include/vmc_hip.h, "per-move objects"): the state the reference
// keeps in its class members lives in device memory, and each call is one
// single-workgroup launch (256 threads) on the object's own stream, followed
// by the copy of its result.  Everything is FP64 with the reference's
// physics: one plane-wave determinant over all n particles, one Pade a,
// alpha = 6/L, the A&S 7.1.26 erfc.  Layout in device memory:
//   vmc_slater  k [3][nk], orbital -> k index and type [n], sin and cos cache
//               [n][TS], saved trig row [2][nk], D, LU and the transposed
//               inverse [n][S] (S, TS: n and nk padded to 8 doubles, the
//               reference's strides), pivots [n], a row [n], positions
//               [3][n], grad/lap [4][n], scalars
//   vmc_pade    positions [3][n], grad/lap [4][n], scalars (grown on demand)
//   vmc_ewald   G [3][ng], weights [ng], S_G re/im [ng], positions and
//               grad/lap as above, scalars {V_real, V_recip}
#include "../../include/vmc_hip.h"
#include "vmc_device.hpp"
#include "vmc_tables.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <new>
#include <numbers>
#include <vector>

namespace vmc {
int set_last_error(int code, const char *msg);
}

namespace {

using vmc::block_argmax;
using vmc::block_sum;
using vmc::wave_sum;

constexpr int kT = 256;  // threads of every launch in this file

size_t pad8(size_t n) { return (n + 7) & ~size_t(7); }

int err(int code, const char *msg) { return vmc::set_last_error(code, msg); }

#define SINGLE_TRY(expr)                                                          \
    do {                                                                          \
        if ((expr) != hipSuccess) return err(VMC_ERR_DEVICE, "device call failed: " #expr); \
    } while (0)

// Minimum image as the reference writes it: d += L [d <= -L/2] - L [d > L/2]
// (jastrow_pade.cpp:36-38, energy_tracking.cpp:120-122).
__device__ inline double wrap(double d, double L) {
    return d + ((d <= -0.5 * L ? L : 0.0) + (d > 0.5 * L ? -L : 0.0));
}

struct Pair {
    double dx, dy, dz, r;
};

__device__ inline Pair pair(double xi, double yi, double zi, double xj, double yj, double zj, double L) {
    Pair p;
    p.dx = wrap(xi - xj, L);
    p.dy = wrap(yi - yj, L);
    p.dz = wrap(zi - zj, L);
    p.r = sqrt(p.dx * p.dx + p.dy * p.dy + p.dz * p.dz);
    return p;
}

// erfc(alpha r) / r with the reference's guard (energy_tracking.cpp:125-137).
__device__ inline double ewald_term(double r, double alpha) {
    const double inv_r = r < 1e-12 ? 1.0 : 1.0 / r;
    return vmc::erfc_as7126<double>(alpha * r) * inv_r;
}

// Pade pair derivatives (jastrow_pade.cpp:96-118): grad factor u'/r and
// laplacian u'' + 2u'/r, zero for the particle itself and at r < 1e-12.
struct PadeDeriv {
    double g, l;
};

__device__ inline PadeDeriv pade_deriv(double r, bool self, double a, double b) {
    const bool degenerate = r < 1e-12;
    const double inv_r = degenerate ? 1.0 : 1.0 / r;
    const double mask = (degenerate || self) ? 0.0 : 1.0;
    const double den = 1.0 / (1.0 + b * r);
    const double den2 = den * den;
    const double d1 = a * den2;
    const double d2 = -2.0 * a * b * den2 * den;
    return {mask * d1 * inv_r, mask * (d2 + 2.0 * d1 * inv_r)};
}

// ------------------------------------------------------------ determinant
struct SlaterDev {
    int n, S, nk, TS;
    const double *k;        // [3][nk]
    const int *kidx, *type; // [n]
    double *sinc, *cosc;    // [n][TS]
    double *saved;          // [2][nk]
    double *D, *LU, *INV;   // [n][S]
    int *piv;               // [n]
    double *row;            // [n]
    const double *pos;      // [3][n]
    double *gl;             // [4][n]
    double *out;            // scalars
};

__device__ inline void trig_row(const SlaterDev &s, int p, double x, double y, double z) {
    for (int k = threadIdx.x; k < s.nk; k += kT) {
        const double dot = s.k[k] * x + s.k[s.nk + k] * y + s.k[2 * s.nk + k] * z;
        double sn, cs;
        sincos(dot, &sn, &cs);
        s.sinc[p * s.TS + k] = sn;
        s.cosc[p * s.TS + k] = cs;
    }
}

// phi_j(r_p) from the trig cache: cos + type (sin - cos) (slater_plane_wave.cpp:247-256).
__device__ inline double orbital_of(const SlaterDev &s, int p, int j) {
    const int k = s.kidx[j];
    const double c = s.cosc[p * s.TS + k], sn = s.sinc[p * s.TS + k];
    return c + static_cast<double>(s.type[j]) * (sn - c);
}

// log_abs_det: trig cache and D for every particle, partial-pivot LU with
// the reference's tie rule (first row of the largest |value|) and pivot
// tolerance, log|det| and, when finite, inverse columns by LU solves
// (slater_plane_wave.cpp:202-318, matrix.hpp:13-96).
__global__ void __launch_bounds__(kT) k_slater_log_det(SlaterDev s) {
    __shared__ double red[32];
    __shared__ int skip;
    const int n = s.n, S = s.S, t = threadIdx.x;
    for (int p = 0; p < n; ++p) trig_row(s, p, s.pos[p], s.pos[n + p], s.pos[2 * n + p]);
    __syncthreads();
    for (int e = t; e < n * n; e += kT) {
        const int p = e / n, j = e % n;
        const double v = orbital_of(s, p, j);
        s.D[p * S + j] = v;
        s.LU[p * S + j] = v;
    }
    for (int r = t; r < n; r += kT) s.piv[r] = r;
    __syncthreads();
    for (int col = 0; col < n; ++col) {
        double best = -1.0;
        int brow = n;
        for (int r = col + t; r < n; r += kT) {
            const double v = fabs(s.LU[r * S + col]);
            if (v > best) { best = v; brow = r; }
        }
        block_argmax(best, brow, red);
        if (t == 0) {
            skip = best < 1e-12;
            if (skip) s.LU[col * S + col] = 0.0;
        }
        __syncthreads();
        if (skip) continue;
        if (brow != col) {
            for (int c = t; c < n; c += kT) {
                const double a = s.LU[col * S + c];
                s.LU[col * S + c] = s.LU[brow * S + c];
                s.LU[brow * S + c] = a;
            }
            if (t == 0) {
                const int a = s.piv[col];
                s.piv[col] = s.piv[brow];
                s.piv[brow] = a;
            }
            __syncthreads();
        }
        const double pv = s.LU[col * S + col];
        for (int r = col + 1 + t; r < n; r += kT) s.LU[r * S + col] /= pv;
        __syncthreads();
        const int w = n - col - 1;
        for (int e = t; e < w * w; e += kT) {
            const int r = col + 1 + e / w, c = col + 1 + e % w;
            s.LU[r * S + c] -= s.LU[r * S + col] * s.LU[col * S + c];
        }
        __syncthreads();
    }
    double ld = 0.0;
    for (int i = t; i < n; i += kT) ld += log(fabs(s.LU[i * S + i]));
    ld = block_sum(ld);
    const bool finite = isfinite(ld);
    if (t == 0) s.out[0] = finite ? ld : -std::numeric_limits<double>::infinity();
    if (!finite) return;
    // Column c of the inverse (row c of the transposed inverse): solve
    // LU x = P e_c, x kept in INV[c][*].
    for (int c = t; c < n; c += kT) {
        double *x = s.INV + c * S;
        for (int r = 0; r < n; ++r) x[r] = s.piv[r] == c ? 1.0 : 0.0;
        for (int r = 0; r < n; ++r) {
            double v = x[r];
            for (int q = 0; q < r; ++q) v -= s.LU[r * S + q] * x[q];
            x[r] = v;
        }
        for (int r = n - 1; r >= 0; --r) {
            double v = x[r];
            for (int q = r + 1; q < n; ++q) v -= s.LU[r * S + q] * x[q];
            x[r] = v / s.LU[r * S + r];
        }
    }
}

// op 0: update_trig_cache of particle p at (x, y, z); 1: save_trig_row;
// 2: restore_trig_row; 3: build_row into s.row (slater_plane_wave.cpp:154-200,320-349).
__global__ void __launch_bounds__(kT) k_slater_row_op(SlaterDev s, int op, int p, double x, double y, double z) {
    const int t = threadIdx.x;
    if (op == 0) {
        trig_row(s, p, x, y, z);
    } else if (op == 1) {
        for (int k = t; k < s.nk; k += kT) {
            s.saved[k] = s.sinc[p * s.TS + k];
            s.saved[s.nk + k] = s.cosc[p * s.TS + k];
        }
    } else if (op == 2) {
        for (int k = t; k < s.nk; k += kT) {
            s.sinc[p * s.TS + k] = s.saved[k];
            s.cosc[p * s.TS + k] = s.saved[s.nk + k];
        }
    } else {
        for (int j = t; j < s.n; j += kT) s.row[j] = orbital_of(s, p, j);
    }
}

// determinant_ratio: R = sum_j row_j invT[p][j] (slater_plane_wave.cpp:351-365).
__global__ void __launch_bounds__(kT) k_slater_ratio(SlaterDev s, int p) {
    double v = 0.0;
    for (int j = threadIdx.x; j < s.n; j += kT) v += s.row[j] * s.INV[p * s.S + j];
    v = block_sum(v);
    if (threadIdx.x == 0) s.out[1] = v;
}

// accept_move: invT[k] -= invT[p] (row . invT[k]) / R for k != p, invT[p] /= R,
// D[p] = row (slater_plane_wave.cpp:367-423).  One wavefront per row k; the
// old invT[p] is staged in LDS first, so no row is read after it changes.
__global__ void __launch_bounds__(kT) k_slater_accept(SlaterDev s, int p, double ratio) {
    extern __shared__ double col[];
    const int n = s.n, S = s.S, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const double inv_r = 1.0 / ratio;
    for (int j = t; j < n; j += kT) col[j] = s.INV[p * S + j];
    __syncthreads();
    for (int k = wv; k < n; k += kT / 64) {
        double *rk = s.INV + k * S;
        if (k == p) {
            for (int j = lane; j < n; j += 64) rk[j] = col[j] * inv_r;
            continue;
        }
        double sk = 0.0;
        for (int m = lane; m < n; m += 64) sk += s.row[m] * rk[m];
        const double f = wave_sum(sk) * inv_r;
        for (int j = lane; j < n; j += 64) rk[j] -= col[j] * f;
    }
    for (int j = t; j < n; j += kT) s.D[p * S + j] = s.row[j];
}

// add_derivatives (slater_plane_wave.cpp:425-499): one wavefront per
// particle, lanes over orbitals; grad += g, lap += (sum invT k^2 phi'') - |g|^2.
__global__ void __launch_bounds__(kT) k_slater_derivs(SlaterDev s) {
    const int n = s.n, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int p = wv; p < n; p += kT / 64) {
        double gx = 0.0, gy = 0.0, gz = 0.0, lp = 0.0;
        for (int j = lane; j < n; j += 64) {
            const int k = s.kidx[j];
            const double kx = s.k[k], ky = s.k[s.nk + k], kz = s.k[2 * s.nk + k];
            const double k2 = kx * kx + ky * ky + kz * kz;
            const double ty = static_cast<double>(s.type[j]);
            const double c = s.cosc[p * s.TS + k], sn = s.sinc[p * s.TS + k];
            const double gf = -sn + ty * (sn + c), lf = -c + ty * (c - sn);
            const double w = s.INV[p * s.S + j];
            gx += w * kx * gf;
            gy += w * ky * gf;
            gz += w * kz * gf;
            lp += w * k2 * lf;
        }
        gx = wave_sum(gx);
        gy = wave_sum(gy);
        gz = wave_sum(gz);
        lp = wave_sum(lp);
        if (lane == 0) {
            s.gl[p] += gx;
            s.gl[n + p] += gy;
            s.gl[2 * n + p] += gz;
            s.gl[3 * n + p] += lp - (gx * gx + gy * gy + gz * gz);
        }
    }
}

// ------------------------------------------------------------ Pade Jastrow
struct PadeDev {
    int n;
    double L, a, b;
    const double *pos;  // [3][n]
    double *gl;         // [4][n]
    double *out;
};

__device__ inline Pair pade_pair(const PadeDev &j, int i, int m) {
    const double *x = j.pos, *y = j.pos + j.n, *z = j.pos + 2 * j.n;
    return pair(x[i], y[i], z[i], x[m], y[m], z[m], j.L);
}

// value: J = 1/2 sum_{i != j} u(r_ij) (jastrow_pade.cpp:7-55).
__global__ void __launch_bounds__(kT) k_pade_value(PadeDev j) {
    double v = 0.0;
    for (int e = threadIdx.x; e < j.n * j.n; e += kT) {
        const int i = e / j.n, m = e % j.n;
        if (i == m) continue;
        const double r = pade_pair(j, i, m).r;
        v += j.a * r * (1.0 / (1.0 + j.b * r));
    }
    v = block_sum(v);
    if (threadIdx.x == 0) j.out[0] = 0.5 * v;
}

// add_derivatives (jastrow_pade.cpp:57-138): thread per particle i.
__global__ void __launch_bounds__(kT) k_pade_derivs(PadeDev j) {
    const int n = j.n;
    for (int i = threadIdx.x; i < n; i += kT) {
        double gx = 0.0, gy = 0.0, gz = 0.0, lp = 0.0;
        for (int m = 0; m < n; ++m) {
            const Pair q = pade_pair(j, i, m);
            const PadeDeriv d = pade_deriv(q.r, i == m, j.a, j.b);
            gx += d.g * q.dx;
            gy += d.g * q.dy;
            gz += d.g * q.dz;
            lp += d.l;
        }
        j.gl[i] += gx;
        j.gl[n + i] += gy;
        j.gl[2 * n + i] += gz;
        j.gl[3 * n + i] += lp;
    }
}

// delta_value and update_derivatives_for_move (jastrow_pade.cpp:140-317):
// positions hold the moved particle at its new place, (ox, oy, oz) is the old
// one.  Thread per partner m; the moved particle's own row is the sum over
// partners of the new pairs.
__global__ void __launch_bounds__(kT) k_pade_move(PadeDev j, int moved, double ox, double oy, double oz,
                                                  int update) {
    const int n = j.n;
    const double *x = j.pos, *y = j.pos + n, *z = j.pos + 2 * n;
    double dv = 0.0, mx = 0.0, my = 0.0, mz = 0.0, ml = 0.0;
    for (int m = threadIdx.x; m < n; m += kT) {
        const bool self = m == moved;
        const Pair po = pair(ox, oy, oz, x[m], y[m], z[m], j.L);
        const Pair pn = pair(x[moved], y[moved], z[moved], x[m], y[m], z[m], j.L);
        if (!self) dv += j.a * (pn.r * (1.0 / (1.0 + j.b * pn.r)) - po.r * (1.0 / (1.0 + j.b * po.r)));
        if (update) {
            const PadeDeriv d_o = pade_deriv(po.r, self, j.a, j.b);
            const PadeDeriv d_n = pade_deriv(pn.r, self, j.a, j.b);
            mx += d_n.g * pn.dx;
            my += d_n.g * pn.dy;
            mz += d_n.g * pn.dz;
            ml += d_n.l;
            if (!self) {
                j.gl[m] += d_o.g * po.dx - d_n.g * pn.dx;
                j.gl[n + m] += d_o.g * po.dy - d_n.g * pn.dy;
                j.gl[2 * n + m] += d_o.g * po.dz - d_n.g * pn.dz;
                j.gl[3 * n + m] += d_n.l - d_o.l;
            }
        }
    }
    dv = block_sum(dv);
    if (update) {
        mx = block_sum(mx);
        my = block_sum(my);
        mz = block_sum(mz);
        ml = block_sum(ml);
    }
    if (threadIdx.x == 0) {
        j.out[0] = dv;
        if (update) {
            j.gl[moved] = mx;
            j.gl[n + moved] = my;
            j.gl[2 * n + moved] = mz;
            j.gl[3 * n + moved] = ml;
        }
    }
}

// ------------------------------------------------------------ Ewald
struct EwaldDev {
    int n, ng;
    double L, alpha, pref;  // pref = 1 / (2 pi L^3)
    const double *g;        // [3][ng]
    const double *w;        // [ng]
    double *sre, *sim;      // [ng]
    const double *pos;      // [3][n]
    const double *gl;       // [4][n]
    double *state;          // {V_real, V_recip}
    double *out;
};

// initialize_structure_factors (energy_tracking.cpp:146-191): thread per G.
__global__ void __launch_bounds__(kT) k_ewald_sf_init(EwaldDev e) {
    const double *x = e.pos, *y = e.pos + e.n, *z = e.pos + 2 * e.n;
    for (int q = threadIdx.x; q < e.ng; q += kT) {
        const double gx = e.g[q], gy = e.g[e.ng + q], gz = e.g[2 * e.ng + q];
        double cs = 0.0, sn = 0.0;
        for (int m = 0; m < e.n; ++m) {
            double s, c;
            sincos(gx * x[m] + gy * y[m] + gz * z[m], &s, &c);
            cs += c;
            sn += s;
        }
        e.sre[q] = cs;
        e.sim[q] = sn;
    }
}

// initialize_reciprocal_energy (energy_tracking.cpp:72-91).
__global__ void __launch_bounds__(kT) k_ewald_recip_init(EwaldDev e) {
    double v = 0.0;
    for (int q = threadIdx.x; q < e.ng; q += kT) v += e.w[q] * (e.sre[q] * e.sre[q] + e.sim[q] * e.sim[q]);
    v = block_sum(v);
    if (threadIdx.x == 0) e.state[1] = e.pref * v;
}

// update_structure_factors (energy_tracking.cpp:193-251): S_G += dS and
// V_recip += pref sum w (2 Re(conj(S) dS) + |dS|^2), S before the update.
__global__ void __launch_bounds__(kT) k_ewald_sf_update(EwaldDev e, double ox, double oy, double oz, double nx,
                                                        double ny, double nz) {
    double v = 0.0;
    for (int q = threadIdx.x; q < e.ng; q += kT) {
        const double gx = e.g[q], gy = e.g[e.ng + q], gz = e.g[2 * e.ng + q];
        double sn, cn, so, co;
        sincos(gx * nx + gy * ny + gz * nz, &sn, &cn);
        sincos(gx * ox + gy * oy + gz * oz, &so, &co);
        const double dr = cn - co, di = sn - so;
        v += e.w[q] * (2.0 * (e.sre[q] * dr + e.sim[q] * di) + dr * dr + di * di);
        e.sre[q] += dr;
        e.sim[q] += di;
    }
    v = block_sum(v);
    if (threadIdx.x == 0) e.state[1] += e.pref * v;
}

// initialize_real_energy (energy_tracking.cpp:93-144): pairs i < j.
__global__ void __launch_bounds__(kT) k_ewald_real_init(EwaldDev e) {
    const double *x = e.pos, *y = e.pos + e.n, *z = e.pos + 2 * e.n;
    double v = 0.0;
    for (int q = threadIdx.x; q < e.n * e.n; q += kT) {
        const int i = q / e.n, m = q % e.n;
        if (m <= i) continue;
        v += ewald_term(pair(x[i], y[i], z[i], x[m], y[m], z[m], e.L).r, e.alpha);
    }
    v = block_sum(v);
    if (threadIdx.x == 0) e.state[0] = v;
}

// update_real_energy (energy_tracking.cpp:280-365).
__global__ void __launch_bounds__(kT) k_ewald_real_update(EwaldDev e, int moved, double ox, double oy, double oz) {
    const double *x = e.pos, *y = e.pos + e.n, *z = e.pos + 2 * e.n;
    double v = 0.0;
    for (int m = threadIdx.x; m < e.n; m += kT) {
        if (m == moved) continue;
        const double ro = pair(ox, oy, oz, x[m], y[m], z[m], e.L).r;
        const double rn = pair(x[moved], y[moved], z[moved], x[m], y[m], z[m], e.L).r;
        v += ewald_term(rn, e.alpha) - ewald_term(ro, e.alpha);
    }
    v = block_sum(v);
    if (threadIdx.x == 0) e.state[0] += v;
}

// eval_total_energy (energy_tracking.cpp:253-278,367-382): kinetic from the
// grad/lap arrays plus the cached potentials and constants.
__global__ void __launch_bounds__(kT) k_ewald_total(EwaldDev e, double constants) {
    double v = 0.0;
    for (int i = threadIdx.x; i < e.n; i += kT) {
        const double gx = e.gl[i], gy = e.gl[e.n + i], gz = e.gl[2 * e.n + i];
        v += e.gl[3 * e.n + i] + (gx * gx + gy * gy + gz * gz);
    }
    v = block_sum(v);
    if (threadIdx.x == 0) e.out[0] = -0.5 * v + (e.state[0] + e.state[1] + constants);
}

// ------------------------------------------------------------ host helpers
// One device allocation carved into aligned pieces, plus the object's stream.
struct Arena {
    unsigned char *base = nullptr;
    size_t used = 0, cap = 0;
    hipStream_t stream = nullptr;

    int init(size_t bytes) {
        cap = bytes;
        if (hipMalloc(&base, bytes) != hipSuccess) {
            base = nullptr;
            return err(VMC_ERR_MEMORY, "per-move object: device allocation failed");
        }
        if (hipMemset(base, 0, bytes) != hipSuccess) return err(VMC_ERR_DEVICE, "per-move object: memset failed");
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
            return err(VMC_ERR_DEVICE, "per-move object: stream creation failed");
        return VMC_OK;
    }
    template <typename T>
    T *take(size_t count) {
        const size_t off = (used + 255) & ~size_t(255);
        used = off + count * sizeof(T);
        return reinterpret_cast<T *>(base + off);
    }
    void release() {
        if (stream) (void)hipStreamDestroy(stream);
        if (base) (void)hipFree(base);
        base = nullptr;
        stream = nullptr;
    }
    // Bytes the pieces of the given sizes take, with alignment.
    static size_t bytes_for(std::initializer_list<size_t> pieces) {
        size_t b = 0;
        for (size_t p : pieces) b += ((p + 255) & ~size_t(255)) + 256;
        return b;
    }
    int up(void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream) == hipSuccess
                   ? VMC_OK
                   : err(VMC_ERR_DEVICE, "per-move object: upload failed");
    }
    int down(void *dst, const void *src, size_t bytes) {
        if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream) != hipSuccess)
            return err(VMC_ERR_DEVICE, "per-move object: download failed");
        return VMC_OK;
    }
    int sync() {
        if (hipGetLastError() != hipSuccess) return err(VMC_ERR_DEVICE, "per-move object: launch failed");
        return hipStreamSynchronize(stream) == hipSuccess ? VMC_OK
                                                          : err(VMC_ERR_DEVICE, "per-move object: device error");
    }
};

int upload_positions(Arena &a, double *dpos, int n, const double *x, const double *y, const double *z) {
    if (!x || !y || !z) return err(VMC_ERR_INVALID, "per-move object: null position array");
    int rc = a.up(dpos, x, sizeof(double) * n);
    if (!rc) rc = a.up(dpos + n, y, sizeof(double) * n);
    if (!rc) rc = a.up(dpos + 2 * n, z, sizeof(double) * n);
    return rc;
}

int upload_gl(Arena &a, double *dgl, int n, const double *gx, const double *gy, const double *gz, const double *lp) {
    if (!gx || !gy || !gz || !lp) return err(VMC_ERR_INVALID, "per-move object: null derivative array");
    int rc = a.up(dgl, gx, sizeof(double) * n);
    if (!rc) rc = a.up(dgl + n, gy, sizeof(double) * n);
    if (!rc) rc = a.up(dgl + 2 * n, gz, sizeof(double) * n);
    if (!rc) rc = a.up(dgl + 3 * n, lp, sizeof(double) * n);
    return rc;
}

int download_gl(Arena &a, const double *dgl, int n, double *gx, double *gy, double *gz, double *lp) {
    int rc = a.down(gx, dgl, sizeof(double) * n);
    if (!rc) rc = a.down(gy, dgl + n, sizeof(double) * n);
    if (!rc) rc = a.down(gz, dgl + 2 * n, sizeof(double) * n);
    if (!rc) rc = a.down(lp, dgl + 3 * n, sizeof(double) * n);
    return rc;
}

}  // namespace

struct vmc_slater {
    Arena a;
    SlaterDev d{};
    vmc::ShellTable shells;
    std::vector<double> kvec;  // [nk][3]
};

struct vmc_pade {
    Arena a;
    PadeDev d{};
    int cap = 0;
};

struct vmc_ewald {
    Arena a;
    EwaldDev d{};
    int cap = 0;
    double self_term = 0.0, background = 0.0;
    std::vector<double> g;  // [3][ng] (host copy for ensure())
    std::vector<double> w;
};

namespace {

bool valid_n(int n) { return n >= 1 && n <= VMC_SINGLE_MAX_N; }

int check_p(const vmc_slater *s, int p) {
    return (p >= 0 && p < s->d.n) ? VMC_OK : err(VMC_ERR_INVALID, "vmc_slater: particle index out of range");
}

// Grows the per-call position and derivative buffers of a Pade or Ewald
// object to n particles (a fresh allocation holding the fixed tables too for
// the Ewald object).
int pade_ensure(vmc_pade *j, int n) {
    if (n <= j->cap) {
        j->d.n = n;
        return VMC_OK;
    }
    if (j->a.base) {
        SINGLE_TRY(hipStreamSynchronize(j->a.stream));
        j->a.release();
    }
    j->a = Arena{};
    int rc = j->a.init(Arena::bytes_for({3 * sizeof(double) * n, 4 * sizeof(double) * n, 8 * sizeof(double)}));
    if (rc) return rc;
    j->d.pos = j->a.take<double>(3 * static_cast<size_t>(n));
    j->d.gl = j->a.take<double>(4 * static_cast<size_t>(n));
    j->d.out = j->a.take<double>(8);
    j->cap = n;
    j->d.n = n;
    return VMC_OK;
}

int ewald_ensure(vmc_ewald *t, int n) {
    if (n <= t->cap) {
        t->d.n = n;
        return VMC_OK;
    }
    const int ng = t->d.ng;
    std::vector<double> sre(static_cast<size_t>(ng)), sim(static_cast<size_t>(ng)), state(2, 0.0);
    if (t->a.base) {  // keep S_G and the cached energies across the move
        int rc = t->a.down(sre.data(), t->d.sre, sizeof(double) * ng);
        if (!rc) rc = t->a.down(sim.data(), t->d.sim, sizeof(double) * ng);
        if (!rc) rc = t->a.down(state.data(), t->d.state, 2 * sizeof(double));
        if (!rc) rc = t->a.sync();
        if (rc) return rc;
        t->a.release();
    }
    t->a = Arena{};
    const size_t g = static_cast<size_t>(ng > 0 ? ng : 1);
    int rc = t->a.init(Arena::bytes_for({3 * sizeof(double) * g, sizeof(double) * g, sizeof(double) * g,
                                          sizeof(double) * g, 3 * sizeof(double) * n, 4 * sizeof(double) * n,
                                          2 * sizeof(double), 8 * sizeof(double)}));
    if (rc) return rc;
    double *gd = t->a.take<double>(3 * g);
    double *wd = t->a.take<double>(g);
    t->d.g = gd;
    t->d.w = wd;
    t->d.sre = t->a.take<double>(g);
    t->d.sim = t->a.take<double>(g);
    t->d.pos = t->a.take<double>(3 * static_cast<size_t>(n));
    t->d.gl = t->a.take<double>(4 * static_cast<size_t>(n));
    t->d.state = t->a.take<double>(2);
    t->d.out = t->a.take<double>(8);
    if (ng > 0) {
        rc = t->a.up(gd, t->g.data(), sizeof(double) * 3 * ng);
        if (!rc) rc = t->a.up(wd, t->w.data(), sizeof(double) * ng);
        if (!rc) rc = t->a.up(t->d.sre, sre.data(), sizeof(double) * ng);
        if (!rc) rc = t->a.up(t->d.sim, sim.data(), sizeof(double) * ng);
    }
    if (!rc) rc = t->a.up(t->d.state, state.data(), 2 * sizeof(double));
    if (!rc) rc = t->a.sync();
    if (rc) return rc;
    t->cap = n;
    t->d.n = n;
    return VMC_OK;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ vmc_slater
int vmc_slater_create(int32_t n, double box_length, vmc_slater **out) {
    if (!out) return err(VMC_ERR_INVALID, "vmc_slater_create: null output");
    *out = nullptr;
    if (!valid_n(n)) return err(VMC_ERR_INVALID, "vmc_slater_create: n must be in [1, VMC_SINGLE_MAX_N]");
    if (!(box_length > 0.0) || !std::isfinite(box_length))
        return err(VMC_ERR_INVALID, "vmc_slater_create: box length must be > 0");
    vmc_slater *s = new (std::nothrow) vmc_slater();
    if (!s) return err(VMC_ERR_MEMORY, "vmc_slater_create: host allocation failed");
    s->shells = vmc::fill_shells(n);
    const int nk = s->shells.nk;
    const int S = static_cast<int>(pad8(n)), TS = static_cast<int>(pad8(nk));
    // k = 2 pi (1/L) n, in the reference's order of operations (slater_plane_wave.cpp:140-149).
    const double inv_L = 1.0 / box_length;
    std::vector<double> k3(3 * static_cast<size_t>(nk));
    s->kvec.resize(3 * static_cast<size_t>(nk));
    for (int k = 0; k < nk; ++k)
        for (int c = 0; c < 3; ++c) {
            const double v = 2 * std::numbers::pi * inv_L * static_cast<double>(s->shells.nvec[k][c]);
            k3[static_cast<size_t>(c) * nk + k] = v;
            s->kvec[3 * static_cast<size_t>(k) + c] = v;
        }
    std::vector<int> kidx(s->shells.orb_k.begin(), s->shells.orb_k.end());
    std::vector<int> type(s->shells.orb_type.begin(), s->shells.orb_type.end());
    const size_t nn = static_cast<size_t>(n) * S, nt = static_cast<size_t>(n) * TS;
    int rc = s->a.init(Arena::bytes_for({3 * sizeof(double) * nk, sizeof(int) * n, sizeof(int) * n,
                                          sizeof(double) * nt, sizeof(double) * nt, 2 * sizeof(double) * nk,
                                          sizeof(double) * nn, sizeof(double) * nn, sizeof(double) * nn,
                                          sizeof(int) * n, sizeof(double) * n, 3 * sizeof(double) * n,
                                          4 * sizeof(double) * n, 8 * sizeof(double)}));
    if (rc) {
        s->a.release();
        delete s;
        return rc;
    }
    SlaterDev &d = s->d;
    d.n = n;
    d.S = S;
    d.nk = nk;
    d.TS = TS;
    double *kd = s->a.take<double>(3 * static_cast<size_t>(nk));
    int *kid = s->a.take<int>(n);
    int *tyd = s->a.take<int>(n);
    d.k = kd;
    d.kidx = kid;
    d.type = tyd;
    d.sinc = s->a.take<double>(nt);
    d.cosc = s->a.take<double>(nt);
    d.saved = s->a.take<double>(2 * static_cast<size_t>(nk));
    d.D = s->a.take<double>(nn);
    d.LU = s->a.take<double>(nn);
    d.INV = s->a.take<double>(nn);
    d.piv = s->a.take<int>(n);
    d.row = s->a.take<double>(n);
    d.pos = s->a.take<double>(3 * static_cast<size_t>(n));
    d.gl = s->a.take<double>(4 * static_cast<size_t>(n));
    d.out = s->a.take<double>(8);
    rc = s->a.up(kd, k3.data(), sizeof(double) * 3 * nk);
    if (!rc) rc = s->a.up(kid, kidx.data(), sizeof(int) * n);
    if (!rc) rc = s->a.up(tyd, type.data(), sizeof(int) * n);
    if (!rc) rc = s->a.sync();
    if (rc) {
        s->a.release();
        delete s;
        return rc;
    }
    *out = s;
    return VMC_OK;
}

int vmc_slater_destroy(vmc_slater *s) {
    if (!s) return VMC_OK;
    if (s->a.stream) (void)hipStreamSynchronize(s->a.stream);
    s->a.release();
    delete s;
    return VMC_OK;
}

int vmc_slater_shape(const vmc_slater *s, int32_t *n_unique_k, int32_t *matrix_stride, int32_t *trig_stride) {
    if (!s) return err(VMC_ERR_INVALID, "vmc_slater_shape: null object");
    if (n_unique_k) *n_unique_k = s->d.nk;
    if (matrix_stride) *matrix_stride = s->d.S;
    if (trig_stride) *trig_stride = s->d.TS;
    return VMC_OK;
}

int vmc_slater_tables(const vmc_slater *s, int32_t *n_vec, double *k_vec, int32_t *orbital_k, uint8_t *orbital_type) {
    if (!s) return err(VMC_ERR_INVALID, "vmc_slater_tables: null object");
    const int nk = s->d.nk, n = s->d.n;
    for (int k = 0; k < nk; ++k)
        for (int c = 0; c < 3; ++c) {
            if (n_vec) n_vec[3 * k + c] = s->shells.nvec[static_cast<size_t>(k)][c];
            if (k_vec) k_vec[3 * k + c] = s->kvec[3 * static_cast<size_t>(k) + c];
        }
    for (int j = 0; j < n; ++j) {
        if (orbital_k) orbital_k[j] = s->shells.orb_k[static_cast<size_t>(j)];
        if (orbital_type) orbital_type[j] = s->shells.orb_type[static_cast<size_t>(j)];
    }
    return VMC_OK;
}

int vmc_slater_log_abs_det(vmc_slater *s, const double *x, const double *y, const double *z, double *log_abs_det) {
    if (!s || !log_abs_det) return err(VMC_ERR_INVALID, "vmc_slater_log_abs_det: null argument");
    int rc = upload_positions(s->a, const_cast<double *>(s->d.pos), s->d.n, x, y, z);
    if (rc) return rc;
    hipLaunchKernelGGL(k_slater_log_det, dim3(1), dim3(kT), 0, s->a.stream, s->d);
    rc = s->a.down(log_abs_det, s->d.out, sizeof(double));
    return rc ? rc : s->a.sync();
}

static int slater_row_op(vmc_slater *s, int op, int32_t p, double x, double y, double z) {
    if (!s) return err(VMC_ERR_INVALID, "vmc_slater: null object");
    int rc = check_p(s, p);
    if (rc) return rc;
    hipLaunchKernelGGL(k_slater_row_op, dim3(1), dim3(kT), 0, s->a.stream, s->d, op, static_cast<int>(p), x, y, z);
    return s->a.sync();
}

int vmc_slater_update_trig(vmc_slater *s, int32_t p, double x, double y, double z) {
    return slater_row_op(s, 0, p, x, y, z);
}

int vmc_slater_save_trig(vmc_slater *s, int32_t p) { return slater_row_op(s, 1, p, 0.0, 0.0, 0.0); }

int vmc_slater_restore_trig(vmc_slater *s, int32_t p) { return slater_row_op(s, 2, p, 0.0, 0.0, 0.0); }

int vmc_slater_build_row(vmc_slater *s, int32_t p, double *row) {
    if (!s || !row) return err(VMC_ERR_INVALID, "vmc_slater_build_row: null argument");
    int rc = check_p(s, p);
    if (rc) return rc;
    hipLaunchKernelGGL(k_slater_row_op, dim3(1), dim3(kT), 0, s->a.stream, s->d, 3, static_cast<int>(p), 0.0, 0.0,
                       0.0);
    rc = s->a.down(row, s->d.row, sizeof(double) * s->d.n);
    return rc ? rc : s->a.sync();
}

int vmc_slater_ratio(vmc_slater *s, int32_t p, const double *row, double *ratio) {
    if (!s || !row || !ratio) return err(VMC_ERR_INVALID, "vmc_slater_ratio: null argument");
    int rc = check_p(s, p);
    if (!rc) rc = s->a.up(s->d.row, row, sizeof(double) * s->d.n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_slater_ratio, dim3(1), dim3(kT), 0, s->a.stream, s->d, static_cast<int>(p));
    rc = s->a.down(ratio, s->d.out + 1, sizeof(double));
    return rc ? rc : s->a.sync();
}

int vmc_slater_accept(vmc_slater *s, int32_t p, const double *row, double ratio) {
    if (!s || !row) return err(VMC_ERR_INVALID, "vmc_slater_accept: null argument");
    int rc = check_p(s, p);
    if (!rc) rc = s->a.up(s->d.row, row, sizeof(double) * s->d.n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_slater_accept, dim3(1), dim3(kT), sizeof(double) * s->d.n, s->a.stream, s->d,
                       static_cast<int>(p), ratio);
    return s->a.sync();
}

int vmc_slater_add_derivatives(vmc_slater *s, double *grad_x, double *grad_y, double *grad_z, double *lap) {
    if (!s) return err(VMC_ERR_INVALID, "vmc_slater_add_derivatives: null object");
    int rc = upload_gl(s->a, s->d.gl, s->d.n, grad_x, grad_y, grad_z, lap);
    if (rc) return rc;
    hipLaunchKernelGGL(k_slater_derivs, dim3(1), dim3(kT), 0, s->a.stream, s->d);
    rc = download_gl(s->a, s->d.gl, s->d.n, grad_x, grad_y, grad_z, lap);
    return rc ? rc : s->a.sync();
}

int vmc_slater_read(const vmc_slater *s, int32_t which, double *out) {
    if (!s || !out) return err(VMC_ERR_INVALID, "vmc_slater_read: null argument");
    const size_t nn = static_cast<size_t>(s->d.n) * s->d.S, nt = static_cast<size_t>(s->d.n) * s->d.TS;
    const double *src = nullptr;
    size_t count = 0;
    switch (which) {
        case VMC_SLATER_D: src = s->d.D, count = nn; break;
        case VMC_SLATER_INV: src = s->d.INV, count = nn; break;
        case VMC_SLATER_LU: src = s->d.LU, count = nn; break;
        case VMC_SLATER_SIN: src = s->d.sinc, count = nt; break;
        case VMC_SLATER_COS: src = s->d.cosc, count = nt; break;
        default: return err(VMC_ERR_INVALID, "vmc_slater_read: unknown array");
    }
    Arena &a = const_cast<vmc_slater *>(s)->a;
    int rc = a.down(out, src, sizeof(double) * count);
    return rc ? rc : a.sync();
}

int vmc_slater_pivots(const vmc_slater *s, int32_t *pivots) {
    if (!s || !pivots) return err(VMC_ERR_INVALID, "vmc_slater_pivots: null argument");
    Arena &a = const_cast<vmc_slater *>(s)->a;
    int rc = a.down(pivots, s->d.piv, sizeof(int) * s->d.n);
    return rc ? rc : a.sync();
}

// ------------------------------------------------------------ vmc_pade
int vmc_pade_create(double box_length, double a, double b, vmc_pade **out) {
    if (!out) return err(VMC_ERR_INVALID, "vmc_pade_create: null output");
    *out = nullptr;
    if (!(box_length > 0.0) || !std::isfinite(box_length) || !std::isfinite(a) || !std::isfinite(b))
        return err(VMC_ERR_INVALID, "vmc_pade_create: box length must be > 0 and a, b finite");
    vmc_pade *j = new (std::nothrow) vmc_pade();
    if (!j) return err(VMC_ERR_MEMORY, "vmc_pade_create: host allocation failed");
    j->d.L = box_length;
    j->d.a = a;
    j->d.b = b;
    *out = j;
    return VMC_OK;
}

int vmc_pade_destroy(vmc_pade *j) {
    if (!j) return VMC_OK;
    if (j->a.stream) (void)hipStreamSynchronize(j->a.stream);
    j->a.release();
    delete j;
    return VMC_OK;
}

static int pade_begin(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z) {
    if (!j) return err(VMC_ERR_INVALID, "vmc_pade: null object");
    if (!valid_n(n)) return err(VMC_ERR_INVALID, "vmc_pade: n must be in [1, VMC_SINGLE_MAX_N]");
    int rc = pade_ensure(j, n);
    return rc ? rc : upload_positions(j->a, const_cast<double *>(j->d.pos), n, x, y, z);
}

int vmc_pade_value(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z, double *value) {
    if (!value) return err(VMC_ERR_INVALID, "vmc_pade_value: null output");
    int rc = pade_begin(j, n, x, y, z);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pade_value, dim3(1), dim3(kT), 0, j->a.stream, j->d);
    rc = j->a.down(value, j->d.out, sizeof(double));
    return rc ? rc : j->a.sync();
}

int vmc_pade_add_derivatives(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z,
                             double *grad_x, double *grad_y, double *grad_z, double *lap) {
    int rc = pade_begin(j, n, x, y, z);
    if (!rc) rc = upload_gl(j->a, j->d.gl, n, grad_x, grad_y, grad_z, lap);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pade_derivs, dim3(1), dim3(kT), 0, j->a.stream, j->d);
    rc = download_gl(j->a, j->d.gl, n, grad_x, grad_y, grad_z, lap);
    return rc ? rc : j->a.sync();
}

int vmc_pade_delta_value(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z, int32_t moved,
                         double old_x, double old_y, double old_z, double *delta) {
    if (!delta) return err(VMC_ERR_INVALID, "vmc_pade_delta_value: null output");
    if (moved < 0 || moved >= n) return err(VMC_ERR_INVALID, "vmc_pade_delta_value: moved index out of range");
    int rc = pade_begin(j, n, x, y, z);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pade_move, dim3(1), dim3(kT), 0, j->a.stream, j->d, static_cast<int>(moved), old_x, old_y,
                       old_z, 0);
    rc = j->a.down(delta, j->d.out, sizeof(double));
    return rc ? rc : j->a.sync();
}

int vmc_pade_update_for_move(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z,
                             int32_t moved, double old_x, double old_y, double old_z, double *grad_x,
                             double *grad_y, double *grad_z, double *lap) {
    if (moved < 0 || moved >= n) return err(VMC_ERR_INVALID, "vmc_pade_update_for_move: moved index out of range");
    int rc = pade_begin(j, n, x, y, z);
    if (!rc) rc = upload_gl(j->a, j->d.gl, n, grad_x, grad_y, grad_z, lap);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pade_move, dim3(1), dim3(kT), 0, j->a.stream, j->d, static_cast<int>(moved), old_x, old_y,
                       old_z, 1);
    rc = download_gl(j->a, j->d.gl, n, grad_x, grad_y, grad_z, lap);
    return rc ? rc : j->a.sync();
}

// ------------------------------------------------------------ vmc_ewald
int vmc_ewald_create(double box_length, double num_particles, vmc_ewald **out) {
    if (!out) return err(VMC_ERR_INVALID, "vmc_ewald_create: null output");
    *out = nullptr;
    if (!(box_length > 0.0) || !std::isfinite(box_length) || !std::isfinite(num_particles))
        return err(VMC_ERR_INVALID, "vmc_ewald_create: box length must be > 0");
    vmc_ewald *t = new (std::nothrow) vmc_ewald();
    if (!t) return err(VMC_ERR_MEMORY, "vmc_ewald_create: host allocation failed");
    constexpr double pi = std::numbers::pi;
    const vmc::GTable gt = vmc::ewald_gvectors(box_length, 6.0 / box_length, 0);
    const int ng = static_cast<int>(gt.m.size());
    const double unit = 2.0 * pi / box_length;
    t->g.resize(3 * static_cast<size_t>(ng));
    for (int q = 0; q < ng; ++q)
        for (int c = 0; c < 3; ++c) t->g[static_cast<size_t>(c) * ng + q] = unit * static_cast<double>(gt.m[q][c]);
    t->w = gt.w;
    t->d.ng = ng;
    t->d.L = box_length;
    t->d.alpha = 6.0 / box_length;
    t->d.pref = 1.0 / (2.0 * pi * box_length * box_length * box_length);
    // Self and background terms as the reference's constructor writes them
    // (energy_tracking.cpp:6-12).
    t->self_term = -6.0 * num_particles / (std::sqrt(pi) * box_length);
    t->background = -pi * num_particles * num_particles / (72.0 * box_length);
    const int n0 = num_particles >= 1.0 && num_particles <= VMC_SINGLE_MAX_N ? static_cast<int>(num_particles) : 1;
    const int rc = ewald_ensure(t, n0);
    if (rc) {
        t->a.release();
        delete t;
        return rc;
    }
    *out = t;
    return VMC_OK;
}

int vmc_ewald_destroy(vmc_ewald *t) {
    if (!t) return VMC_OK;
    if (t->a.stream) (void)hipStreamSynchronize(t->a.stream);
    t->a.release();
    delete t;
    return VMC_OK;
}

static int ewald_begin(vmc_ewald *t, int32_t n, const double *x, const double *y, const double *z) {
    if (!t) return err(VMC_ERR_INVALID, "vmc_ewald: null object");
    if (!valid_n(n)) return err(VMC_ERR_INVALID, "vmc_ewald: n must be in [1, VMC_SINGLE_MAX_N]");
    int rc = ewald_ensure(t, n);
    return rc ? rc : upload_positions(t->a, const_cast<double *>(t->d.pos), n, x, y, z);
}

int vmc_ewald_init_structure_factors(vmc_ewald *t, int32_t n, const double *x, const double *y, const double *z) {
    int rc = ewald_begin(t, n, x, y, z);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ewald_sf_init, dim3(1), dim3(kT), 0, t->a.stream, t->d);
    return t->a.sync();
}

int vmc_ewald_init_reciprocal(vmc_ewald *t) {
    if (!t) return err(VMC_ERR_INVALID, "vmc_ewald: null object");
    hipLaunchKernelGGL(k_ewald_recip_init, dim3(1), dim3(kT), 0, t->a.stream, t->d);
    return t->a.sync();
}

int vmc_ewald_init_real(vmc_ewald *t, int32_t n, const double *x, const double *y, const double *z) {
    int rc = ewald_begin(t, n, x, y, z);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ewald_real_init, dim3(1), dim3(kT), 0, t->a.stream, t->d);
    return t->a.sync();
}

int vmc_ewald_update_structure_factors(vmc_ewald *t, double old_x, double old_y, double old_z, double new_x,
                                       double new_y, double new_z) {
    if (!t) return err(VMC_ERR_INVALID, "vmc_ewald: null object");
    hipLaunchKernelGGL(k_ewald_sf_update, dim3(1), dim3(kT), 0, t->a.stream, t->d, old_x, old_y, old_z, new_x, new_y,
                       new_z);
    return t->a.sync();
}

int vmc_ewald_update_real(vmc_ewald *t, int32_t n, const double *x, const double *y, const double *z, int32_t moved,
                          double old_x, double old_y, double old_z) {
    if (moved < 0 || moved >= n) return err(VMC_ERR_INVALID, "vmc_ewald_update_real: moved index out of range");
    int rc = ewald_begin(t, n, x, y, z);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ewald_real_update, dim3(1), dim3(kT), 0, t->a.stream, t->d, static_cast<int>(moved), old_x,
                       old_y, old_z);
    return t->a.sync();
}

int vmc_ewald_total_energy(vmc_ewald *t, int32_t n, const double *grad_x, const double *grad_y, const double *grad_z,
                           const double *lap, double *energy) {
    if (!t || !energy) return err(VMC_ERR_INVALID, "vmc_ewald_total_energy: null argument");
    if (!valid_n(n)) return err(VMC_ERR_INVALID, "vmc_ewald: n must be in [1, VMC_SINGLE_MAX_N]");
    int rc = ewald_ensure(t, n);
    if (!rc) rc = upload_gl(t->a, const_cast<double *>(t->d.gl), n, grad_x, grad_y, grad_z, lap);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ewald_total, dim3(1), dim3(kT), 0, t->a.stream, t->d, t->self_term + t->background);
    rc = t->a.down(energy, t->d.out, sizeof(double));
    return rc ? rc : t->a.sync();
}

int vmc_ewald_state(const vmc_ewald *t, double *v_real, double *v_recip, double *self_term, double *background,
                    int32_t *n_g, double *s_re, double *s_im) {
    if (!t) return err(VMC_ERR_INVALID, "vmc_ewald_state: null object");
    Arena &a = const_cast<vmc_ewald *>(t)->a;
    double st[2] = {0.0, 0.0};
    int rc = a.down(st, t->d.state, sizeof(st));
    if (!rc && s_re && t->d.ng) rc = a.down(s_re, t->d.sre, sizeof(double) * t->d.ng);
    if (!rc && s_im && t->d.ng) rc = a.down(s_im, t->d.sim, sizeof(double) * t->d.ng);
    if (!rc) rc = a.sync();
    if (rc) return rc;
    if (v_real) *v_real = st[0];
    if (v_recip) *v_recip = st[1];
    if (self_term) *self_term = t->self_term;
    if (background) *background = t->background;
    if (n_g) *n_g = t->d.ng;
    return VMC_OK;
}

}  // extern "C"
