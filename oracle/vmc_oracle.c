/*
 * vmc_oracle.c -- FP64 CPU oracle for the VMC sweep of the 3D homogeneous
 * electron gas.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline of bench.py.  The product library never links or calls it.
 *
 * Citations are reference-repository paths (file:line).  Loops run in the
 * reference's order; the reference's `#pragma omp simd` reductions may sum in
 * a different order, so agreement with the reference build is at rounding
 * level (about 1e-14 relative), not bitwise.
 */
#define _GNU_SOURCE
#include "vmc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* =================================================================== RNG */

/* mt19937_64 (Matsumoto & Nishimura; parameters of C++ [rand.predef]).  The
   reference seeds one per chain: simulation.hpp:33, simulation.cpp:18. */
#define MT_N 312
#define MT_M 156
#define MT_UPPER 0xFFFFFFFF80000000ULL
#define MT_LOWER 0x000000007FFFFFFFULL

void orc_mt64_seed(orc_mt64 *g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < MT_N; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = MT_N;
}

static void mt64_twist(orc_mt64 *g) {
    for (int i = 0; i < MT_N; ++i) {
        uint64_t x = (g->mt[i] & MT_UPPER) | (g->mt[(i + 1) % MT_N] & MT_LOWER);
        uint64_t xa = x >> 1;
        if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
        g->mt[i] = g->mt[(i + MT_M) % MT_N] ^ xa;
    }
    g->idx = 0;
}

uint64_t orc_mt64_next(orc_mt64 *g) {
    if (g->idx >= MT_N) mt64_twist(g);
    uint64_t x = g->mt[g->idx++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= x >> 43;
    return x;
}

/* std::generate_canonical<double, 53> over a 64-bit engine as libstdc++
   implements it: one draw, (double)x / 2^64, clamped below 1.  Used by
   std::uniform_real_distribution (simulation.hpp:34-35). */
double orc_std_canonical(orc_mt64 *g) {
    const double r = 18446744073709551616.0; /* 2^64 */
    double sum = (double)orc_mt64_next(g);
    double ret = sum / r;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

double orc_std_uniform_real(orc_mt64 *g, double a, double b) {
    return orc_std_canonical(g) * (b - a) + a;
}

/* std::uniform_int_distribution<size_t>{0, n-1} over mt19937_64 as libstdc++
   implements it (Lemire's nearly-divisionless method with a 128-bit product);
   simulation.hpp:36, simulation.cpp:91. */
uint64_t orc_std_uniform_index(orc_mt64 *g, uint64_t n) {
    const uint64_t range = n; /* urange + 1 */
    unsigned __int128 product = (unsigned __int128)orc_mt64_next(g) * range;
    uint64_t low = (uint64_t)product;
    if (low < range) {
        uint64_t threshold = (uint64_t)(-range) % range;
        while (low < threshold) {
            product = (unsigned __int128)orc_mt64_next(g) * range;
            low = (uint64_t)product;
        }
    }
    return (uint64_t)(product >> 64);
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): the per-walker counter
   RNG of the ensemble sweep.  Multipliers and Weyl keys are the published
   ones. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ========================================================= shell filling */

/* matrix.hpp:101-116 */
static int is_canonical(int x, int y, int z) {
    if (x > 0) return 1;
    if (x < 0) return 0;
    if (y > 0) return 1;
    if (y < 0) return 0;
    if (z > 0) return 1;
    if (z < 0) return 0;
    return 1;
}

typedef struct { int x, y, z, m; } ncand;

static int ncand_cmp(const void *pa, const void *pb) {
    const ncand *a = (const ncand *)pa, *b = (const ncand *)pb;
    if (a->m != b->m) return a->m < b->m ? -1 : 1;
    if (a->x != b->x) return a->x < b->x ? -1 : 1;
    if (a->y != b->y) return a->y < b->y ? -1 : 1;
    if (a->z != b->z) return a->z < b->z ? -1 : 1;
    return 0;
}

/* slater_plane_wave.cpp:30-152: canonical n sorted by (|n|^2, nx, ny, nz);
   orbital 0 = cos(0); each further k gives (cos, sin). */
int orc_shell_fill(int n, int *nvec, int *orb_k, unsigned char *orb_type) {
    const int nmax = (int)ceil(cbrt((double)n)) + 2;
    const int side = 2 * nmax + 1;
    ncand *c = (ncand *)malloc(sizeof(ncand) * (size_t)side * side * side);
    int nc = 0;
    for (int x = -nmax; x <= nmax; ++x)
        for (int y = -nmax; y <= nmax; ++y)
            for (int z = -nmax; z <= nmax; ++z) {
                if (!is_canonical(x, y, z)) continue;
                c[nc].x = x; c[nc].y = y; c[nc].z = z; c[nc].m = x * x + y * y + z * z;
                ++nc;
            }
    qsort(c, (size_t)nc, sizeof(ncand), ncand_cmp);
    int orb = 0, k = 0;
    for (int i = 0; i < nc && orb < n; ++i) {
        nvec[3 * k + 0] = c[i].x;
        nvec[3 * k + 1] = c[i].y;
        nvec[3 * k + 2] = c[i].z;
        orb_k[orb] = k;
        orb_type[orb] = 0;
        ++orb;
        if (c[i].m != 0 && orb < n) {
            orb_k[orb] = k;
            orb_type[orb] = 1;
            ++orb;
        }
        ++k;
    }
    free(c);
    return k;
}

/* ============================================================ LU (matrix.hpp) */

/* matrix.hpp:13-61: in-place partial-pivot Doolittle LU, pivot tolerance 1e-12. */
static int lu_decomp(double *a, int *piv, int n) {
    int swaps = 0;
    for (int r = 0; r < n; ++r) piv[r] = r;
    for (int col = 0; col < n; ++col) {
        int prow = col;
        double maxabs = fabs(a[col * n + col]);
        for (int r = col + 1; r < n; ++r) {
            double v = fabs(a[r * n + col]);
            if (v > maxabs) { maxabs = v; prow = r; }
        }
        if (maxabs < 1e-12) {
            a[col * n + col] = 0.0;
            continue;
        }
        if (prow != col) {
            for (int c2 = 0; c2 < n; ++c2) {
                double t = a[col * n + c2];
                a[col * n + c2] = a[prow * n + c2];
                a[prow * n + c2] = t;
            }
            int t = piv[col]; piv[col] = piv[prow]; piv[prow] = t;
            ++swaps;
        }
        const double pv = a[col * n + col];
        for (int r = col + 1; r < n; ++r) {
            a[r * n + col] /= pv;
            const double m = a[r * n + col];
            for (int c2 = col + 1; c2 < n; ++c2) a[r * n + c2] -= m * a[col * n + c2];
        }
    }
    return swaps;
}

/* matrix.hpp:69-96 */
static void lu_solve(const double *lu, const int *piv, const double *b, double *x, int n) {
    for (int r = 0; r < n; ++r) x[r] = b[piv[r]];
    for (int r = 0; r < n; ++r) {
        double s = x[r];
        for (int c = 0; c < r; ++c) s -= lu[r * n + c] * x[c];
        x[r] = s;
    }
    for (int rev = 0; rev < n; ++rev) {
        int r = n - 1 - rev;
        double s = x[r];
        for (int c = r + 1; c < n; ++c) s -= lu[r * n + c] * x[c];
        x[r] = s / lu[r * n + r];
    }
}

/* ==================================================================== walker */

typedef struct {
    int n, nk, first;
    int *nvec, *orb_k;
    unsigned char *orb_type;
    double *kx, *ky, *kz;
    double *D, *inv, *lu, *rhs, *sol;
    int *piv;
    double *cosc, *sinc;      /* [n][nk] trig cache, slater_plane_wave.hpp:47-48 */
    double *new_c, *new_s;    /* trig row of a pending move */
    double *row;              /* build_row buffer, slater_plane_wave.hpp:155-157 */
    double *col;              /* inv_d_col scratch, slater_plane_wave.cpp:386-390 */
    double logdet;
} det_t;

struct orc_walker {
    orc_system sys;
    int N, ndet;
    det_t det[2];
    double *x, *y, *z;
    double *gx, *gy, *gz, *lap;     /* Particles grad/lap, particles.hpp:14-22 */
    double *jgx, *jgy, *jgz, *jlap; /* wavefunction.hpp:19-21 */
    int jastrow_cache_valid;
    size_t steps_since_refresh;
    double alpha, self, bg, pref;
    int ng;
    double *Gx, *Gy, *Gz, *Gw;
    double *Sre, *Sim, *dre, *dim;
    double vrecip, vreal;
};

static int spin_of(const struct orc_walker *w, int i) {
    return (w->ndet == 2 && i >= w->sys.n_up) ? 1 : 0;
}

static double pair_a(const struct orc_walker *w, int i, int j) {
    return spin_of(w, i) == spin_of(w, j) ? w->sys.a_same : w->sys.a_anti;
}

static void det_init(det_t *d, int n, int first, double L) {
    memset(d, 0, sizeof(*d));
    d->n = n;
    d->first = first;
    if (n == 0) return;
    d->nvec = (int *)calloc((size_t)n * 3, sizeof(int));
    d->orb_k = (int *)calloc((size_t)n, sizeof(int));
    d->orb_type = (unsigned char *)calloc((size_t)n, 1);
    d->nk = orc_shell_fill(n, d->nvec, d->orb_k, d->orb_type);
    d->kx = (double *)calloc((size_t)d->nk, sizeof(double));
    d->ky = (double *)calloc((size_t)d->nk, sizeof(double));
    d->kz = (double *)calloc((size_t)d->nk, sizeof(double));
    const double inv_L = 1.0 / L;
    /* slater_plane_wave.cpp:136-144: k = 2 pi / L * n */
    for (int i = 0; i < d->nk; ++i) {
        d->kx[i] = 2 * M_PI * inv_L * (double)d->nvec[3 * i + 0];
        d->ky[i] = 2 * M_PI * inv_L * (double)d->nvec[3 * i + 1];
        d->kz[i] = 2 * M_PI * inv_L * (double)d->nvec[3 * i + 2];
    }
    size_t nn = (size_t)n * n;
    d->D = (double *)calloc(nn, sizeof(double));
    d->inv = (double *)calloc(nn, sizeof(double));
    d->lu = (double *)calloc(nn, sizeof(double));
    d->rhs = (double *)calloc((size_t)n, sizeof(double));
    d->sol = (double *)calloc((size_t)n, sizeof(double));
    d->piv = (int *)calloc((size_t)n, sizeof(int));
    d->cosc = (double *)calloc((size_t)n * d->nk, sizeof(double));
    d->sinc = (double *)calloc((size_t)n * d->nk, sizeof(double));
    d->new_c = (double *)calloc((size_t)d->nk, sizeof(double));
    d->new_s = (double *)calloc((size_t)d->nk, sizeof(double));
    d->row = (double *)calloc((size_t)n, sizeof(double));
    d->col = (double *)calloc((size_t)n, sizeof(double));
}

static void det_free(det_t *d) {
    free(d->nvec); free(d->orb_k); free(d->orb_type);
    free(d->kx); free(d->ky); free(d->kz);
    free(d->D); free(d->inv); free(d->lu); free(d->rhs); free(d->sol); free(d->piv);
    free(d->cosc); free(d->sinc); free(d->new_c); free(d->new_s); free(d->row); free(d->col);
}

/* slater_plane_wave.cpp:202-318: trig cache, D, LU, log|det| and the
   transposed inverse by n LU solves. */
static double det_log_abs_det(det_t *d, const double *x, const double *y, const double *z) {
    const int n = d->n, nk = d->nk;
    if (n == 0) return 0.0;
    for (int p = 0; p < n; ++p) {
        const double px = x[d->first + p], py = y[d->first + p], pz = z[d->first + p];
        for (int k = 0; k < nk; ++k) {
            const double dot = d->kx[k] * px + d->ky[k] * py + d->kz[k] * pz;
            sincos(dot, &d->sinc[p * nk + k], &d->cosc[p * nk + k]);
        }
        for (int o = 0; o < n; ++o) {
            const int ki = d->orb_k[o];
            const double type = (double)d->orb_type[o];
            const double ct = d->cosc[p * nk + ki], st = d->sinc[p * nk + ki];
            d->D[p * n + o] = ct + type * (st - ct);
        }
    }
    memcpy(d->lu, d->D, sizeof(double) * (size_t)n * n);
    (void)lu_decomp(d->lu, d->piv, n);
    double ld = 0.0;
    for (int i = 0; i < n; ++i) ld += log(fabs(d->lu[i * n + i]));
    if (!isfinite(ld)) {
        d->logdet = -INFINITY;
        return d->logdet;
    }
    for (int c = 0; c < n; ++c) {
        for (int r = 0; r < n; ++r) d->rhs[r] = 0.0;
        d->rhs[c] = 1.0;
        lu_solve(d->lu, d->piv, d->rhs, d->sol, n);
        for (int r = 0; r < n; ++r) d->inv[c * n + r] = d->sol[r];
    }
    d->logdet = ld;
    return ld;
}

/* slater_plane_wave.cpp:172-200 (trig row of the moved particle),
   320-349 (build_row), 351-365 (determinant_ratio). */
static double det_ratio(det_t *d, int pl, double px, double py, double pz) {
    const int n = d->n, nk = d->nk;
    for (int k = 0; k < nk; ++k) {
        const double dot = d->kx[k] * px + d->ky[k] * py + d->kz[k] * pz;
        sincos(dot, &d->new_s[k], &d->new_c[k]);
    }
    for (int o = 0; o < n; ++o) {
        const int ki = d->orb_k[o];
        const double type = (double)d->orb_type[o];
        d->row[o] = d->new_c[ki] + type * (d->new_s[ki] - d->new_c[ki]);
    }
    double ratio = 0.0;
    for (int j = 0; j < n; ++j) ratio += d->row[j] * d->inv[pl * n + j];
    return ratio;
}

/* slater_plane_wave.cpp:367-423: Sherman-Morrison on the transposed inverse,
   then patch row pl of D; the trig row is committed (update_trig_cache). */
static void det_accept(det_t *d, int pl, double ratio) {
    const int n = d->n, nk = d->nk;
    const double inv_ratio = 1.0 / ratio;
    for (int j = 0; j < n; ++j) d->col[j] = d->inv[pl * n + j];
    for (int k = 0; k < n; ++k) {
        if (k == pl) continue;
        double s = 0.0;
        for (int m = 0; m < n; ++m) s += d->row[m] * d->inv[k * n + m];
        const double f = s * inv_ratio;
        for (int j = 0; j < n; ++j) d->inv[k * n + j] -= d->col[j] * f;
    }
    for (int j = 0; j < n; ++j) d->inv[pl * n + j] = d->col[j] * inv_ratio;
    for (int j = 0; j < n; ++j) d->D[pl * n + j] = d->row[j];
    for (int k = 0; k < nk; ++k) {
        d->cosc[pl * nk + k] = d->new_c[k];
        d->sinc[pl * nk + k] = d->new_s[k];
    }
}

/* slater_plane_wave.cpp:425-499 */
static void det_add_derivatives(const det_t *d, double *gx, double *gy, double *gz, double *lap) {
    const int n = d->n, nk = d->nk;
    for (int p = 0; p < n; ++p) {
        double dx = 0, dy = 0, dz = 0, lt = 0;
        for (int o = 0; o < n; ++o) {
            const int ki = d->orb_k[o];
            const double kx = d->kx[ki], ky = d->ky[ki], kz = d->kz[ki];
            const double ksq = kx * kx + ky * ky + kz * kz;
            const double type = (double)d->orb_type[o];
            const double ct = d->cosc[p * nk + ki], st = d->sinc[p * nk + ki];
            const double gf = -st + type * (st + ct);
            const double lf = -ct + type * (ct - st);
            const double wgt = d->inv[p * n + o];
            dx += wgt * kx * gf;
            dy += wgt * ky * gf;
            dz += wgt * kz * gf;
            lt += wgt * ksq * lf;
        }
        const double g2 = dx * dx + dy * dy + dz * dz;
        gx[d->first + p] += dx;
        gy[d->first + p] += dy;
        gz[d->first + p] += dz;
        lap[d->first + p] += (lt - g2);
    }
}

/* Minimum image by masks: jastrow_pade.cpp:37-39, energy_tracking.cpp:127-129. */
static double min_image(double d, double L) {
    const double half = 0.5 * L;
    return d + L * (double)(d <= -half) + (-L) * (double)(d > half);
}

/* jastrow_pade.cpp:7-55 (value, the full N x N sum halved); the pair a is
   the spin-resolved extension (parallel a_same, antiparallel a_anti). */
static double jastrow_value(const struct orc_walker *w) {
    const int N = w->N;
    const double L = w->sys.L, b = w->sys.b;
    double J = 0.0;
    for (int i = 0; i < N; ++i) {
        double loc = 0.0;
        for (int j = 0; j < N; ++j) {
            const double mask = (i == j) ? 0.0 : 1.0;
            const double dx = min_image(w->x[i] - w->x[j], L);
            const double dy = min_image(w->y[i] - w->y[j], L);
            const double dz = min_image(w->z[i] - w->z[j], L);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double inv_den = 1.0 / (1.0 + b * r);
            loc += mask * pair_a(w, i, j) * r * inv_den;
        }
        J += loc;
    }
    return 0.5 * J;
}

/* jastrow_pade.cpp:57-138 */
static void jastrow_add_derivatives(const struct orc_walker *w, double *gx, double *gy, double *gz,
                                    double *lap) {
    const int N = w->N;
    const double L = w->sys.L, b = w->sys.b;
    for (int i = 0; i < N; ++i) {
        double ax = 0, ay = 0, az = 0, al = 0;
        for (int j = 0; j < N; ++j) {
            const double valid = (i == j) ? 0.0 : 1.0;
            const double dx = min_image(w->x[i] - w->x[j], L);
            const double dy = min_image(w->y[i] - w->y[j], L);
            const double dz = min_image(w->z[i] - w->z[j], L);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const int degen = r < 1e-12;
            const double inv_r = degen ? 1.0 : 1.0 / r;
            const double mask = degen ? 0.0 : valid;
            const double a = pair_a(w, i, j);
            const double den = 1.0 / (1.0 + b * r);
            const double den2 = den * den, den3 = den2 * den;
            const double d1 = a * den2;
            const double d2 = -2.0 * a * b * den3;
            const double gf = mask * d1 * inv_r;
            const double lp = mask * (d2 + 2.0 * d1 * inv_r);
            ax += gf * dx;
            ay += gf * dy;
            az += gf * dz;
            al += lp;
        }
        gx[i] += ax;
        gy[i] += ay;
        gz[i] += az;
        lap[i] += al;
    }
}

/* jastrow_pade.cpp:140-200; x[p] holds the new position, old is passed. */
static double jastrow_delta(const struct orc_walker *w, int p, double ox, double oy, double oz,
                            double nx, double ny, double nz) {
    const int N = w->N;
    const double L = w->sys.L, b = w->sys.b;
    double delta = 0.0;
    for (int j = 0; j < N; ++j) {
        const double valid = (j == p) ? 0.0 : 1.0;
        const double a = pair_a(w, p, j);
        const double dxo = min_image(ox - w->x[j], L);
        const double dyo = min_image(oy - w->y[j], L);
        const double dzo = min_image(oz - w->z[j], L);
        const double ro = sqrt(dxo * dxo + dyo * dyo + dzo * dzo);
        const double deno = 1.0 / (1.0 + b * ro);
        const double dxn = min_image(nx - w->x[j], L);
        const double dyn = min_image(ny - w->y[j], L);
        const double dzn = min_image(nz - w->z[j], L);
        const double rn = sqrt(dxn * dxn + dyn * dyn + dzn * dzn);
        const double denn = 1.0 / (1.0 + b * rn);
        delta += valid * a * (rn * denn - ro * deno);
    }
    return delta;
}

/* jastrow_pade.cpp:203-316; positions already hold the new r_p. */
static void jastrow_update_for_move(const struct orc_walker *w, int p, double ox, double oy,
                                    double oz, double *gx, double *gy, double *gz, double *lap) {
    gx[p] = gy[p] = gz[p] = lap[p] = 0.0;
    const int N = w->N;
    const double L = w->sys.L, b = w->sys.b;
    const double nx = w->x[p], ny = w->y[p], nz = w->z[p];
    double mx = 0, my = 0, mz = 0, ml = 0;
    for (int j = 0; j < N; ++j) {
        const int moved = (j == p);
        const double a = pair_a(w, p, j);
        const double m2ab = -2.0 * a * b;
        const double dxo = min_image(ox - w->x[j], L);
        const double dyo = min_image(oy - w->y[j], L);
        const double dzo = min_image(oz - w->z[j], L);
        const double ro = sqrt(dxo * dxo + dyo * dyo + dzo * dzo);
        const double inv_ro = (ro < 1e-12) ? 1.0 : 1.0 / ro;
        const double mask_o = (moved || ro < 1e-12) ? 0.0 : 1.0;
        const double ido = 1.0 / (1.0 + b * ro);
        const double ido2 = ido * ido;
        const double d1o = a * ido2;
        const double d2o = m2ab * ido2 * ido;
        const double gfo = mask_o * d1o * inv_ro;
        const double lpo = mask_o * (d2o + 2.0 * d1o * inv_ro);
        const double dxn = min_image(nx - w->x[j], L);
        const double dyn = min_image(ny - w->y[j], L);
        const double dzn = min_image(nz - w->z[j], L);
        const double rn = sqrt(dxn * dxn + dyn * dyn + dzn * dzn);
        const double inv_rn = (rn < 1e-12) ? 1.0 : 1.0 / rn;
        const double mask_n = (moved || rn < 1e-12) ? 0.0 : 1.0;
        const double idn = 1.0 / (1.0 + b * rn);
        const double idn2 = idn * idn;
        const double d1n = a * idn2;
        const double d2n = m2ab * idn2 * idn;
        const double gfn = mask_n * d1n * inv_rn;
        const double lpn = mask_n * (d2n + 2.0 * d1n * inv_rn);
        mx += gfn * dxn;
        my += gfn * dyn;
        mz += gfn * dzn;
        ml += lpn;
        gx[j] += gfo * dxo - gfn * dxn;
        gy[j] += gfo * dyo - gfn * dyn;
        gz[j] += gfo * dzo - gfn * dzn;
        lap[j] += lpn - lpo;
    }
    gx[p] = mx;
    gy[p] = my;
    gz[p] = mz;
    lap[p] = ml;
}

/* ================================================================= Ewald */

static double ewald_alpha(const orc_system *sys) {
    return sys->alpha > 0.0 ? sys->alpha : 6.0 / sys->L; /* energy_tracking.cpp:7 */
}

static int ntot(const orc_system *sys) { return sys->n_up + sys->n_dn; }

/* energy_tracking.cpp:8-9 (alpha = 6/L) and the general-alpha forms
   -alpha N / sqrt(pi) and -pi N^2 / (2 alpha^2 L^3) they reduce to. */
double orc_ewald_self(const orc_system *sys) {
    const double N = (double)ntot(sys), L = sys->L;
    if (sys->alpha <= 0.0) return -6.0 * N / (sqrt(M_PI) * L);
    return -sys->alpha * N / sqrt(M_PI);
}

double orc_ewald_background(const orc_system *sys) {
    const double N = (double)ntot(sys), L = sys->L;
    if (sys->alpha <= 0.0) return -M_PI * N * N / (72.0 * L);
    return -M_PI * N * N / (2.0 * sys->alpha * sys->alpha * L * L * L);
}

/* energy_tracking.cpp:13-69: half-space G vectors in (mx, my, mz) loop order
   with weights 8 pi^2 / G^2 exp(-G^2 / 4 alpha^2).  recip_nmax2 > 0 replaces
   the tolerance filter by |n|^2 <= recip_nmax2. */
int orc_ewald_gvectors(const orc_system *sys, double *gx, double *gy, double *gz, double *w,
                       int *nvec, int cap, double *alpha_out) {
    const double L = sys->L;
    const double alpha = ewald_alpha(sys);
    if (alpha_out) *alpha_out = alpha;
    const double tpl = 2.0 * M_PI / L;
    const double fas = 4.0 * alpha * alpha;
    const double cutoff_factor = -log(1.0e-6);
    const double gmax2 = fas * cutoff_factor;
    int mmax;
    if (sys->recip_nmax2 > 0)
        mmax = (int)floor(sqrt((double)sys->recip_nmax2));
    else
        mmax = (int)ceil(sqrt(gmax2) / tpl) + 1;
    int count = 0;
    for (int mx = -mmax; mx <= mmax; ++mx)
        for (int my = -mmax; my <= mmax; ++my)
            for (int mz = -mmax; mz <= mmax; ++mz) {
                if (mx < 0) continue;
                if (mx == 0 && my < 0) continue;
                if (mx == 0 && my == 0 && mz <= 0) continue;
                const double cx = tpl * (double)mx, cy = tpl * (double)my, cz = tpl * (double)mz;
                const double g2 = cx * cx + cy * cy + cz * cz;
                if (sys->recip_nmax2 > 0) {
                    if (mx * mx + my * my + mz * mz > sys->recip_nmax2) continue;
                } else if (g2 > gmax2) {
                    continue;
                }
                if (gx && count < cap) {
                    gx[count] = cx;
                    gy[count] = cy;
                    gz[count] = cz;
                    w[count] = 8.0 * M_PI * M_PI / g2 * exp(-g2 / fas);
                    if (nvec) {
                        nvec[3 * count + 0] = mx;
                        nvec[3 * count + 1] = my;
                        nvec[3 * count + 2] = mz;
                    }
                }
                ++count;
            }
    return count;
}

/* energy_tracking.cpp:101-107,133-137: Abramowitz & Stegun 7.1.26. */
static double erfc_as(double x) {
    const double p = 0.3275911;
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
                 a5 = 1.061405429;
    const double t = 1.0 / (1.0 + p * x);
    const double tau = t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5))));
    return tau * exp(-x * x);
}

static double pair_real(const struct orc_walker *w, double r) {
    const double inv_r = (r < 1e-12) ? 1.0 : 1.0 / r;
    const double arg = w->alpha * r;
    const double e = (w->sys.erfc_kind == 0) ? erfc_as(arg) : erfc(arg);
    return e * inv_r;
}

static int in_cut(const struct orc_walker *w, double r) {
    return w->sys.rcut <= 0.0 || r < w->sys.rcut;
}

/* energy_tracking.cpp:93-144 (+ the optional spherical cutoff). */
static double real_energy_full(const struct orc_walker *w) {
    const int N = w->N;
    const double L = w->sys.L;
    double sum = 0.0;
    for (int i = 0; i < N; ++i) {
        double loc = 0.0;
        for (int j = i + 1; j < N; ++j) {
            const double dx = min_image(w->x[i] - w->x[j], L);
            const double dy = min_image(w->y[i] - w->y[j], L);
            const double dz = min_image(w->z[i] - w->z[j], L);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (!in_cut(w, r)) continue;
            loc += pair_real(w, r);
        }
        sum += loc;
    }
    return sum;
}

/* energy_tracking.cpp:146-191 */
static void structure_factors_full(struct orc_walker *w) {
    for (int g = 0; g < w->ng; ++g) {
        double cs = 0.0, sn = 0.0;
        for (int j = 0; j < w->N; ++j) {
            const double dot = w->Gx[g] * w->x[j] + w->Gy[g] * w->y[j] + w->Gz[g] * w->z[j];
            double s, c;
            sincos(dot, &s, &c);
            cs += c;
            sn += s;
        }
        w->Sre[g] = cs;
        w->Sim[g] = sn;
    }
}

/* energy_tracking.cpp:72-91 */
static double reciprocal_energy_full(const struct orc_walker *w) {
    double sum = 0.0;
    for (int g = 0; g < w->ng; ++g)
        sum += w->Gw[g] * (w->Sre[g] * w->Sre[g] + w->Sim[g] * w->Sim[g]);
    return w->pref * sum;
}

/* energy_tracking.cpp:193-251 */
static void structure_factors_update(struct orc_walker *w, double ox, double oy, double oz,
                                     double nx, double ny, double nz) {
    double delta = 0.0;
    for (int g = 0; g < w->ng; ++g) {
        const double od = w->Gx[g] * ox + w->Gy[g] * oy + w->Gz[g] * oz;
        const double nd = w->Gx[g] * nx + w->Gy[g] * ny + w->Gz[g] * nz;
        double ns, nc, os, oc;
        sincos(nd, &ns, &nc);
        sincos(od, &os, &oc);
        w->dre[g] = nc - oc;
        w->dim[g] = ns - os;
    }
    for (int g = 0; g < w->ng; ++g) {
        const double dr = w->dre[g], di = w->dim[g];
        delta += w->Gw[g] * (2.0 * (w->Sre[g] * dr + w->Sim[g] * di) + dr * dr + di * di);
        w->Sre[g] += dr;
        w->Sim[g] += di;
    }
    w->vrecip += w->pref * delta;
}

/* energy_tracking.cpp:280-365; positions already hold the new r_p. */
static void real_energy_update(struct orc_walker *w, int p, double ox, double oy, double oz) {
    const double L = w->sys.L;
    const double nx = w->x[p], ny = w->y[p], nz = w->z[p];
    double delta = 0.0;
    for (int j = 0; j < w->N; ++j) {
        const double valid = (j == p) ? 0.0 : 1.0;
        const double dxo = min_image(ox - w->x[j], L);
        const double dyo = min_image(oy - w->y[j], L);
        const double dzo = min_image(oz - w->z[j], L);
        const double ro = sqrt(dxo * dxo + dyo * dyo + dzo * dzo);
        const double eo = in_cut(w, ro) ? pair_real(w, ro) : 0.0;
        const double dxn = min_image(nx - w->x[j], L);
        const double dyn = min_image(ny - w->y[j], L);
        const double dzn = min_image(nz - w->z[j], L);
        const double rn = sqrt(dxn * dxn + dyn * dyn + dzn * dzn);
        const double en = in_cut(w, rn) ? pair_real(w, rn) : 0.0;
        delta += valid * (en - eo);
    }
    w->vreal += delta;
}

/* =========================================================== walker API */

orc_walker *orc_walker_new(const orc_system *sys) {
    struct orc_walker *w = (struct orc_walker *)calloc(1, sizeof(*w));
    w->sys = *sys;
    w->N = sys->n_up + sys->n_dn;
    w->ndet = sys->n_dn > 0 ? 2 : 1;
    det_init(&w->det[0], sys->n_up, 0, sys->L);
    det_init(&w->det[1], sys->n_dn, sys->n_up, sys->L);
    const size_t N = (size_t)w->N;
    w->x = (double *)calloc(N, sizeof(double));
    w->y = (double *)calloc(N, sizeof(double));
    w->z = (double *)calloc(N, sizeof(double));
    w->gx = (double *)calloc(N, sizeof(double));
    w->gy = (double *)calloc(N, sizeof(double));
    w->gz = (double *)calloc(N, sizeof(double));
    w->lap = (double *)calloc(N, sizeof(double));
    w->jgx = (double *)calloc(N, sizeof(double));
    w->jgy = (double *)calloc(N, sizeof(double));
    w->jgz = (double *)calloc(N, sizeof(double));
    w->jlap = (double *)calloc(N, sizeof(double));
    w->ng = orc_ewald_gvectors(sys, NULL, NULL, NULL, NULL, NULL, 0, &w->alpha);
    const size_t ng = (size_t)(w->ng > 0 ? w->ng : 1);
    w->Gx = (double *)calloc(ng, sizeof(double));
    w->Gy = (double *)calloc(ng, sizeof(double));
    w->Gz = (double *)calloc(ng, sizeof(double));
    w->Gw = (double *)calloc(ng, sizeof(double));
    w->Sre = (double *)calloc(ng, sizeof(double));
    w->Sim = (double *)calloc(ng, sizeof(double));
    w->dre = (double *)calloc(ng, sizeof(double));
    w->dim = (double *)calloc(ng, sizeof(double));
    orc_ewald_gvectors(sys, w->Gx, w->Gy, w->Gz, w->Gw, NULL, w->ng, &w->alpha);
    w->self = orc_ewald_self(sys);
    w->bg = orc_ewald_background(sys);
    w->pref = 1.0 / (2.0 * M_PI * sys->L * sys->L * sys->L); /* energy_tracking.cpp:74 */
    return w;
}

void orc_walker_free(orc_walker *w) {
    if (!w) return;
    det_free(&w->det[0]);
    det_free(&w->det[1]);
    free(w->x); free(w->y); free(w->z);
    free(w->gx); free(w->gy); free(w->gz); free(w->lap);
    free(w->jgx); free(w->jgy); free(w->jgz); free(w->jlap);
    free(w->Gx); free(w->Gy); free(w->Gz); free(w->Gw);
    free(w->Sre); free(w->Sim); free(w->dre); free(w->dim);
    free(w);
}

int orc_walker_num_electrons(const orc_walker *w) { return w->N; }
int orc_walker_num_g(const orc_walker *w) { return w->ng; }
int orc_walker_num_k(const orc_walker *w, int spin) { return w->det[spin].nk; }

static void refresh_ewald(struct orc_walker *w) {
    structure_factors_full(w);
    w->vrecip = reciprocal_energy_full(w);
    w->vreal = real_energy_full(w);
}

int orc_walker_set_positions(orc_walker *w, const double *pos) {
    for (int i = 0; i < w->N; ++i) {
        w->x[i] = pos[3 * i + 0];
        w->y[i] = pos[3 * i + 1];
        w->z[i] = pos[3 * i + 2];
    }
    int singular = 0;
    for (int d = 0; d < w->ndet; ++d) {
        double ld = det_log_abs_det(&w->det[d], w->x, w->y, w->z);
        if (!isfinite(ld)) singular = 1;
    }
    refresh_ewald(w);
    w->jastrow_cache_valid = 0;
    w->steps_since_refresh = 0;
    return singular;
}

void orc_walker_get_positions(const orc_walker *w, double *pos) {
    for (int i = 0; i < w->N; ++i) {
        pos[3 * i + 0] = w->x[i];
        pos[3 * i + 1] = w->y[i];
        pos[3 * i + 2] = w->z[i];
    }
}

double orc_walker_logdet(const orc_walker *w, int spin) { return w->det[spin].logdet; }

void orc_walker_inverse(const orc_walker *w, int spin, double *inv_t) {
    const det_t *d = &w->det[spin];
    memcpy(inv_t, d->inv, sizeof(double) * (size_t)d->n * d->n);
}

void orc_walker_dmatrix(const orc_walker *w, int spin, double *dm) {
    const det_t *d = &w->det[spin];
    memcpy(dm, d->D, sizeof(double) * (size_t)d->n * d->n);
}

double orc_walker_jastrow(const orc_walker *w) { return jastrow_value(w); }
double orc_walker_vreal(const orc_walker *w) { return w->vreal; }
double orc_walker_vrecip(const orc_walker *w) { return w->vrecip; }

void orc_walker_structure_factor(const orc_walker *w, double *s_re, double *s_im) {
    memcpy(s_re, w->Sre, sizeof(double) * (size_t)w->ng);
    memcpy(s_im, w->Sim, sizeof(double) * (size_t)w->ng);
}

static void copy_out_derivs(const struct orc_walker *w, double *grad, double *lap) {
    for (int i = 0; i < w->N; ++i) {
        if (grad) {
            grad[3 * i + 0] = w->gx[i];
            grad[3 * i + 1] = w->gy[i];
            grad[3 * i + 2] = w->gz[i];
        }
        if (lap) lap[i] = w->lap[i];
    }
}

/* wavefunction.cpp:12-58 */
static void derivatives_full(struct orc_walker *w) {
    const size_t N = (size_t)w->N;
    memset(w->gx, 0, N * sizeof(double));
    memset(w->gy, 0, N * sizeof(double));
    memset(w->gz, 0, N * sizeof(double));
    memset(w->lap, 0, N * sizeof(double));
    memset(w->jgx, 0, N * sizeof(double));
    memset(w->jgy, 0, N * sizeof(double));
    memset(w->jgz, 0, N * sizeof(double));
    memset(w->jlap, 0, N * sizeof(double));
    for (int d = 0; d < w->ndet; ++d) det_add_derivatives(&w->det[d], w->gx, w->gy, w->gz, w->lap);
    jastrow_add_derivatives(w, w->jgx, w->jgy, w->jgz, w->jlap);
    for (int i = 0; i < w->N; ++i) {
        w->gx[i] += w->jgx[i];
        w->gy[i] += w->jgy[i];
        w->gz[i] += w->jgz[i];
        w->lap[i] += w->jlap[i];
    }
    w->jastrow_cache_valid = 1;
    w->steps_since_refresh = 0;
}

void orc_walker_derivatives(orc_walker *w, double *grad, double *lap) {
    derivatives_full(w);
    copy_out_derivs(w, grad, lap);
}

/* wavefunction.cpp:60-114 */
void orc_walker_derivatives_after_step(orc_walker *w, int accepted, int p, const double oldpos[3],
                                       double *grad, double *lap) {
    if (!w->jastrow_cache_valid || w->steps_since_refresh >= 500) {
        derivatives_full(w);
        copy_out_derivs(w, grad, lap);
        return;
    }
    w->steps_since_refresh += 1;
    if (accepted) {
        jastrow_update_for_move(w, p, oldpos[0], oldpos[1], oldpos[2], w->jgx, w->jgy, w->jgz,
                                w->jlap);
        const size_t N = (size_t)w->N;
        memset(w->gx, 0, N * sizeof(double));
        memset(w->gy, 0, N * sizeof(double));
        memset(w->gz, 0, N * sizeof(double));
        memset(w->lap, 0, N * sizeof(double));
        for (int d = 0; d < w->ndet; ++d)
            det_add_derivatives(&w->det[d], w->gx, w->gy, w->gz, w->lap);
        for (int i = 0; i < w->N; ++i) {
            w->gx[i] += w->jgx[i];
            w->gy[i] += w->jgy[i];
            w->gz[i] += w->jgz[i];
            w->lap[i] += w->jlap[i];
        }
    }
    copy_out_derivs(w, grad, lap);
}

/* energy_tracking.cpp:253-278 */
double orc_walker_kinetic(const orc_walker *w) {
    double t = 0.0;
    for (int i = 0; i < w->N; ++i) {
        const double g2 = w->gx[i] * w->gx[i] + w->gy[i] * w->gy[i] + w->gz[i] * w->gz[i];
        t += (w->lap[i] + g2);
    }
    return -0.5 * t;
}

/* energy_tracking.cpp:367-382 */
static double potential(const struct orc_walker *w) {
    return w->vreal + w->vrecip + w->self + w->bg;
}

double orc_walker_local_energy(orc_walker *w, double *kinetic, double *pot) {
    derivatives_full(w);
    const double t = orc_walker_kinetic(w);
    const double v = potential(w);
    if (kinetic) *kinetic = t;
    if (pot) *pot = v;
    return t + v;
}

static det_t *det_of(struct orc_walker *w, int p, int *pl) {
    const int s = spin_of(w, p);
    *pl = p - w->det[s].first;
    return &w->det[s];
}

double orc_walker_ratio(orc_walker *w, int p, const double np[3]) {
    int pl;
    det_t *d = det_of(w, p, &pl);
    return det_ratio(d, pl, np[0], np[1], np[2]);
}

double orc_walker_delta_j(const orc_walker *w, int p, const double np[3]) {
    return jastrow_delta(w, p, w->x[p], w->y[p], w->z[p], np[0], np[1], np[2]);
}

/* simulation.cpp:133-145 */
void orc_walker_accept(orc_walker *w, int p, const double np[3]) {
    int pl;
    det_t *d = det_of(w, p, &pl);
    const double ratio = det_ratio(d, pl, np[0], np[1], np[2]);
    const double ox = w->x[p], oy = w->y[p], oz = w->z[p];
    w->x[p] = np[0];
    w->y[p] = np[1];
    w->z[p] = np[2];
    d->logdet += log(fabs(ratio));
    det_accept(d, pl, ratio);
    structure_factors_update(w, ox, oy, oz, np[0], np[1], np[2]);
    real_energy_update(w, p, ox, oy, oz);
}

/* ====================================================== blocking analysis */

/* blocking_analysis.cpp:3-5 */
void orc_blocking_init(orc_blocking *b, size_t block_size) {
    memset(b, 0, sizeof(*b));
    b->block_size = block_size;
}

/* blocking_analysis.cpp:17-39 */
void orc_blocking_add(orc_blocking *b, double x) {
    b->block_sum += x;
    ++b->in_block;
    if (b->in_block == b->block_size) {
        const double bm = b->block_sum / (double)b->block_size;
        ++b->num_blocks;
        const double delta = bm - b->running_mean;
        b->running_mean += delta / (double)b->num_blocks;
        b->running_m2 += delta * (bm - b->running_mean);
        b->block_sum = 0.0;
        b->in_block = 0;
    }
}

/* blocking_analysis.cpp:41 */
int orc_blocking_ready(const orc_blocking *b) { return b->num_blocks >= 2; }

/* blocking_analysis.cpp:7-15 */
int orc_blocking_mean_se(const orc_blocking *b, double *mean, double *se) {
    if (b->num_blocks < 2) return 1;
    const double var = b->running_m2 / (double)(b->num_blocks - 1);
    *se = sqrt(var / (double)b->num_blocks);
    *mean = b->running_mean;
    return 0;
}

/* ================================================ reference simulation run */

typedef struct {
    orc_mt64 rng;
    double step;
} ref_rng;

/* simulation.cpp:80-156 */
static int ref_metropolis_step(struct orc_walker *w, ref_rng *r, int *moved, double old[3]) {
    const int N = w->N;
    const double L = w->sys.L;
    const double inv_L = 1.0 / L;
    const int p = (int)orc_std_uniform_index(&r->rng, (uint64_t)N);
    const double ox = w->x[p], oy = w->y[p], oz = w->z[p];
    double nx = ox + orc_std_uniform_real(&r->rng, -r->step, r->step);
    double ny = oy + orc_std_uniform_real(&r->rng, -r->step, r->step);
    double nz = oz + orc_std_uniform_real(&r->rng, -r->step, r->step);
    nx -= L * floor(nx * inv_L);
    ny -= L * floor(ny * inv_L);
    nz -= L * floor(nz * inv_L);
    int pl;
    det_t *d = det_of(w, p, &pl);
    const double ratio = det_ratio(d, pl, nx, ny, nz);
    const double dj = jastrow_delta(w, p, ox, oy, oz, nx, ny, nz);
    const double lr2 = 2.0 * log(fabs(ratio)) + 2.0 * dj;
    double u = orc_std_canonical(&r->rng);
    if (u < DBL_MIN) u = DBL_MIN;
    const double lu = log(u);
    const double mt = lr2 < 0.0 ? lr2 : 0.0;
    const int acc = lu < mt;
    *moved = p;
    old[0] = ox; old[1] = oy; old[2] = oz;
    if (acc) {
        w->x[p] = nx; w->y[p] = ny; w->z[p] = nz;
        d->logdet += log(fabs(ratio));
        det_accept(d, pl, ratio);
        structure_factors_update(w, ox, oy, oz, nx, ny, nz);
        real_energy_update(w, p, ox, oy, oz);
    }
    return acc;
}

int orc_ref_simulation_run(const orc_ref_config *cfg, orc_ref_summary *out, double *frames_e,
                           double *frames_mean, double *frames_se, double *frames_acc,
                           double *frames_pos) {
    orc_system sys;
    memset(&sys, 0, sizeof(sys));
    sys.n_up = (int)cfg->num_particles;
    sys.n_dn = 0;
    sys.L = cfg->box_length;
    sys.a_same = cfg->jastrow_a;
    sys.a_anti = cfg->jastrow_a;
    sys.b = cfg->jastrow_b;
    sys.alpha = 0.0;
    sys.recip_nmax2 = 0;
    sys.rcut = 0.0;
    sys.erfc_kind = 0;
    struct orc_walker *w = orc_walker_new(&sys);
    ref_rng r;
    orc_mt64_seed(&r.rng, cfg->seed);
    r.step = cfg->step_size;
    const int N = w->N;
    const double L = cfg->box_length;
    /* simulation.cpp:43-76 */
    double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    int ok = 0;
    for (int attempt = 0; attempt < 100; ++attempt) {
        for (int i = 0; i < N; ++i) {
            pos[3 * i + 0] = orc_std_canonical(&r.rng) * L;
            pos[3 * i + 1] = orc_std_canonical(&r.rng) * L;
            pos[3 * i + 2] = orc_std_canonical(&r.rng) * L;
        }
        int sing = orc_walker_set_positions(w, pos);
        const double lpsi = w->det[0].logdet + jastrow_value(w);
        if (!sing && isfinite(lpsi)) { ok = 1; break; }
    }
    free(pos);
    if (!ok) { orc_walker_free(w); return 1; }
    /* simulation.cpp:159-196 */
    size_t wp = 0, wa = 0;
    int moved;
    double old[3];
    for (size_t i = 0; i < cfg->warmup_steps; ++i) {
        ++wp;
        if (ref_metropolis_step(w, &r, &moved, old)) ++wa;
        if (wp % (size_t)N == 0) {
            const double rate = (double)wa / (double)wp;
            r.step *= exp(0.25 * (rate - 0.5));
            if (r.step > 0.5 * L) r.step = 0.5 * L;
            wa = 0;
            wp = 0;
        }
    }
    /* simulation.cpp:198-263 */
    orc_blocking blk;
    orc_blocking_init(&blk, cfg->block_size);
    size_t proposed = 0, accepted = 0;
    double run_sum = 0.0, final_mean = 0.0, final_se = 0.0;
    int have_se = 0;
    for (size_t i = 0; i < cfg->measure_steps; ++i) {
        ++proposed;
        const int acc = ref_metropolis_step(w, &r, &moved, old);
        if (acc) ++accepted;
        orc_walker_derivatives_after_step(w, acc, moved, old, NULL, NULL);
        const double e = orc_walker_kinetic(w) + potential(w);
        run_sum += e;
        orc_blocking_add(&blk, e);
        const double rmean = run_sum / (double)(i + 1);
        final_mean = rmean;
        double fse = 0.0;
        if (orc_blocking_ready(&blk)) {
            double bm, se;
            orc_blocking_mean_se(&blk, &bm, &se);
            final_mean = bm;
            final_se = se;
            have_se = 1;
            fse = se;
        }
        if (frames_e) frames_e[i] = e;
        if (frames_mean) frames_mean[i] = rmean;
        if (frames_se) frames_se[i] = fse;
        if (frames_acc) frames_acc[i] = (double)accepted / (double)proposed;
        if (frames_pos) orc_walker_get_positions(w, frames_pos + 3 * (size_t)N * i);
    }
    out->mean_energy = final_mean;
    out->standard_error = final_se;
    out->has_standard_error = have_se;
    out->acceptance_rate = proposed ? (double)accepted / (double)proposed : 0.0;
    out->accepted = accepted;
    out->proposed = proposed;
    out->final_step_size = r.step;
    orc_walker_free(w);
    return 0;
}

/* ================================================= ensemble sweep (GPU twin) */

/* Stream tags of the Philox counter (word 3). */
#define TAG_MOVE 0x4D4F5645u /* "MOVE" */
#define TAG_ACCP 0x41434350u /* "ACCP" */
#define TAG_INIT 0x494E4954u /* "INIT" */

static double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

static void philox_draw(uint64_t seed, uint32_t c0, uint64_t walker, uint64_t c2, uint32_t tag,
                        uint32_t out[4]) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {c0, (uint32_t)walker, (uint32_t)c2, tag};
    orc_philox4x32_10(ctr, key, out);
}

int orc_init_positions(const orc_system *sys, uint64_t seed, uint64_t walker, double *pos) {
    struct orc_walker *w = orc_walker_new(sys);
    const int N = w->N;
    int used = -1;
    for (int attempt = 0; attempt < 100; ++attempt) {
        for (int i = 0; i < N; ++i) {
            uint32_t r[4];
            philox_draw(seed, (uint32_t)i, walker, (uint64_t)attempt, TAG_INIT, r);
            pos[3 * i + 0] = u01(r[0]) * sys->L;
            pos[3 * i + 1] = u01(r[1]) * sys->L;
            pos[3 * i + 2] = u01(r[2]) * sys->L;
        }
        if (orc_walker_set_positions(w, pos) == 0) { used = attempt; break; }
    }
    orc_walker_free(w);
    return used;
}

/* Electron moved at step t of a sweep: 0..N-1 in order, except that with
   equal spins the two spins alternate (up 0, down 0, up 1, down 1, ...). */
static int sweep_visit(int n_up, int n_dn, int t) {
    if (n_up != n_dn) return t;
    return (t & 1) ? n_up + (t >> 1) : (t >> 1);
}

int orc_sweep(orc_walker *w, const orc_sweep_config *cfg, uint64_t walker, uint64_t sweep,
              double *step, int adapt, double *e_local) {
    const int N = w->N;
    const double L = w->sys.L;
    const double inv_L = 1.0 / L;
    int accepted = 0;
    for (int t = 0; t < N; ++t) {
        const int p = sweep_visit(w->sys.n_up, w->sys.n_dn, t);
        uint32_t r[4], ra[4];
        philox_draw(cfg->seed, (uint32_t)p, walker, sweep, TAG_MOVE, r);
        philox_draw(cfg->seed, (uint32_t)p, walker, sweep, TAG_ACCP, ra);
        double dx, dy, dz;
        if (cfg->move_kind == ORC_MOVE_GAUSSIAN) {
            const double rho0 = sqrt(-2.0 * log(u01(r[0])));
            const double rho1 = sqrt(-2.0 * log(u01(r[2])));
            const double t0 = 2.0 * M_PI * u01(r[1]);
            const double t1 = 2.0 * M_PI * u01(r[3]);
            dx = *step * (rho0 * cos(t0));
            dy = *step * (rho0 * sin(t0));
            dz = *step * (rho1 * cos(t1));
        } else {
            dx = *step * (2.0 * u01(r[0]) - 1.0);
            dy = *step * (2.0 * u01(r[1]) - 1.0);
            dz = *step * (2.0 * u01(r[2]) - 1.0);
        }
        double np[3] = {w->x[p] + dx, w->y[p] + dy, w->z[p] + dz};
        np[0] -= L * floor(np[0] * inv_L);
        np[1] -= L * floor(np[1] * inv_L);
        np[2] -= L * floor(np[2] * inv_L);
        int pl;
        det_t *d = det_of(w, p, &pl);
        const double ratio = det_ratio(d, pl, np[0], np[1], np[2]);
        const double dj = jastrow_delta(w, p, w->x[p], w->y[p], w->z[p], np[0], np[1], np[2]);
        const double prob = ratio * ratio * exp(2.0 * dj);
        if (u01(ra[0]) < prob) {
            const double ox = w->x[p], oy = w->y[p], oz = w->z[p];
            w->x[p] = np[0]; w->y[p] = np[1]; w->z[p] = np[2];
            d->logdet += log(fabs(ratio));
            det_accept(d, pl, ratio);
            (void)ox; (void)oy; (void)oz;
            ++accepted;
        }
    }
    if (adapt) {
        *step *= exp(0.25 * ((double)accepted / (double)N - 0.5));
        if (*step > 0.5 * L) *step = 0.5 * L;
    }
    if (e_local) {
        refresh_ewald(w);
        *e_local = orc_walker_local_energy(w, NULL, NULL);
    }
    return accepted;
}

/* ============================================== threaded CPU baseline timing */

#include <pthread.h>
#include <time.h>

typedef struct {
    const orc_system *sys;
    orc_sweep_config cfg;
    uint64_t walker;
    int sweeps;
    int measure;
    double seconds;
    long long moves;
    double e_sum;
} bench_job;

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void *bench_thread(void *arg) {
    bench_job *j = (bench_job *)arg;
    double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)(j->sys->n_up + j->sys->n_dn));
    orc_init_positions(j->sys, j->cfg.seed, j->walker, pos);
    struct orc_walker *w = orc_walker_new(j->sys);
    orc_walker_set_positions(w, pos);
    double step = j->cfg.step0;
    const double t0 = now_s();
    for (int s = 0; s < j->sweeps; ++s) {
        double e = 0.0;
        orc_sweep(w, &j->cfg, j->walker, (uint64_t)s, &step, 0, j->measure ? &e : NULL);
        j->e_sum += e;
    }
    j->seconds = now_s() - t0;
    j->moves = (long long)j->sweeps * w->N;
    orc_walker_free(w);
    free(pos);
    return NULL;
}

/* One independent chain per thread (main.cpp:50-68), each `sweeps` sweeps of
   orc_sweep with a measurement per sweep when measure != 0.  Returns the
   aggregate proposed moves per second over the slowest thread's time. */
double orc_bench_threads(const orc_system *sys, const orc_sweep_config *cfg, int threads,
                         int sweeps, int measure, double *elapsed) {
    pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    bench_job *jobs = (bench_job *)calloc((size_t)threads, sizeof(bench_job));
    const double t0 = now_s();
    for (int t = 0; t < threads; ++t) {
        jobs[t].sys = sys;
        jobs[t].cfg = *cfg;
        jobs[t].walker = (uint64_t)t;
        jobs[t].sweeps = sweeps;
        jobs[t].measure = measure;
        pthread_create(&tid[t], NULL, bench_thread, &jobs[t]);
    }
    long long moves = 0;
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], NULL);
        moves += jobs[t].moves;
    }
    const double wall = now_s() - t0;
    if (elapsed) *elapsed = wall;
    free(tid);
    free(jobs);
    return (double)moves / wall;
}
