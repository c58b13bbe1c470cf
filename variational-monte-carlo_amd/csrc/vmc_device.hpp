// vmc_device.hpp -- wavefront-level building blocks of the gfx950 VMC sweep.
//
// One wavefront (64 lanes) owns one walker.  Lanes stride over electrons,
// orbitals or G vectors; sums are butterfly-reduced so every lane holds the
// identical value, which keeps accept/reject decisions wave-uniform.
//
// `Real` is the working precision of orbitals, ratios, Sherman-Morrison,
// Jastrow and real-space Ewald (double in the FP64 path, float in the mixed
// path).  Sums of physical terms always accumulate in double.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vmc {

constexpr int kWave = 64;
constexpr int kStatCount = 16;

// ------------------------------------------------------------------ system
// Steps of one structure-factor column segment (DevSys::sfseg): a fixed count
// per thread, absent steps carrying zero weights.
constexpr int kSfSegSteps = 4;

struct DevSys {
    int N, n_up, n_dn, ndet;
    int n[2];        // electrons (= orbitals) per determinant
    int first[2];    // index of the determinant's first electron
    int inv_off[2];  // offset of the determinant's inverse in a walker's block
    int ninv;        // n[0]^2 + n[1]^2
    int nmax;
    int ng;
    int erfc_kind;   // 0 = Abramowitz-Stegun 7.1.26, 1 = erfc
    int move_kind;   // 0 = uniform cube, 1 = Gaussian
    int use_cut;
    int reinv_k;     // due threshold of the walker's accepted-move counter, summed over its
                     // determinants: K x ndet (0 = never re-invert)
    double L, inv_L, half_L;
    double a_same, a_anti, b;
    double alpha, rcut;
    double e_const;  // Ewald self + background
    double pref;     // 1 / (2 pi L^3)
    double kunit;    // 2 pi / L
    uint32_t key0, key1;
    const int4 *orb;     // [ndet][nmax] = {nx, ny, nz, type (0 cos, 1 sin)}
    const int4 *gvec;    // [ng] = {mx, my, mz, 0}
    const double *gw;    // [ng] weights 8 pi^2 / G^2 exp(-G^2 / 4 alpha^2)
    // The same G vectors as column segments (vmc_tables.hpp sf_segments):
    // [nseg] = {mx, my, a0, steps} and [nseg][kSfSegSteps][2] weights.
    int nseg;
    const int4 *sfseg;
    const double *sfw;
};

// Per-ensemble device state (walker-major SoA in HBM).
struct DevState {
    int W;
    double *pos;      // [W][3][N]
    void *inv;        // Real [W][ninv]
    double *sigma;    // [W]
    double *stats;    // [W][kStatCount]
    double *logdet;   // [W][2]
    int *since;       // [W] accepted moves since the last FP64 re-inversion
    int *flags;       // [W] bit0: singular determinant met
    double *scratch;  // [W][nmax^2] FP64 workspace
    double *drift;    // [W] largest |I - D^-1 D| met before a re-inversion
    float *resid;     // [W][P^2] Newton-Schulz residual above 80 per spin (mixed), else null
    double *jrow;     // [W][N] Jastrow row sums U_i kept by the register sweeps between sweeps
    unsigned *due;    // [2] walkers listed in due_list, finished workgroups of the re-inversion
    int *due_list;    // [W] walkers a register sweep found due for re-inversion
    // [W][ninv] FP64 inverses for the FP64 local energy of a mixed-precision
    // measured sweep (multi-wavefront and generic kernels).  Set only in the
    // state handed to that sweep's re-inversion, which then re-inverts every
    // walker and writes its FP64 result here; null otherwise.
    double *inv64;
    // With inv64 set: 1 when only inv64 is written (the measurement entry
    // point, which leaves the walkers' state alone), 0 when the re-inversion
    // also replaces the stored inverse and its bookkeeping (a measured sweep).
    int inv64_only = 0;
    // Ensemble index of this state's walker 0: the register sweep keys its
    // Philox draws by w0 + w, so a view of a walker range (a walker group
    // swept on a stream of its own) draws exactly what the whole ensemble
    // would.
    int w0 = 0;
};

// Caller-supplied moves of one sweep (vmc_ensemble_sweep_given): electron e
// of walker w is proposed at prop[(w N + e) 3 + c] and accepted when
// u[w N + e] < R^2 e^{2 dJ}; trace[(w N + e) kTraceCount + k] receives the
// move's ratio, Jastrow change and decision.  All null in the Philox sweeps.
struct SweepGiven {
    const double *prop;
    const double *u;
    double *trace;
};

// The reference's measurement cadence for one walker (vmc_ensemble_sweep_traced):
// after each of walker `walker`'s moves, in sweep order t, its local energy
// e_local[t], the moved electron moved[t], the decision accepted[t] and, with
// pos set, its positions pos[t][N][3].  walker < 0: no trace.
struct StepTrace {
    long long walker;
    double *e_local;
    int *moved;
    int *accepted;
    double *pos;
};
constexpr int kTraceCount = 3;  // VMC_MV_COUNT

// Per-walker parts of the local energy written by a measurement launched
// for vmc_ensemble_measure (statistics untouched): kinetic, V_real, V_recip,
// E_L, log|det| of each determinant, J.
constexpr int kPartCount = 7;  // VMC_PART_COUNT

// log|det| of the two determinants tracked through the accepted moves'
// ratios: the product of |R| kept as mantissa and binary exponent (no log
// on the move's path), folded into log|det| at the end of the sweep.
struct LogDetTrack {
    double m[2] = {1.0, 1.0};
    int ex[2] = {0, 0};
    __device__ void add(int h, double R) {
        int e;
        m[h] = frexp(m[h] * fabs(R), &e);
        ex[h] += e;
    }
    __device__ double fold(int h, double ld) const {
        return ld + log(m[h]) + (double)ex[h] * 0.69314718055994531;
    }
};

// Internal sweep flag (beside VMC_SWEEP_*): the register sweep lists the
// walkers that reach K accepted moves in DevState::due_list for the compacted
// re-inversion launch that follows it.
constexpr int kSweepListDue = 1 << 30;

// Electron moved at step t of a sweep: 0..N-1 in order, except that with
// equal spins the two spins alternate (up 0, down 0, up 1, down 1, ...), so a
// register sweep can evaluate an up and a down move together.
__host__ __device__ constexpr int sweep_visit(int n_up, int n_dn, int t) {
    return n_up != n_dn ? t : ((t & 1) ? n_up + (t >> 1) : (t >> 1));
}

// Walker flag bits (DevState::flags).
constexpr int kFlagSingular = 1;  // a singular determinant was met
constexpr int kFlagJrowValid = 2;  // jrow holds the current positions' row sums
// The register sweeps recompute the Jastrow row sums from the positions every
// kJrowRefresh sweeps (bounding the drift of their incremental updates) and
// whenever another kernel moved the walker.
constexpr uint32_t kJrowRefresh = 16;

// One system of a batch group: its tables and state, whether its inverses
// live in LDS (generic kernel), and its walkers' range in the group's walker
// numbering.
struct BatchItem {
    DevSys s;
    DevState st;
    int inv_lds;
    int walker_begin;
    int walker_end;
};

// Index of the item whose walker range holds g (items in walker order).
__device__ inline int batch_find(const BatchItem *items, int n_items, int g) {
    int i = 0;
    while (i + 1 < n_items && g >= items[i].walker_end) ++i;
    return i;
}

// -------------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ inline int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ inline double wave_uniform(double v) {
    union { double d; int i[2]; } u;
    u.d = v;
    u.i[0] = __builtin_amdgcn_readfirstlane(u.i[0]);
    u.i[1] = __builtin_amdgcn_readfirstlane(u.i[1]);
    return u.d;
}

// Wave-wide barrier for LDS traffic between lanes of one wavefront.  The
// kernels run one wavefront per workgroup, so the workgroup barrier is the
// wavefront barrier.
__device__ inline void wave_sync() { __syncthreads(); }

// Sum over the whole workgroup (one walker may own several wavefronts): the
// wave sums, then the per-wave totals in a fixed order, so every thread gets
// the same value.
__device__ inline double block_sum(double v) {
    v = wave_sum(v);
    if (blockDim.x == kWave) return v;
    __shared__ double red[16];
    const int wv = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) red[wv] = v;
    __syncthreads();
    double t = red[0];
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i) t += red[i];
    return t;
}

// Block-wide max of v (NaN wins); every thread gets the result.
__device__ inline double block_max(double v, double *red) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = (o > v || o != o) ? o : v;
    }
    const int wv = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) red[wv] = v;
    __syncthreads();
    double r = red[0];
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i) r = (red[i] > r || red[i] != red[i]) ? red[i] : r;
    return r;
}

// Block-wide argmax of (a, idx), ties to the lower index.
__device__ inline void block_argmax(double &a, int &idx, double *red) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double oa = __shfl_xor(a, off, kWave);
        const int oi = __shfl_xor(idx, off, kWave);
        if (oa > a || (oa == a && oi < idx)) { a = oa; idx = oi; }
    }
    const int wv = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[2 * wv] = a;
        red[2 * wv + 1] = (double)idx;
    }
    __syncthreads();
    a = red[0];
    idx = (int)red[1];
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i) {
        const double oa = red[2 * i];
        const int oi = (int)red[2 * i + 1];
        if (oa > a || (oa == a && oi < idx)) { a = oa; idx = oi; }
    }
}

// Gauss-Jordan inversion with partial pivoting of the n x n row-major FP64
// matrix A (LDS or global memory) in place, by the whole workgroup; the
// reference's LU tolerance, a pivot below 1e-12 is singular
// (src/utilities/matrix.hpp:35-39).  gjcol holds n doubles and perm n ints of
// workspace.  Returns log|det A|, or -inf for a singular matrix (A is then
// left part-way).
__device__ inline double gj_invert(double *A, int n, double *gjcol, int *perm, double *red) {
    const int t = threadIdx.x, nt = blockDim.x;
    double logdet = 0.0;
    for (int c = 0; c < n; ++c) {
        double best = -1.0;
        int bi = n;
        for (int r = c + t; r < n; r += nt) {
            const double a = fabs(A[r * n + c]);
            if (a > best) { best = a; bi = r; }
        }
        block_argmax(best, bi, red);
        if (!(best >= 1e-12)) return -INFINITY;
        if (t == 0) perm[c] = bi;
        if (bi != c)
            for (int j = t; j < n; j += nt) {
                const double tmp = A[c * n + j];
                A[c * n + j] = A[bi * n + j];
                A[bi * n + j] = tmp;
            }
        __syncthreads();
        const double piv = A[c * n + c];
        logdet += log(fabs(piv));
        const double pinv = 1.0 / piv;
        for (int r = t; r < n; r += nt) gjcol[r] = A[r * n + c];
        __syncthreads();
        for (int j = t; j < n; j += nt) A[c * n + j] = (j == c ? 1.0 : A[c * n + j]) * pinv;
        __syncthreads();
        for (int e = t; e < n * n; e += nt) {
            const int r = e / n, j = e - (e / n) * n;
            if (r == c) continue;
            A[e] = (j == c ? 0.0 : A[e]) - gjcol[r] * A[c * n + j];
        }
        __syncthreads();
    }
    for (int c = n - 1; c >= 0; --c) {
        const int q = perm[c];
        if (q != c)
            for (int r = t; r < n; r += nt) {
                const double tmp = A[r * n + c];
                A[r * n + c] = A[r * n + q];
                A[r * n + q] = tmp;
            }
        __syncthreads();
    }
    return logdet;
}

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) with the published
// multipliers 0xD2511F53, 0xCD9E8D57 and Weyl keys 0x9E3779B9, 0xBB67AE85.
__device__ inline uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c.x;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Stream tags (counter word 3): which draw of a (electron, walker, sweep).
constexpr uint32_t kTagMove = 0x4D4F5645u;  // "MOVE"
constexpr uint32_t kTagAccept = 0x41434350u; // "ACCP"
constexpr uint32_t kTagInit = 0x494E4954u;   // "INIT"

__device__ inline double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

// ----------------------------------------------------------------- geometry
// Minimum image by the reference's mask form (jastrow_pade.cpp:37-39):
// d + L [d <= -L/2] - L [d > L/2], valid for |d| < L.
__device__ inline double min_image(double d, const DevSys &s) {
    return d + s.L * (double)(d <= -s.half_L) - s.L * (double)(d > s.half_L);
}

// Fractional phase n . r / L reduced to [-1/2, 1/2]; the orbital and
// structure-factor phases are 2 pi times it.
__device__ inline double frac_phase(int nx, int ny, int nz, double ux, double uy, double uz) {
    double t = (double)nx * ux + (double)ny * uy + (double)nz * uz;
    return t - rint(t);
}

template <typename Real>
__device__ inline void sincos_2pi(double t, Real *s, Real *c);

template <>
__device__ inline void sincos_2pi<double>(double t, double *s, double *c) {
    sincospi(2.0 * t, s, c);
}

template <>
__device__ inline void sincos_2pi<float>(double t, float *s, float *c) {
    sincospif((float)(2.0 * t), s, c);
}

template <typename Real>
__device__ inline Real rsqrt_r(Real x);
template <>
__device__ inline double rsqrt_r<double>(double x) { return 1.0 / sqrt(x); }
template <>
__device__ inline float rsqrt_r<float>(float x) { return 1.0f / sqrtf(x); }

// Working-precision sqrt, reciprocal and exp: IEEE in double; in float the
// hardware v_sqrt_f32 / v_rcp_f32 / v_exp_f32 (about 1 ulp), well inside the
// mixed path's error budget.
template <typename Real>
__device__ __forceinline__ Real wsqrt(Real x);
template <>
__device__ __forceinline__ double wsqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float wsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename Real>
__device__ __forceinline__ Real wrcp(Real x);
template <>
__device__ __forceinline__ double wrcp<double>(double x) { return 1.0 / x; }
template <>
__device__ __forceinline__ float wrcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
// FP64 reciprocal and inverse square root from the hardware approximations
// (v_rcp_f64, v_rsq_f64) with two Newton steps each: within about an ulp,
// without the scaling and special-case paths of IEEE division and sqrt.
// rcp_nr needs a finite non-zero argument, rsqrt_nr a positive normal one.
__device__ __forceinline__ double rcp_nr(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}
__device__ __forceinline__ double rsqrt_nr(double x) {
    const double hx = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = fma(y, fma(-hx * y, y, 0.5), y);
    return fma(y, fma(-hx * y, y, 0.5), y);
}
template <typename Real>
__device__ __forceinline__ Real wexp(Real x);
template <>
__device__ __forceinline__ double wexp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float wexp<float>(float x) { return __expf(x); }

// Abramowitz & Stegun 7.1.26 with the reference's constants
// (energy_tracking.cpp:101-107,133-137).
template <typename Real>
__device__ inline Real erfc_as7126(Real x) {
    const Real p = (Real)0.3275911;
    const Real a1 = (Real)0.254829592, a2 = (Real)-0.284496736, a3 = (Real)1.421413741,
               a4 = (Real)-1.453152027, a5 = (Real)1.061405429;
    const Real t = wrcp<Real>((Real)1 + p * x);
    const Real tau = t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5))));
    return tau * wexp<Real>(-x * x);
}

// FP32 erfc for 0 <= x <= kErfcFitMax (alpha r inside the cutoff radius L/2
// with alpha = 5/L): erfc(x) = exp(-x^2) P(t), t = (x - 1.5) / (x + 1.5),
// where P is a degree-8 least-squares Chebyshev fit of erfcx(x) = exp(x^2)
// erfc(x) (coefficients and their FP32 error from fit_erfc.py: P within
// 2.0e-7 relative).  exp(-x^2) keeps the rounding of x^2 and of x^2 log2(e)
// as a first-order correction, so the exponential factor adds about one
// rounding of v_exp_f32.
constexpr float kErfcFitMax = 2.5f;

__device__ inline float erfc_fit(float x) {
    const float t = (x - 1.5f) * __builtin_amdgcn_rcpf(x + 1.5f);
    float p = 1.670535858e-04f;
    p = fmaf(p, t, 1.198613783e-03f);
    p = fmaf(p, t, 2.984628081e-03f);
    p = fmaf(p, t, -4.575148632e-04f);
    p = fmaf(p, t, -1.864113472e-02f);
    p = fmaf(p, t, -9.285487235e-03f);
    p = fmaf(p, t, 1.944909096e-01f);
    p = fmaf(p, t, -4.908687174e-01f);
    p = fmaf(p, t, 3.215854168e-01f);
    constexpr float kLog2e = 1.44269502e+00f, kLog2eLo = 1.92596303e-08f, kLn2 = 6.93147182e-01f;
    const float h = x * x, l = fmaf(x, x, -h);       // x^2 = h + l
    const float y = -h * kLog2e;                      // exponent of 2, rounded
    float ye = fmaf(-h, kLog2e, -y);                  // its rounding
    ye = fmaf(-h, kLog2eLo, ye);
    ye = fmaf(-l, kLog2e, ye);
    const float e = __builtin_amdgcn_exp2f(y);
    return fmaf(e, ye * kLn2, e) * p;
}

// FP64 erfc for 0 <= x <= kErfcFitMax: Q(t), t = 0.8 x - 1, the degree-24
// Chebyshev interpolant of erfc on [0, 2.5] in the power basis of t
// (fit_erfc.py: within 4e-16 of erfc in absolute terms, which is what the
// Ewald sum of erfc(alpha r) / r over pairs needs).  One branch-free path for
// every lane with no exponential and no division, where the library erfc
// splits the range into intervals that the lanes of a wavefront take in
// turn.  Q(t) = E(t^2) + t O(t^2): the even and odd coefficients make two
// independent Horner chains of 12 and 11 steps instead of one of 24.
__device__ inline double erfc_fit64(double x) {
    // c0 .. c24
    constexpr double c[25] = {
        0.07709987174354177, -0.2956514030591187, 0.46195531727986267,
        -0.3272183497393615, 0.015037607984594071, 0.14398509643356755,
        -0.08125790773493699, -0.017290032190568103, 0.03396080776560146,
        -0.006538901006467077, -0.007390152521692959, 0.003771362281381614,
        0.0007674433199906389, -0.0010155399183195398, 6.853680201713117e-05,
        0.00018224272805793717, -4.804828007235314e-05, -2.2665916030465085e-05,
        1.1729907997669007e-05, 1.681147304115916e-06, -1.945561418732383e-06,
        -4.204118927684107e-09, 2.3109297654588739e-07, -1.0547604630012188e-08,
        -1.5789738524349255e-08};
    const double t = fma(x, 0.8, -1.0), u = t * t;
    double e = c[24], o = c[23];
#pragma unroll
    for (int k = 22; k >= 0; k -= 2) e = fma(e, u, c[k]);
#pragma unroll
    for (int k = 21; k >= 1; k -= 2) o = fma(o, u, c[k]);
    return fma(o, t, e);
}

template <typename Real>
__device__ inline Real erfc_r(Real x, int kind) {
    if (kind == 0) return erfc_as7126<Real>(x);
    if constexpr (sizeof(Real) == 4) {
        if (x >= 0.0f && x <= kErfcFitMax) return erfc_fit(x);
    } else {
        if (x >= 0.0 && x <= (double)kErfcFitMax) return erfc_fit64(x);
    }
    return erfc(x);
}

// Pade pair value u(r) = a r / (1 + b r)   (jastrow_pade.hpp:15-16).
template <typename Real>
__device__ inline Real pade_u(Real r, Real a, Real b) {
    return a * r / ((Real)1 + b * r);
}

// Real-space Ewald pair term erfc(alpha r) / r with the reference's r < 1e-12
// guard (energy_tracking.cpp:132,139) and an optional spherical cutoff.
template <typename Real>
__device__ inline Real ewald_pair(Real r, const DevSys &s) {
    if (s.use_cut && !(r < (Real)s.rcut)) return (Real)0;
    const Real inv_r = (r < (Real)1e-12) ? (Real)1 : (Real)1 / r;
    return erfc_r<Real>((Real)s.alpha * r, s.erfc_kind) * inv_r;
}

}  // namespace vmc
