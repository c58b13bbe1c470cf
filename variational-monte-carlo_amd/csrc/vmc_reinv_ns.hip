// vmc_reinv_ns.hip -- periodic FP64 re-inversion of FP32 inverses on the
// matrix cores, for determinants of up to 160 electrons (the multi-wavefront
// register kernels' 33 to 147 per spin, and the generic kernel's sizes).
//
// One 256-thread workgroup per due walker and determinant up to 80 electrons
// per spin (E in LDS; the walker's next sweep clears its move count), one per
// due walker above (accepted moves since its last re-inversion >= K); walkers
// not due leave at once.  Per determinant of order
// n, padded to P = 16 NT with an identity block:
//   E = D^T in FP64 (E[m][i] = phi_m(r_i), axis phasors of the FP64
//       positions), Y = the FP32 inverse invT (= E^-1 up to its drift);
//   repeat: Z = E Y on v_mfma_f64_16x16x4f64; the drift d = max|I - Z| over
//       the n x n block (recorded on the first pass); R = I - Z (in E's
//       place, E rebuilt if another step follows);
//       Y <- Y + Y R (FP64 products, stored back in FP32);
//   until d < 5e-4, after which the step just taken leaves d^2 < 2.5e-7, at
//   the rounding floor of the FP32 inverse.  A drift of 0.5 or more, a non-finite
//   one, or no convergence within kNSSteps falls back to Gauss-Jordan from
//   scratch, which also refreshes log|det D|.
// With an FP64 output (a mixed measured sweep's FP64 local energy) and a
// first-pass drift d with n d^3 <= 1e-8, one pass gives it instead:
// Y (I + R + R^2), residual R^3, with R^2 and the product by Y on the FP32
// matrix cores (both are corrections of order d; their FP32 rounding stays
// far below the residual).
// The identity padding keeps the padded block exact: Z, R and Y stay the
// identity there and zero across, so the n x n block evolves as if unpadded.
// The reference never re-inverts: it keeps its FP64 inverse by
// Sherman-Morrison updates alone (src/slater_plane_wave/slater_plane_wave.cpp:367-423).
#include "vmc_device.hpp"
#include "vmc_rr_launch.hpp"
#include "vmc_sweep_rr.hpp"

namespace vmc {
namespace {

typedef double ns_v4d __attribute__((ext_vector_type(4)));

constexpr int kNSThreads = 256;
constexpr int kNSWaves = kNSThreads / kWave;
constexpr int kNSSteps = 6;
// A step taken from drift d leaves about d^2 plus the rounding floor of an
// FP32 inverse (max|I - E Y| of a correctly rounded inverse is a few 1e-7):
// from d < 5e-4 one step reaches that floor.
constexpr double kNSDone = 5e-4;
// With an FP64 output (DevState::inv64: the FP64 local energy of a mixed
// measured sweep) the steps go on until a step starts from a drift d with
// n d^2 <= kRefineTarget (the register kernels' in-kernel refinement rule):
// its FP64 result Y + Y R has a residual of at most n d^2.
__device__ __forceinline__ bool ns_exact_from(int n, double dm) { return (double)n * dm * dm <= kRefineTarget; }
// One pass Y (I + R + R^2) has the residual R^3, at most n d^3: the FP64
// output in one pass when that is within the same target.
__device__ __forceinline__ bool ns_series_from(int n, double dm) {
    return (double)n * dm * dm * dm <= kRefineTarget;
}

typedef float ns_v4f __attribute__((ext_vector_type(4)));

// Output layouts of the 16 x 16 x 4 matrix instructions (lane l, register q):
// v_mfma_f64_16x16x4f64 holds C[4q + (l >> 4)][l & 15], v_mfma_f32_16x16x4f32
// holds C[4 (l >> 4) + q][l & 15]; both take A[l & 15][l >> 4] and
// B[l >> 4][l & 15].
__device__ __forceinline__ int ns_row32(int q, int lane) { return 4 * (lane >> 4) + q; }

// R = I - Z takes E's storage once Z = E Y is in registers (E is rebuilt
// for a further step, rarely needed), so a workgroup holds E and Y only:
// 50 KB at NT = 4, three workgroups per CU.
template <int NT>
struct NSShared {
    static constexpr int P = 16 * NT;   // padded order
    static constexpr int PE = P + 1;    // row pitch of E in doubles
    static constexpr int PF = P + 1;    // row pitch of Y and R in floats
    double E[P * PE];                   // also R (floats, pitch PF)
    float Y[P * PF];                    // also Gauss-Jordan's gjcol and perm
    double red[2 * kNSWaves];
};

// Output tiles of wave wv: with NT even, the wave's (NT/2) x (NT/2) block of
// the 2 x 2 block grid (each A and B fragment loaded once per k-step for NT/2
// tiles); with NT odd, tiles wv, wv + 4, ... in row-major order.
template <int NT>
struct NSTiles {
    static constexpr bool kBlock = NT % 2 == 0;
    static constexpr int kHalf = NT / 2;
    static constexpr int kCount = kBlock ? kHalf * kHalf : (NT * NT + kNSWaves - 1) / kNSWaves;
    __device__ static bool valid(int i, int wv) { return kBlock || wv + kNSWaves * i < NT * NT; }
    __device__ static int row(int i, int wv) {
        return kBlock ? (wv >> 1) * kHalf + i / kHalf : (wv + kNSWaves * i) / NT;
    }
    __device__ static int col(int i, int wv) {
        return kBlock ? (wv & 1) * kHalf + i % kHalf : (wv + kNSWaves * i) % NT;
    }
};

// acc[i] = (A B) on tile i of this wave; A and B are row-major LDS matrices of
// order P with pitches PA and PB.  Operand layout of v_mfma_f64_16x16x4f64:
// lane l supplies A[l & 15][l >> 4] and B[l >> 4][l & 15]; output register q
// of lane l is C[4q + (l >> 4)][l & 15].
template <int NT, typename TA, int PA, typename TB, int PB>
__device__ __forceinline__ void ns_mm(const TA *A, const TB *B, ns_v4d acc[NSTiles<NT>::kCount], int wv,
                                      int lane) {
    using T = NSTiles<NT>;
    const int lo = lane & 15, hi = lane >> 4;
#pragma unroll
    for (int i = 0; i < T::kCount; ++i) acc[i] = ns_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
    for (int kk = 0; kk < 4 * NT; ++kk) {
        const int k = 4 * kk + hi;
        if constexpr (T::kBlock) {
            double a[T::kHalf], b[T::kHalf];
#pragma unroll
            for (int u = 0; u < T::kHalf; ++u) {
                a[u] = (double)A[(16 * ((wv >> 1) * T::kHalf + u) + lo) * PA + k];
                b[u] = (double)B[k * PB + 16 * ((wv & 1) * T::kHalf + u) + lo];
            }
#pragma unroll
            for (int u = 0; u < T::kHalf; ++u)
#pragma unroll
                for (int v = 0; v < T::kHalf; ++v)
                    acc[u * T::kHalf + v] =
                        __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[v], acc[u * T::kHalf + v], 0, 0, 0);
        } else {
#pragma unroll
            for (int i = 0; i < T::kCount; ++i) {
                if (T::valid(i, wv)) {
                    const double a = (double)A[(16 * T::row(i, wv) + lo) * PA + k];
                    const double b = (double)B[k * PB + 16 * T::col(i, wv) + lo];
                    acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
                }
            }
        }
    }
}

// acc[i] = (A B) on tile i of this wave, FP32 operands and accumulation on
// v_mfma_f32_16x16x4f32 (A, B row-major LDS matrices, pitch PF).
template <int NT, int PF>
__device__ __forceinline__ void ns_mm32(const float *A, const float *B, ns_v4f acc[NSTiles<NT>::kCount], int wv,
                                        int lane) {
    using T = NSTiles<NT>;
    const int lo = lane & 15, hi = lane >> 4;
#pragma unroll
    for (int i = 0; i < T::kCount; ++i) acc[i] = ns_v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
    for (int kk = 0; kk < 4 * NT; ++kk) {
        const int k = 4 * kk + hi;
#pragma unroll
        for (int i = 0; i < T::kCount; ++i) {
            if (T::valid(i, wv)) {
                const float a = A[(16 * T::row(i, wv) + lo) * PF + k];
                const float b = B[k * PF + 16 * T::col(i, wv) + lo];
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
            }
        }
    }
}

// E = D^T of determinant d in FP64 (row pitch `pitch`): four threads per
// electron i, each writing every fourth orbital of column i from the axis
// phasors of the electron's FP64 position.  With `pad`, rows and columns n..P
// are set to the identity.
template <int NT>
__device__ void ns_build(const DevSys &s, const double *gp, int d, double *E, int pitch, bool pad) {
    constexpr int P = 16 * NT;
    const int n = s.n[d], first = s.first[d];
    if (pad)
        for (int e = threadIdx.x; e < P * P; e += kNSThreads) {
            const int i = e / P, j = e - (e / P) * P;
            if (i >= n || j >= n) E[i * pitch + j] = i == j ? 1.0 : 0.0;
        }
    for (int e = threadIdx.x; e < 4 * n; e += kNSThreads) {
        const int c = e / n, i = e - c * n;
        rr::Cx<double> a[3], a2[3], a3[3];
        a[0] = rr::axis_phasor<double>(gp[first + i], s.inv_L);
        a[1] = rr::axis_phasor<double>(gp[s.N + first + i], s.inv_L);
        a[2] = rr::axis_phasor<double>(gp[2 * s.N + first + i], s.inv_L);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            a2[q] = rr::cmul(a[q], a[q]);
            a3[q] = rr::cmul(a2[q], a[q]);
        }
        for (int m = c; m < n; m += 4) {
            const int4 o = s.orb[d * s.nmax + m];
            const rr::Cx<double> ph = rr::orbital_phasor3<double>(o.x, o.y, o.z, a, a2, a3);
            E[m * pitch + i] = o.w ? ph.im : ph.re;
        }
    }
}

template <int NT>
__device__ __forceinline__ void reinv_ns_body(const DevSys &s, const DevState &st, int w, int d_only) {
    using S = NSShared<NT>;
    using T = NSTiles<NT>;
    constexpr int P = S::P, PE = S::PE, PF = S::PF;
    const bool force = st.inv64 != nullptr;  // every walker, FP64 output
    if (!force && (s.reinv_k <= 0 || st.since[w] < s.reinv_k)) return;
    __shared__ S sh;
    float *R = reinterpret_cast<float *>(sh.E);
    const int t = threadIdx.x, lane = t & (kWave - 1), wv = t / kWave;
    const double *gp = st.pos + (size_t)w * 3 * s.N;
    float *inv = reinterpret_cast<float *>(st.inv) + (size_t)w * s.ninv;
    double drift = 0.0, ld[2] = {st.logdet[2 * w + 0], st.logdet[2 * w + 1]};
    bool singular = false;
    for (int d = d_only; d <= d_only; ++d) {
        const int n = s.n[d];
        float *Yg = inv + s.inv_off[d];
        double *Y64 = force ? st.inv64 + (size_t)w * s.ninv + s.inv_off[d] : nullptr;
        ns_build<NT>(s, gp, d, sh.E, PE, true);
        for (int e = t; e < P * P; e += kNSThreads) {
            const int i = e / P, j = e - (e / P) * P;
            sh.Y[i * PF + j] = (i < n && j < n) ? Yg[i * n + j] : (i == j ? 1.0f : 0.0f);
        }
        __syncthreads();
        double d0 = 0.0;
        bool ok = false, series = false;
#pragma unroll 1
        for (int it = 0; it < kNSSteps; ++it) {
            ns_v4d z[T::kCount];
            ns_mm<NT, double, PE, float, PF>(sh.E, sh.Y, z, wv, lane);
            // The lane's row and column offsets, opaque to loop-invariant code
            // motion (else every tile's addresses and bound tests are hoisted
            // out of the step loop and held in registers).
            int hl = lane >> 4, ll = lane & 15;
            asm volatile("" : "+v"(hl), "+v"(ll));
            __syncthreads();  // every wave is done reading E: R takes its place
            double dm = 0.0;
#pragma unroll
            for (int i = 0; i < T::kCount; ++i) {
                if (!T::valid(i, wv)) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = 16 * T::row(i, wv) + 4 * q + hl;
                    const int c = 16 * T::col(i, wv) + ll;
                    const double dv = (r == c ? 1.0 : 0.0) - z[i][q];
                    if (r < n && c < n) dm = (fabs(dv) > dm || dv != dv) ? fabs(dv) : dm;
                    R[r * PF + c] = (float)dv;
                }
            }
            dm = block_max(dm, sh.red);  // its barriers also publish R
            if (it == 0) d0 = dm;
            if (!(dm < 0.5)) break;
            if (dm < 1e-15) {
                if (Y64)
                    for (int e = t; e < n * n; e += kNSThreads) Y64[e] = (double)sh.Y[(e / n) * PF + e % n];
                ok = true;
                break;
            }
            if (Y64 && ns_series_from(n, dm)) {
                series = true;
                break;
            }
            // Y R on the FP32 matrix cores: a correction of order d, whose FP32
            // rounding (about 1e-7 of d) stays below what the step leaves.
            ns_v4f y[T::kCount];
            ns_mm32<NT, PF>(sh.Y, R, y, wv, lane);
            const bool exact = Y64 && ns_exact_from(n, dm);  // this step's FP64 result is the output
#pragma unroll
            for (int i = 0; i < T::kCount; ++i) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = 16 * T::row(i, wv) + ns_row32(q, lane);
                    const int c = 16 * T::col(i, wv) + ll;
                    const double v = T::valid(i, wv) ? (double)sh.Y[r * PF + c] + (double)y[i][q] : 0.0;
                    y[i][q] = (float)v;
                    if (exact && T::valid(i, wv) && r < n && c < n) Y64[r * n + c] = v;
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < T::kCount; ++i) {
                if (!T::valid(i, wv)) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    sh.Y[(16 * T::row(i, wv) + ns_row32(q, lane)) * PF + 16 * T::col(i, wv) + ll] = y[i][q];
            }
            if (Y64 ? exact : dm < kNSDone) {
                ok = true;
                __syncthreads();
                break;
            }
            ns_build<NT>(s, gp, d, sh.E, PE, true);  // R held E's place
            __syncthreads();
        }
        if (series) {
            // R <- R + R R, then Y (I + R + R^2) = Y + Y R', in FP32 on the
            // matrix cores, each lane updating its own output entries.
            const int ll = lane & 15;
            ns_v4f acc[T::kCount];
            ns_mm32<NT, PF>(R, R, acc, wv, lane);
            __syncthreads();
#pragma unroll
            for (int i = 0; i < T::kCount; ++i) {
                if (!T::valid(i, wv)) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    R[(16 * T::row(i, wv) + ns_row32(q, lane)) * PF + 16 * T::col(i, wv) + ll] += acc[i][q];
            }
            __syncthreads();
            ns_mm32<NT, PF>(sh.Y, R, acc, wv, lane);
#pragma unroll
            for (int i = 0; i < T::kCount; ++i) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = 16 * T::row(i, wv) + ns_row32(q, lane);
                    const int c = 16 * T::col(i, wv) + ll;
                    const double v = T::valid(i, wv) ? (double)sh.Y[r * PF + c] + (double)acc[i][q] : 0.0;
                    acc[i][q] = (float)v;
                    if (T::valid(i, wv) && r < n && c < n) Y64[r * n + c] = v;
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < T::kCount; ++i) {
                if (!T::valid(i, wv)) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    sh.Y[(16 * T::row(i, wv) + ns_row32(q, lane)) * PF + 16 * T::col(i, wv) + ll] = acc[i][q];
            }
            __syncthreads();
            ok = true;
        }
        drift = (d0 > drift || d0 != d0) ? d0 : drift;
        if (ok) {
            for (int e = t; e < n * n && !st.inv64_only; e += kNSThreads) {
                const int i = e / n, j = e - (e / n) * n;
                Yg[e] = sh.Y[i * PF + j];
            }
        } else {
            // Gauss-Jordan from scratch on E (row pitch n) in FP64.
            __syncthreads();
            ns_build<NT>(s, gp, d, sh.E, n, false);
            __syncthreads();
            double *gjcol = reinterpret_cast<double *>(sh.Y);
            int *perm = reinterpret_cast<int *>(gjcol + P);
            const double l = gj_invert(sh.E, n, gjcol, perm, sh.red);
            if (!(l > -INFINITY)) {
                singular = true;
                break;
            }
            ld[d] = l;
            for (int e = t; e < n * n && !st.inv64_only; e += kNSThreads) Yg[e] = (float)sh.E[e];
            if (Y64)
                for (int e = t; e < n * n; e += kNSThreads) Y64[e] = sh.E[e];
        }
        __syncthreads();
    }
    // One workgroup per determinant: the walker's next sweep clears its move
    // count (a forced re-inversion clears it here); the drift is an atomic
    // maximum (a non-negative double, or +inf / NaN, orders as its bit
    // pattern).
    if (t == 0 && !st.inv64_only) {
        if (force) st.since[w] = 0;
        if (d_only == 0) st.stats[(size_t)w * kStatCount + 15] += 1.0;
        if (singular) atomicOr(&st.flags[w], kFlagSingular);
        else st.logdet[2 * w + d_only] = ld[d_only];
        atomicMax(reinterpret_cast<unsigned long long *>(&st.drift[w]),
                  (unsigned long long)__double_as_longlong(drift));
    }
}

// Determinants of 97 to 160 electrons: the same Newton-Schulz steps with only
// Y (FP32) in LDS; E (FP64, row pitch n, the identity padding supplied on
// load) lives in the walker's Gauss-Jordan scratch and R = I - Z (FP32, row
// pitch P) in its residual buffer, both read through the caches.  The wave
// grid is 2 x 2, each wave an (NT/2) x (NT/2) block of output tiles with the
// next k-step's operands loaded before the current step's MFMAs.
template <int NT>
struct NSGShared {
    static constexpr int P = 16 * NT;
    static constexpr int PF = P + 1;
    float Y[P * PF];  // also Gauss-Jordan's gjcol and perm
    double red[2 * kNSWaves];
};

// The wave's output tiles r0 .. r0 + H - 1 by columns c0 + CB .. c0 + CB +
// NC - 1 (acc[u * NC + v]).
template <int NT, int CB, int NC, typename LA, typename LB>
__device__ __forceinline__ void nsg_mm(LA load_a, LB load_b, ns_v4d acc[(NT / 2) * NC], int wv, int lane) {
    constexpr int H = NT / 2;
    const int lo = lane & 15, hi = lane >> 4;
    const int r0 = (wv >> 1) * H, c0 = (wv & 1) * H + CB;
#pragma unroll
    for (int i = 0; i < H * NC; ++i) acc[i] = ns_v4d{0.0, 0.0, 0.0, 0.0};
    double a[H], b[NC], an[H], bn[NC];
#pragma unroll
    for (int u = 0; u < H; ++u) a[u] = load_a(16 * (r0 + u) + lo, hi);
#pragma unroll
    for (int v = 0; v < NC; ++v) b[v] = load_b(hi, 16 * (c0 + v) + lo);
#pragma unroll 1
    for (int kk = 0; kk < 4 * NT; ++kk) {
        const int kn = 4 * (kk + 1 < 4 * NT ? kk + 1 : kk) + hi;
#pragma unroll
        for (int u = 0; u < H; ++u) an[u] = load_a(16 * (r0 + u) + lo, kn);
#pragma unroll
        for (int v = 0; v < NC; ++v) bn[v] = load_b(kn, 16 * (c0 + v) + lo);
#pragma unroll
        for (int u = 0; u < H; ++u)
#pragma unroll
            for (int v = 0; v < NC; ++v)
                acc[u * NC + v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[v], acc[u * NC + v], 0, 0, 0);
#pragma unroll
        for (int u = 0; u < H; ++u) a[u] = an[u];
#pragma unroll
        for (int v = 0; v < NC; ++v) b[v] = bn[v];
    }
}

// Z = E Y over the wave's column tiles CB .. CB + NC - 1 (nsg_mm), then
// their part of R = I - Z (FP32, row pitch P) and of max|R| over the n x n
// block into dm.
template <int NT, int CB, int NC, typename LA, typename LB>
__device__ __forceinline__ void nsg_resid(LA load_e, LB load_y, float *R, int n, int wv, int lane, double &dm) {
    constexpr int H = NT / 2, P = 16 * NT;
    ns_v4d z[H * NC];
    nsg_mm<NT, CB, NC>(load_e, load_y, z, wv, lane);
    // This lane's first row and column, tied to the product's result:
    // otherwise every (tile, register) address and bound test below is
    // formed before the product's k loop and held in registers through it.
    int rl = 16 * ((wv >> 1) * H) + (lane >> 4), cl = 16 * ((wv & 1) * H + CB) + (lane & 15);
    const double z0 = z[0][0];
    asm volatile("" : "+v"(rl), "+v"(cl) : "v"(z0));
#pragma unroll
    for (int i = 0; i < H * NC; ++i) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = rl + 16 * (i / NC) + 4 * q;
            const int c = cl + 16 * (i % NC);
            const double dv = (r == c ? 1.0 : 0.0) - z[i][q];
            if (r < n && c < n) dm = (fabs(dv) > dm || dv != dv) ? fabs(dv) : dm;
            R[r * P + c] = (float)dv;
        }
    }
}

// nsg_mm with FP32 operands and accumulation on v_mfma_f32_16x16x4f32.
template <int NT, typename LA, typename LB>
__device__ __forceinline__ void nsg_mm32(LA load_a, LB load_b, ns_v4f acc[(NT / 2) * (NT / 2)], int wv,
                                         int lane) {
    constexpr int H = NT / 2;
    const int lo = lane & 15, hi = lane >> 4;
    const int r0 = (wv >> 1) * H, c0 = (wv & 1) * H;
#pragma unroll
    for (int i = 0; i < H * H; ++i) acc[i] = ns_v4f{0.0f, 0.0f, 0.0f, 0.0f};
    float a[H], b[H], an[H], bn[H];
#pragma unroll
    for (int u = 0; u < H; ++u) {
        a[u] = load_a(16 * (r0 + u) + lo, hi);
        b[u] = load_b(hi, 16 * (c0 + u) + lo);
    }
#pragma unroll 1
    for (int kk = 0; kk < 4 * NT; ++kk) {
        const int kn = 4 * (kk + 1 < 4 * NT ? kk + 1 : kk) + hi;
#pragma unroll
        for (int u = 0; u < H; ++u) {
            an[u] = load_a(16 * (r0 + u) + lo, kn);
            bn[u] = load_b(kn, 16 * (c0 + u) + lo);
        }
#pragma unroll
        for (int u = 0; u < H; ++u)
#pragma unroll
            for (int v = 0; v < H; ++v)
                acc[u * H + v] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[v], acc[u * H + v], 0, 0, 0);
#pragma unroll
        for (int u = 0; u < H; ++u) {
            a[u] = an[u];
            b[u] = bn[u];
        }
    }
}

template <int NT>
__device__ __forceinline__ void reinv_nsg_body(const DevSys &s, const DevState &st, int w) {
    using S = NSGShared<NT>;
    constexpr int P = S::P, PF = S::PF, H = NT / 2;
    // (an ensemble that re-inverts at these sizes always holds the residual buffer)
    const bool force = st.inv64 != nullptr;  // every walker, FP64 output
    if ((!force && (s.reinv_k <= 0 || st.since[w] < s.reinv_k)) || !st.resid) return;
    __shared__ S sh;
    const int t = threadIdx.x, lane = t & (kWave - 1), wv = t / kWave;
    const int r0 = (wv >> 1) * H, c0 = (wv & 1) * H;
    const double *gp = st.pos + (size_t)w * 3 * s.N;
    float *inv = reinterpret_cast<float *>(st.inv) + (size_t)w * s.ninv;
    double *E = st.scratch + (size_t)w * s.nmax * s.nmax;
    float *R = st.resid + (size_t)w * P * P;
    double drift = 0.0, ld[2] = {st.logdet[2 * w + 0], st.logdet[2 * w + 1]};
    bool singular = false;
    for (int d = 0; d < s.ndet; ++d) {
        const int n = s.n[d];
        float *Yg = inv + s.inv_off[d];
        double *Y64 = force ? st.inv64 + (size_t)w * s.ninv + s.inv_off[d] : nullptr;
        ns_build<NT>(s, gp, d, E, n, false);
        for (int e = t; e < P * P; e += kNSThreads) {
            const int i = e / P, j = e - (e / P) * P;
            sh.Y[i * PF + j] = (i < n && j < n) ? Yg[i * n + j] : (i == j ? 1.0f : 0.0f);
        }
        __syncthreads();
        // E with its identity padding; the load stays inside the n x n block.
        const auto load_e = [&](int r, int k) {
            const bool in = r < n && k < n;
            const double v = E[in ? r * n + k : 0];
            return in ? v : (r == k ? 1.0 : 0.0);
        };
        const auto load_y = [&](int r, int k) { return (double)sh.Y[r * PF + k]; };
        double d0 = 0.0;
        bool ok = false, series = false;
#pragma unroll 1
        for (int it = 0; it < kNSSteps; ++it) {
            // All H x H of the wave's tiles in one pass: 25 MFMAs per k step
            // at 147 per spin cover the next step's loads (two passes of 15
            // and 10 measured 9 % slower there).
            double dm = 0.0;
            nsg_resid<NT, 0, H>(load_e, load_y, R, n, wv, lane, dm);
            dm = block_max(dm, sh.red);  // its barriers also publish R
            if (it == 0) d0 = dm;
            if (!(dm < 0.5)) break;
            if (dm < 1e-15) {
                if (Y64)
                    for (int e = t; e < n * n; e += kNSThreads) Y64[e] = (double)sh.Y[(e / n) * PF + e % n];
                ok = true;
                break;
            }
            if (Y64 && ns_series_from(n, dm)) {
                series = true;
                break;
            }
            // Y R on the FP32 matrix cores: a correction of order d, whose FP32
            // rounding (about 1e-7 of d) stays below what the step leaves.
            const auto load_y32 = [&](int r, int k) { return sh.Y[r * PF + k]; };
            const auto load_r32 = [&](int r, int k) { return R[r * P + k]; };
            ns_v4f y[H * H];
            nsg_mm32<NT>(load_y32, load_r32, y, wv, lane);
            // first row of this lane's FP32-layout entries, and its column
            int rl32 = 16 * r0 + 4 * (lane >> 4), cl32 = 16 * c0 + (lane & 15);
            asm volatile("" : "+v"(rl32), "+v"(cl32));
            const bool exact = Y64 && ns_exact_from(n, dm);  // this step's FP64 result is the output
#pragma unroll
            for (int i = 0; i < H * H; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = rl32 + 16 * (i / H) + q, c = cl32 + 16 * (i % H);
                    const double v = (double)sh.Y[r * PF + c] + (double)y[i][q];
                    y[i][q] = (float)v;
                    if (exact && r < n && c < n) Y64[r * n + c] = v;
                }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < H * H; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) sh.Y[(rl32 + 16 * (i / H) + q) * PF + cl32 + 16 * (i % H)] = y[i][q];
            __syncthreads();
            if (Y64 ? exact : dm < kNSDone) {
                ok = true;
                break;
            }
        }
        if (series) {
            // R <- R + R R, then Y (I + R + R^2) = Y + Y R', in FP32 on the
            // matrix cores, each lane updating its own output entries.
            const auto load_y32 = [&](int r, int k) { return sh.Y[r * PF + k]; };
            const auto load_r32 = [&](int r, int k) { return R[r * P + k]; };
            int rl = 16 * r0, cl = 16 * c0 + (lane & 15);
            asm volatile("" : "+v"(rl), "+v"(cl));
            ns_v4f acc[H * H];
            nsg_mm32<NT>(load_r32, load_r32, acc, wv, lane);
            __syncthreads();
#pragma unroll
            for (int i = 0; i < H * H; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    R[(rl + 16 * (i / H) + ns_row32(q, lane)) * P + cl + 16 * (i % H)] += acc[i][q];
            __syncthreads();
            nsg_mm32<NT>(load_y32, load_r32, acc, wv, lane);
#pragma unroll
            for (int i = 0; i < H * H; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = rl + 16 * (i / H) + ns_row32(q, lane), c = cl + 16 * (i % H);
                    const double v = (double)sh.Y[r * PF + c] + (double)acc[i][q];
                    acc[i][q] = (float)v;
                    if (r < n && c < n) Y64[r * n + c] = v;
                }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < H * H; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    sh.Y[(rl + 16 * (i / H) + ns_row32(q, lane)) * PF + cl + 16 * (i % H)] = acc[i][q];
            __syncthreads();
            ok = true;
        }
        drift = (d0 > drift || d0 != d0) ? d0 : drift;
        if (ok) {
            for (int e = t; e < n * n && !st.inv64_only; e += kNSThreads) {
                const int i = e / n, j = e - (e / n) * n;
                Yg[e] = sh.Y[i * PF + j];
            }
        } else {
            // Gauss-Jordan from scratch on E (row pitch n) in FP64.
            __syncthreads();
            ns_build<NT>(s, gp, d, E, n, false);
            __syncthreads();
            double *gjcol = reinterpret_cast<double *>(sh.Y);
            int *perm = reinterpret_cast<int *>(gjcol + P);
            const double l = gj_invert(E, n, gjcol, perm, sh.red);
            if (!(l > -INFINITY)) {
                singular = true;
                break;
            }
            ld[d] = l;
            for (int e = t; e < n * n && !st.inv64_only; e += kNSThreads) Yg[e] = (float)E[e];
            if (Y64)
                for (int e = t; e < n * n; e += kNSThreads) Y64[e] = E[e];
        }
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
}

// Two workgroups per CU up to 96 per spin (37 KB of LDS each at NT = 6).
template <int NT>
__global__ __launch_bounds__(kNSThreads, NT <= 6 ? 2 : 1) void k_reinv_nsg(DevSys s, DevState st) {
    reinv_nsg_body<NT>(s, st, blockIdx.x);
}

template <int NT>
__global__ __launch_bounds__(kNSThreads, NT <= 6 ? 2 : 1) void k_reinv_nsg_b(const BatchItem *__restrict__ items,
                                                                           int n_items) {
    const BatchItem &it = items[batch_find(items, n_items, blockIdx.x)];
    reinv_nsg_body<NT>(it.s, it.st, blockIdx.x - it.walker_begin);
}

template <int NT>
void launch_nsg(const DevSys &d, const DevState &st, const BatchItem *items, int n_items, unsigned grid,
                hipStream_t stream) {
    if (items)
        hipLaunchKernelGGL((k_reinv_nsg_b<NT>), dim3(grid), dim3(kNSThreads), 0, stream, items, n_items);
    else
        hipLaunchKernelGGL((k_reinv_nsg<NT>), dim3(grid), dim3(kNSThreads), 0, stream, d, st);
}

// Workgroups per CU the registers must allow (the LDS allows as many):
// four up to 32 per spin, three up to 64, two up to 80.
template <int NT>
constexpr int ns_blocks() {
    return NT <= 2 ? 4 : NT <= 4 ? 3 : NT == 5 ? 2 : 1;
}

// Workgroup g re-inverts determinant g % ndet of walker g / ndet.
template <int NT>
__global__ __launch_bounds__(kNSThreads, ns_blocks<NT>()) void k_reinv_ns(DevSys s, DevState st) {
    reinv_ns_body<NT>(s, st, blockIdx.x / s.ndet, blockIdx.x % s.ndet);
}

// Batch form over a group of systems of one size (and determinant count).
template <int NT>
__global__ __launch_bounds__(kNSThreads, ns_blocks<NT>()) void k_reinv_ns_b(const BatchItem *__restrict__ items,
                                                                          int n_items) {
    const int ndet = items[0].s.ndet;
    const int g = blockIdx.x / ndet;
    const BatchItem &it = items[batch_find(items, n_items, g)];
    reinv_ns_body<NT>(it.s, it.st, g - it.walker_begin, blockIdx.x % ndet);
}

template <int NT>
void launch_ns(const DevSys &d, const DevState &st, const BatchItem *items, int n_items, unsigned grid,
               hipStream_t stream) {
    const unsigned g = grid * (unsigned)d.ndet;
    if (items)
        hipLaunchKernelGGL((k_reinv_ns_b<NT>), dim3(g), dim3(kNSThreads), 0, stream, items, n_items);
    else
        hipLaunchKernelGGL((k_reinv_ns<NT>), dim3(g), dim3(kNSThreads), 0, stream, d, st);
}

void launch_any(const DevSys &d, const DevState &st, const BatchItem *items, int n_items, unsigned grid,
                hipStream_t stream) {
    if (d.nmax > 128) return launch_nsg<10>(d, st, items, n_items, grid, stream);
    if (d.nmax > 96) return launch_nsg<8>(d, st, items, n_items, grid, stream);
    if (d.nmax > 80) return launch_nsg<6>(d, st, items, n_items, grid, stream);
    switch ((d.nmax + 15) / 16) {
        case 1: launch_ns<1>(d, st, items, n_items, grid, stream); break;
        case 2: launch_ns<2>(d, st, items, n_items, grid, stream); break;
        case 3: launch_ns<3>(d, st, items, n_items, grid, stream); break;
        case 4: launch_ns<4>(d, st, items, n_items, grid, stream); break;
        case 5: launch_ns<5>(d, st, items, n_items, grid, stream); break;
        default: launch_ns<6>(d, st, items, n_items, grid, stream); break;
    }
}

}  // namespace

bool reinv_ns_supported(const DevSys &d, int real_bytes) {
    return real_bytes == 4 && d.nmax <= kNSMaxOrder;
}

void reinv_ns_launch(const DevSys &d, const DevState &st, unsigned grid, hipStream_t stream) {
    launch_any(d, st, nullptr, 0, grid, stream);
}

void reinv_ns_batch_launch(const DevSys &shape, const BatchItem *items, int n_items, unsigned grid,
                           hipStream_t stream) {
    launch_any(shape, DevState{}, items, n_items, grid, stream);
}

}  // namespace vmc
