// vmc_sweep_rr.hpp -- register-resident sweep for systems with at most 32
// electrons per spin (N = 14, 38, 54 unpolarised; polarised up to 32).
//
// One wavefront runs one walker.  Lanes 0..31 own the rows of the spin-up
// transposed inverse, lanes 32..63 those of spin down: lane (h, r) keeps
// invT_h[r][0..n) in VGPRs together with electron r of spin h (position,
// Jastrow row sum U, the sweep's proposal and its axis phasors).  Per move:
//   - the proposal of electron p and its axis phasors e^{2 pi i u'} are read
//     from lane p (v_readlane), every orbital of the row is a product of axis
//     phasors (no transcendental on the move's critical path);
//   - the orbital row is broadcast through 32 LDS words and every row lane
//     forms s_k = sum_m phi_m invT[k][m]; the ratio is s_p (v_readlane);
//   - the Jastrow new-pair sum is a DPP row reduction plus four readlanes, so
//     the accept decision is a scalar branch;
//   - on accept, row p is broadcast through LDS and each row lane applies the
//     Sherman-Morrison update to its own registers.
// The per-sweep local energy uses the same registers plus an LDS mirror of the
// positions.
#pragma once

#include "../../include/vmc_hip.h"
#include "vmc_device.hpp"
#include "vmc_walker.hpp"

#include <array>
#include <cstddef>

namespace vmc {
namespace rr {

struct Orb {
    int x, y, z, t;
};

// Closed-shell filling at compile time: canonical n sorted by
// (|n|^2, nx, ny, nz), orbital 0 = cos(0), then (cos, sin) per nonzero k
// (slater_plane_wave.cpp:30-152; same rule as fill_shells in vmc_tables.hpp).
template <int NO>
__host__ __device__ constexpr std::array<Orb, (NO > 0 ? NO : 1)> shell_orbitals() {
    constexpr int R = 3;
    std::array<std::array<int, 4>, (2 * R + 1) * (2 * R + 1) * (2 * R + 1)> c{};
    int nc = 0;
    for (int x = -R; x <= R; ++x)
        for (int y = -R; y <= R; ++y)
            for (int z = -R; z <= R; ++z) {
                const bool canon = x != 0 ? x > 0 : (y != 0 ? y > 0 : (z != 0 ? z > 0 : true));
                if (canon) c[nc++] = {x * x + y * y + z * z, x, y, z};
            }
    for (int i = 1; i < nc; ++i) {
        auto key = c[i];
        int j = i - 1;
        while (j >= 0 && (c[j][0] > key[0] || (c[j][0] == key[0] && (c[j][1] > key[1] || (c[j][1] == key[1] && (c[j][2] > key[2] || (c[j][2] == key[2] && c[j][3] > key[3]))))))) {
            c[j + 1] = c[j];
            --j;
        }
        c[j + 1] = key;
    }
    std::array<Orb, (NO > 0 ? NO : 1)> out{};
    int o = 0;
    for (int i = 0; i < nc && o < NO; ++i) {
        out[o++] = Orb{c[i][1], c[i][2], c[i][3], 0};
        if (c[i][0] != 0 && o < NO) out[o++] = Orb{c[i][1], c[i][2], c[i][3], 1};
    }
    return out;
}

template <int NO>
__host__ __device__ constexpr int max_component() {
    constexpr auto t = shell_orbitals<NO>();
    int m = 0;
    for (int i = 0; i < NO; ++i) {
        const int a = t[i].x < 0 ? -t[i].x : t[i].x;
        const int b = t[i].y < 0 ? -t[i].y : t[i].y;
        const int c = t[i].z < 0 ? -t[i].z : t[i].z;
        m = a > m ? a : m;
        m = b > m ? b : m;
        m = c > m ? c : m;
    }
    return m;
}

// ------------------------------------------------------- lane primitives
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Sum over each row of 16 lanes, left in every lane of the row:
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror.
template <typename T>
__device__ __forceinline__ T row_sum(T v) {
    v += dpp<0xB1>(v);
    v += dpp<0x4E>(v);
    v += dpp<0x141>(v);
    v += dpp<0x140>(v);
    return v;
}

__device__ __forceinline__ float readlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ double readlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Whole-wave sum as a wave-uniform (scalar) value; all 64 lanes must be active.
template <typename T>
__device__ __forceinline__ T wave_total(T v) {
    v = row_sum(v);
    return (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
}

template <typename Real>
struct Cx {
    Real re, im;
};

template <typename Real>
__device__ __forceinline__ Cx<Real> cmul(Cx<Real> a, Cx<Real> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// a^n for |n| <= 2 given a and a^2 (conjugate for negative n).
template <typename Real>
__device__ __forceinline__ Cx<Real> cpow2(int n, Cx<Real> a1, Cx<Real> a2) {
    const int m = n < 0 ? -n : n;
    Cx<Real> r;
    r.re = m == 0 ? (Real)1 : (m == 1 ? a1.re : a2.re);
    r.im = m == 0 ? (Real)0 : (m == 1 ? a1.im : a2.im);
    r.im = n < 0 ? -r.im : r.im;
    return r;
}

// a^n for |n| <= 3 given a, a^2 and a^3 (conjugate for negative n).
template <typename Real>
__device__ __forceinline__ Cx<Real> cpow3(int n, Cx<Real> a1, Cx<Real> a2, Cx<Real> a3) {
    const int m = n < 0 ? -n : n;
    Cx<Real> r;
    r.re = m == 0 ? (Real)1 : (m == 1 ? a1.re : (m == 2 ? a2.re : a3.re));
    r.im = m == 0 ? (Real)0 : (m == 1 ? a1.im : (m == 2 ? a2.im : a3.im));
    r.im = n < 0 ? -r.im : r.im;
    return r;
}

// Axis phasor e^{2 pi i u} of a coordinate in [0, L).
template <typename Real>
__device__ __forceinline__ Cx<Real> axis_phasor(double x, double inv_L) {
    double t = x * inv_L;
    t -= rint(t);
    Cx<Real> p;
    sincos_2pi<Real>(t, &p.im, &p.re);
    return p;
}

// Compile-time power a^n (n >= 0).
template <int NPOW, typename Real>
__device__ __forceinline__ Cx<Real> cpow_ct(Cx<Real> a) {
    if constexpr (NPOW == 0) {
        return Cx<Real>{(Real)1, (Real)0};
    } else if constexpr (NPOW == 1) {
        return a;
    } else {
        return cmul(a, cpow_ct<NPOW - 1, Real>(a));
    }
}

template <typename Real>
__device__ __forceinline__ Cx<Real> cconj(Cx<Real> a) {
    return {a.re, -a.im};
}

// Orbital phasor e^{2 pi i n.u} from the three axis phasors and their squares.
template <typename Real>
__device__ __forceinline__ Cx<Real> orbital_phasor(int nx, int ny, int nz, const Cx<Real> a[3],
                                                   const Cx<Real> a2[3]) {
    return cmul(cmul(cpow2(nx, a[0], a2[0]), cpow2(ny, a[1], a2[1])), cpow2(nz, a[2], a2[2]));
}

// Orbital phasor for components up to 3 (a3[c] = a[c]^3).
template <typename Real>
__device__ __forceinline__ Cx<Real> orbital_phasor3(int nx, int ny, int nz, const Cx<Real> a[3],
                                                    const Cx<Real> a2[3], const Cx<Real> a3[3]) {
    return cmul(cmul(cpow3(nx, a[0], a2[0], a3[0]), cpow3(ny, a[1], a2[1], a3[1])),
                cpow3(nz, a[2], a2[2], a3[2]));
}

template <typename Real>
__device__ __forceinline__ Real jastrow_pair(Real r, Real a, Real b) {
    return a * r / ((Real)1 + b * r);
}

}  // namespace rr

// Maximum G vectors per lane of the register-resident measurement.
constexpr int kRRMaxGPerLane = 12;
// Electrons whose axis-power tables are staged in LDS at once.
constexpr int kRRChunk = 8;
// Highest power of an axis phasor the structure factor needs (|m| <= 7).
constexpr int kRRMaxPow = 7;

template <typename Real, int NU, int ND>
struct RRShared {
    double pos[3][kWave];          // positions by lane (pair loops of the measurement)
    alignas(16) double prop[4][kWave];  // this sweep's proposal x, y, z and accept uniform
    Real pph[6][kWave];            // axis phasors of the proposal (re, im per axis)
    Real tab[kRRChunk * 3 * (kRRMaxPow + 1) * 2];  // with prop and pph: staging (below)
    double stat[kStatCount];
    double misc[4];                // sigma, log|det| up, log|det| down
    Real phi[2][32];               // orbital rows of the pending moves (by spin)
    Real row[32];                  // inverse row of the moved electron
    double gjcol[32];
    int perm[32];
    // The run prop | pph | tab as one staging area (inverse blocks, the
    // structure-factor tables of the measurement).
    static constexpr size_t kStageBytes =
        sizeof(double) * 4 * kWave + sizeof(Real) * 6 * kWave + sizeof(Real) * kRRChunk * 3 * (kRRMaxPow + 1) * 2;
    __device__ void *stage() { return &prop[0][0]; }
};

// Structure factor S(G) = sum_j e^{i G.r_j} (energy_tracking.cpp:72-91) by
// threads that own G vectors (thread t owns G t + NT k).  Electrons go
// through in chunks of kRRChunk: 3 kRRChunk threads stage the chunk's axis
// phasor powers X^0..X^7, Y^-7..Y^7 and Z^-7..Z^7 (one axis of one electron
// each) in LDS as [entry][electron] (complex, padded pitch), and every thread
// adds X^mx Y^my Z^mz of its G vectors: three conflict-free LDS reads whose
// electron index is the reads' immediate offset, one complex product and four
// fused accumulations.
namespace sf {

constexpr int kEntries = (kRRMaxPow + 1) + 2 * (2 * kRRMaxPow + 1);
constexpr int kY0 = kRRMaxPow + 1 + kRRMaxPow;      // entry of Y^0
constexpr int kZ0 = kY0 + 2 * kRRMaxPow + 1;        // entry of Z^0
// Complex values per entry: one per electron of the chunk and a pad that
// spreads the entries of one FP32 read over distinct LDS banks.
constexpr int kPitch = kRRChunk + 1;

// A table value, aligned so that it moves with one LDS read.
template <typename Real>
struct alignas(2 * sizeof(Real)) Pair {
    Real re, im;
};

template <typename Real>
constexpr int table_reals() {
    return 2 * kEntries * kPitch;
}

// Powers of one axis phasor (coordinate u) of the electron in slot i of the
// chunk (entries centre + m and, for Y and Z, centre - m), or zeros for an
// empty slot.
template <typename Real>
__device__ __forceinline__ void stage_axis(Pair<Real> *tab, int i, int ax, bool present, double u,
                                           double inv_L) {
    const int centre = ax == 0 ? 0 : (ax == 1 ? kY0 : kZ0);
    rr::Cx<Real> a{(Real)0, (Real)0}, p{(Real)0, (Real)0};
    if (present) {
        a = rr::axis_phasor<Real>(u, inv_L);
        p = {(Real)1, (Real)0};
    }
    Pair<Real> *t = tab + i;
#pragma unroll
    for (int m = 0; m <= kRRMaxPow; ++m) {
        t[(centre + m) * kPitch] = {p.re, p.im};
        if (ax != 0) t[(centre - m) * kPitch] = {p.re, -p.im};
        p = rr::cmul(p, a);
    }
}

// term_k += X Y Z of the staged electron at pointer offset i of the entries.
template <typename Real, int GPT>
__device__ __forceinline__ void add_one(const Pair<Real> *tab, int i, int k, const int (&ox)[GPT],
                                        const int (&oy)[GPT], const int (&oz)[GPT], Real (&re)[GPT],
                                        Real (&im)[GPT]) {
    const Pair<Real> X = tab[ox[k] + i], Y = tab[oy[k] + i], Z = tab[oz[k] + i];
    const Real xr = X.re * Y.re - X.im * Y.im, xi = X.re * Y.im + X.im * Y.re;
    re[k] = fma(xr, Z.re, re[k]);
    re[k] = fma(-xi, Z.im, re[k]);
    im[k] = fma(xr, Z.im, im[k]);
    im[k] = fma(xi, Z.re, im[k]);
}

// Adds the kRRChunk staged electrons to the sums of the thread's G vectors
// (ox, oy, oz: table offsets of their entries), two electrons at a time (the
// reads in flight stay within the register budget).  Up to 5 G vectors per
// thread the electron index is the reads' immediate offset; above, the
// electron pairs are a loop, taken once per group of 5 G vectors.
template <typename Real, int GPT>
__device__ __forceinline__ void add(const Pair<Real> *tab, const int (&ox)[GPT], const int (&oy)[GPT],
                                    const int (&oz)[GPT], Real (&re)[GPT], Real (&im)[GPT]) {
    constexpr int kG = 5;
    if constexpr (GPT <= kG) {
#pragma unroll
        for (int i = 0; i < kRRChunk; ++i) {
#pragma unroll
            for (int k = 0; k < GPT; ++k) add_one(tab, i, k, ox, oy, oz, re, im);
            if (i % 2 == 1) {
                asm volatile("" ::: "memory");
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    } else {
#pragma unroll
        for (int k0 = 0; k0 < GPT; k0 += kG) {
#pragma unroll 1
            for (int i = 0; i < kRRChunk; i += 2) {
#pragma unroll
                for (int k = k0; k < (k0 + kG < GPT ? k0 + kG : GPT); ++k) {
                    add_one(tab, i, k, ox, oy, oz, re, im);
                    add_one(tab, i + 1, k, ox, oy, oz, re, im);
                }
            }
        }
    }
}

// sum over the thread's G vectors of w_G |S_G|^2 (to be reduced over the
// threads).  pos_of(ax, e): coordinate ax of electron e (LDS).  tab: LDS of
// table_reals<Real>(); every thread of the workgroup takes part.
template <typename Real, int N, int GPT, int NT, typename PosOf>
__device__ double energy_g(const DevSys &s, Pair<Real> *tab, int tid, PosOf pos_of) {
    constexpr int P = kPitch;
    int ox[GPT], oy[GPT], oz[GPT];
    Real re[GPT], im[GPT];
    // The G table reads stay here (not hoisted to the kernel's start, where
    // their values would be live through the rest of the measurement).
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int k = 0; k < GPT; ++k) {
        const int g = tid + NT * k;
        // Absent G vectors read the entries of G = 0 and get no weight.
        const int4 G = g < s.ng ? s.gvec[g] : make_int4(0, 0, 0, 0);
        ox[k] = G.x * P;
        oy[k] = (kY0 + G.y) * P;
        oz[k] = (kZ0 + G.z) * P;
        re[k] = (Real)0;
        im[k] = (Real)0;
    }
    // The last chunk's empty slots hold zeros, so every chunk is a full one.
    constexpr int kChunks = (N + kRRChunk - 1) / kRRChunk;
#pragma unroll 1
    for (int c = 0; c < kChunks; ++c) {
        __syncthreads();
        // Thread 3 i + ax stages axis ax of slot i.
        if (tid < 3 * kRRChunk) {
            const int i = tid / 3, ax = tid - 3 * i, e = c * kRRChunk + i;
            stage_axis(tab, i, ax, e < N, e < N ? pos_of(ax, e) : 0.0, s.inv_L);
        }
        __syncthreads();
        add<Real, GPT>(tab, ox, oy, oz, re, im);
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < GPT; ++k) {
        const int g = tid + NT * k;
        if (g < s.ng) acc += s.gw[g] * ((double)re[k] * (double)re[k] + (double)im[k] * (double)im[k]);
    }
    __syncthreads();
    return acc;
}

// sum_G w_G |S_G|^2 by column segments (DevSys::sfseg; thread t takes
// segments t, t + NT, ...): per staged electron two LDS reads give
// X^mx Y^my, and each step a = a0, a0 + 1, ... reads Z^a, adding
// X^mx Y^my Z^a to S(mx, my, a) and its conjugate-in-z partner
// X^mx Y^my Z^-a to S(mx, my, -a): the two share their four real products,
// so four fused accumulations per step serve both.  The thread's share is
// reduced by the caller.
template <typename Real, int N, int NT, typename PosOf>
__device__ double energy_cols(const DevSys &s, Pair<Real> *tab, int tid, PosOf pos_of) {
    constexpr int P = kPitch;
    constexpr int S = kSfSegSteps;
    constexpr int kChunks = (N + kRRChunk - 1) / kRRChunk;
    double acc = 0.0;
#pragma unroll 1
    for (int p0 = 0; p0 < s.nseg; p0 += NT) {
        int sg = p0 + tid;
        asm volatile("" : "+v"(sg));  // the segment reads stay in this pass
        const bool act = sg < s.nseg;
        const int4 G = act ? s.sfseg[sg] : make_int4(0, 0, 0, 0);
        const int ox = G.x * P, oy = (kY0 + G.y) * P;
        // Entry of Z^(a0 + k) per step; steps past the segment's end (zero
        // weight) read the highest staged power instead.
        int oz[S];
#pragma unroll
        for (int k = 0; k < S; ++k) oz[k] = (kZ0 + (G.z + k < kRRMaxPow ? G.z + k : kRRMaxPow)) * P;
        // Per step the four product sums of X^mx Y^my = (xr, xi) with
        // Z^a = (zr, zi): sum xr zr, sum xi zi, sum xr zi, sum xi zr.
        // S(mx, my, +a) = (rr - ii, ri + ir) and S(mx, my, -a) =
        // (rr + ii, ir - ri) follow once at the end.
        Real rr[S], ii[S], ri[S], ir[S];
#pragma unroll
        for (int k = 0; k < S; ++k) rr[k] = ii[k] = ri[k] = ir[k] = (Real)0;
#pragma unroll 1
        for (int c = 0; c < kChunks; ++c) {
            __syncthreads();
            if (tid < 3 * kRRChunk) {
                const int i = tid / 3, ax = tid - 3 * i, e = c * kRRChunk + i;
                stage_axis(tab, i, ax, e < N, e < N ? pos_of(ax, e) : 0.0, s.inv_L);
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kRRChunk; ++i) {
                const Pair<Real> X = tab[ox + i], Y = tab[oy + i];
                const Real xr = X.re * Y.re - X.im * Y.im, xi = X.re * Y.im + X.im * Y.re;
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const Pair<Real> Z = tab[oz[k] + i];
                    rr[k] = fma(xr, Z.re, rr[k]);
                    ii[k] = fma(xi, Z.im, ii[k]);
                    ri[k] = fma(xr, Z.im, ri[k]);
                    ir[k] = fma(xi, Z.re, ir[k]);
                }
                if (i % 2 == 1) __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (act) {
            const double *w = s.sfw + (size_t)sg * 2 * S;
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const double pr = (double)rr[k] - (double)ii[k], pi = (double)ri[k] + (double)ir[k];
                const double mr = (double)rr[k] + (double)ii[k], mi = (double)ir[k] - (double)ri[k];
                acc += w[2 * k] * (pr * pr + pi * pi) + w[2 * k + 1] * (mr * mr + mi * mi);
            }
        }
    }
    __syncthreads();
    return acc;
}

// As energy_g with GPT slots per thread, and with fewer when the G set of the
// Ewald sum with alpha = 5/L (at most 5 per wavefront lane) fits them.
template <typename Real, int N, int GPT, int NT, typename PosOf>
__device__ double energy(const DevSys &s, Pair<Real> *tab, int tid, PosOf pos_of) {
    constexpr int kSmall = (5 * kWave + NT - 1) / NT;
    if constexpr (kSmall < GPT) {
        if (s.ng <= kSmall * NT) return energy_g<Real, N, kSmall, NT>(s, tab, tid, pos_of);
    }
    return energy_g<Real, N, GPT, NT>(s, tab, tid, pos_of);
}

}  // namespace sf

// The LDS run prop | pph | tab of RRShared doubles as the stage through which
// inverse blocks move between HBM and the lanes' registers, and holds the
// structure-factor tables of the measurement.
template <typename Real, int NU, int ND>
__device__ __forceinline__ Real *rr_stage(RRShared<Real, NU, ND> &sh) {
    using Sh = RRShared<Real, NU, ND>;
    static_assert(offsetof(Sh, pph) == offsetof(Sh, prop) + sizeof(sh.prop) &&
                      offsetof(Sh, tab) == offsetof(Sh, pph) + sizeof(sh.pph),
                  "stage run must be contiguous");
    constexpr int NMAX = NU > ND ? NU : ND;
    static_assert(sizeof(sh.prop) + sizeof(sh.pph) + sizeof(sh.tab) >= sizeof(Real) * NMAX * NMAX,
                  "stage too small for one inverse block");
    return reinterpret_cast<Real *>(&sh.prop[0][0]);
}

// Inverse blocks of both spins between HBM ([NU x NU][ND x ND], row k =
// invT[k][.]) and the lanes' registers (lane (h, r) holds row r of spin h):
// each block moves with coalesced loads or stores of consecutive elements and
// every lane reads or writes its own row in LDS (odd row stride, no bank
// conflicts).
// kOnly >= 0 moves spin kOnly's block alone.
template <typename Real, int NU, int ND, int NMAX, int kOnly = -1>
__device__ __forceinline__ void rr_inverse_in(const Real *blk, Real *stage, Real (&inv)[NMAX],
                                              int lane, int h, int r) {
#pragma unroll
    for (int m = 0; m < NMAX; ++m) inv[m] = (Real)0;
#pragma unroll
    for (int hp = 0; hp < (ND > 0 ? 2 : 1); ++hp) {
        if (kOnly >= 0 && hp != kOnly) continue;
        constexpr int kU = NU;
        const int np = hp ? ND : kU;
        const Real *src = blk + (hp ? kU * kU : 0);
        for (int e = lane; e < np * np; e += kWave) stage[e] = src[e];
        __syncthreads();
        if (h == hp && r < np) {
#pragma unroll
            for (int m = 0; m < NMAX; ++m)
                if (m < np) inv[m] = stage[r * np + m];
        }
        __syncthreads();
    }
}

template <typename Real, int NU, int ND, int NMAX, int kOnly = -1>
__device__ __forceinline__ void rr_inverse_out(Real *blk, Real *stage, const Real (&inv)[NMAX],
                                               int lane, int h, int r) {
#pragma unroll
    for (int hp = 0; hp < (ND > 0 ? 2 : 1); ++hp) {
        if (kOnly >= 0 && hp != kOnly) continue;
        constexpr int kU = NU;
        const int np = hp ? ND : kU;
        __syncthreads();
        if (h == hp && r < np) {
#pragma unroll
            for (int m = 0; m < NMAX; ++m)
                if (m < np) stage[r * np + m] = inv[m];
        }
        __syncthreads();
        Real *dst = blk + (hp ? kU * kU : 0);
        for (int e = lane; e < np * np; e += kWave) dst[e] = stage[e];
    }
    __syncthreads();
}

// Loop-invariant system scalars, hoisted once per kernel into registers.
template <typename Real>
struct RRConst {
    double L, half_L, inv_L;
    Real a_same, a_anti, b, alpha, rcut;
    int use_cut, erfc_kind;
    __device__ explicit RRConst(const DevSys &s)
        : L(s.L), half_L(s.half_L), inv_L(s.inv_L), a_same((Real)s.a_same),
          a_anti((Real)s.a_anti), b((Real)s.b), alpha((Real)s.alpha), rcut((Real)s.rcut),
          use_cut(s.use_cut), erfc_kind(s.erfc_kind) {}
    __device__ double mi(double d) const {
        return d + L * (double)(d <= -half_L) - L * (double)(d > half_L);
    }
    // Minimum-image separation a - b in the working precision: the
    // difference less the nearest multiple of L (three VALU operations, no
    // compare-select hazards), in FP32 from the FP64 difference rounded once.
    // Against the reference's masked form (jastrow_pade.cpp:18-24) it
    // differs only where |a - b| rounds to L/2, where both are minimum images
    // whose lengths agree to an ulp.
    __device__ Real mi_r(double a, double b) const {
        if constexpr (sizeof(Real) == 8) {
            const double d = a - b;
            return fma(-L, rint(d * inv_L), d);
        } else {
            const Real d = (Real)(a - b);
            return __builtin_fmaf(-(Real)L, __builtin_rintf(d * (Real)inv_L), d);
        }
    }
    // erfc(alpha r) / r with the r < 1e-12 guard and optional cutoff
    // (energy_tracking.cpp:101-139).
    __device__ Real ewald(Real r) const {
        if (use_cut && !(r < rcut)) return (Real)0;
        const Real inv_r = (r < (Real)1e-12) ? (Real)1 : wrcp<Real>(r);
        return erfc_r<Real>(alpha * r, erfc_kind) * inv_r;
    }
    // The same with inv_r from pair_geom.
    __device__ Real ewald_ri(Real r, Real inv_r) const {
        if (use_cut && !(r < rcut)) return (Real)0;
        return erfc_r<Real>(alpha * r, erfc_kind) * inv_r;
    }
    // Pair distance r from its square, 1/r (1 below r = 1e-12, the real-space
    // Ewald guard; the Jastrow derivatives skip those pairs) and the Pade
    // denominator 1 / (1 + b r).  FP64 takes both reciprocals from rsqrt_nr
    // and rcp_nr; a pair closer than 1e-12 gets r = 0, which changes
    // erfc(alpha r) by under 2 alpha 1e-12.
    __device__ void pair_geom(Real r2, Real &r, Real &inv_r, Real &den) const {
        if constexpr (sizeof(Real) == 8) {
            const bool tiny = !(r2 >= 1e-24);
            const double y = rsqrt_nr(tiny ? 1.0 : r2);
            r = tiny ? 0.0 : r2 * y;
            inv_r = tiny ? 1.0 : y;
            den = rcp_nr(1.0 + b * r);
        } else {
            r = wsqrt<Real>(r2);
            inv_r = (r < (Real)1e-12) ? (Real)1 : wrcp<Real>(r);
            den = wrcp<Real>((Real)1 + b * r);
        }
    }
};

// In-LDS FP64 Gauss-Jordan inversion of the n x n matrix A (A = D^T on entry,
// invT on exit) by the whole wave; returns log|det| or -inf below the
// reference's pivot tolerance 1e-12 (matrix.hpp:35-39).
__device__ inline double rr_gauss_jordan(double *A, double *col, int *perm, int n, int lane) {
    double logdet = 0.0;
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
        if (!(best >= 1e-12)) return -INFINITY;
        if (lane == 0) perm[c] = bi;
        if (bi != c)
            for (int j = lane; j < n; j += kWave) {
                const double t = A[c * n + j];
                A[c * n + j] = A[bi * n + j];
                A[bi * n + j] = t;
            }
        __syncthreads();
        const double piv = A[c * n + c];
        logdet += log(fabs(piv));
        const double pinv = 1.0 / piv;
        for (int r = lane; r < n; r += kWave) col[r] = A[r * n + c];
        __syncthreads();
        for (int j = lane; j < n; j += kWave) A[c * n + j] = (j == c ? 1.0 : A[c * n + j]) * pinv;
        __syncthreads();
        for (int e = lane; e < n * n; e += kWave) {
            const int r = e / n, j = e - (e / n) * n;
            if (r == c) continue;
            A[e] = (j == c ? 0.0 : A[e]) - col[r] * A[c * n + j];
        }
        __syncthreads();
    }
    for (int c = n - 1; c >= 0; --c) {
        const int q = perm[c];
        if (q != c)
            for (int r = lane; r < n; r += kWave) {
                const double t = A[r * n + c];
                A[r * n + c] = A[r * n + q];
                A[r * n + q] = t;
            }
        __syncthreads();
    }
    return logdet;
}

// Per-lane walker state of the register-resident kernel.
template <typename Real, int NMAX>
struct RRLane {
    Real inv[NMAX];   // row r of invT of spin h
    double x, y, z;   // electron position
    double U;         // Jastrow row sum
    int orb;          // orbital r: (nx+4) | (ny+4) << 4 | (nz+4) << 8 | type << 12
};

__device__ __forceinline__ int orb_nx(int o) { return (o & 15) - 4; }
__device__ __forceinline__ int orb_ny(int o) { return ((o >> 4) & 15) - 4; }
__device__ __forceinline__ int orb_nz(int o) { return ((o >> 8) & 15) - 4; }
__device__ __forceinline__ int orb_t(int o) { return o >> 12; }

__device__ __attribute__((noinline)) double rr_exp_f64(double x);

template <typename Real, int NU, int ND, int GPL>
struct RRKernel {
    static constexpr int N = NU + ND;
    static constexpr int NMAX = NU > ND ? NU : ND;
    static constexpr auto orb = rr::shell_orbitals<NMAX>();
    static constexpr int kMaxPow = rr::max_component<NMAX>();
    using Sh = RRShared<Real, NU, ND>;
    static_assert(NU <= 32 && ND <= 32, "register-resident path holds at most 32 rows per spin");
    static_assert(ND == 0 || ND == NU, "register-resident path: polarised or equal spins");

    __device__ static int lane_of(int e) { return e < NU ? e : 32 + (e - NU); }

    // Squares of the axis phasors when some orbital needs |n_c| = 2.
    __device__ static void squares(const rr::Cx<Real> a[3], rr::Cx<Real> a2[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a2[c] = kMaxPow >= 2 ? rr::cmul(a[c], a[c]) : a[c];
    }

    // Jastrow row sum, gradient and Laplacian of the lane's electron and its
    // half of the real-space Ewald sum (jastrow_pade.cpp:57-138,
    // energy_tracking.cpp:93-144), reading the other positions from LDS.
    // Pair terms of the local energy with every unordered pair evaluated
    // once: in step k (1 <= k <= N/2) electron e pairs with (e + k) mod N and
    // keeps its own side of the Jastrow gradient and Laplacian; the partner's
    // side (the gradient negated, the same Laplacian term) reaches it through
    // a cross-lane permute from the lane of (e - k) mod N.  For even N the
    // last step's pairs are met from both ends: each keeps its own side, and
    // their energy counts half.  Terms are summed in the working precision
    // within a chunk of kPairChunk steps, chunk sums in FP64.
    template <typename ShT>
    __device__ static void pair_terms(const RRConst<Real> &c, const ShT &sh,
                                      const RRLane<Real, NMAX> &L, int lane, bool has_e, int h,
                                      double g[3], double &lap, double &vreal) {
        g[0] = g[1] = g[2] = 0.0;
        lap = 0.0;
        vreal = 0.0;
        const Real m2b = (Real)-2 * c.b;
        constexpr int K = N / 2;
        constexpr bool kEven = N % 2 == 0;
        constexpr int kPairChunk = sizeof(Real) == 4 ? 8 : 2;
        const int e = h ? NU + (lane & 31) : (lane & 31);
        for (int k0 = 1; k0 <= K; k0 += kPairChunk) {
            Real gx_c = 0, gy_c = 0, gz_c = 0, l_c = 0, v_c = 0;
#pragma unroll
            for (int kk = 0; kk < kPairChunk; ++kk) {
                const int k = k0 + kk;
                if (k > K) break;
                const bool last_even = kEven && k == K;
                int j = e + k;
                j = j >= N ? j - N : j;
                int src = e - k;
                src = src < 0 ? src + N : src;
                const int lj = lane_of(j);
                Real pgx = 0, pgy = 0, pgz = 0, pl = 0;
                if (has_e) {
                    const double xj = sh.pos[0][lj], yj = sh.pos[1][lj], zj = sh.pos[2][lj];
                    const Real dx = c.mi_r(L.x, xj);
                    const Real dy = c.mi_r(L.y, yj);
                    const Real dz = c.mi_r(L.z, zj);
                    Real r, inv_r, den;
                    c.pair_geom(dx * dx + dy * dy + dz * dz, r, inv_r, den);
                    const Real a = ((lj >> 5) == h) ? c.a_same : c.a_anti;
                    const Real v = c.ewald_ri(r, inv_r);
                    v_c += last_even ? (Real)0.5 * v : v;
                    if (r >= (Real)1e-12) {
                        const Real d1 = a * den * den;
                        const Real gf = d1 * inv_r;
                        const Real l = d1 * (m2b * den + (Real)2 * inv_r);
                        gx_c += gf * dx;
                        gy_c += gf * dy;
                        gz_c += gf * dz;
                        l_c += l;
                        if (!last_even) {
                            pgx = -gf * dx;
                            pgy = -gf * dy;
                            pgz = -gf * dz;
                            pl = l;
                        }
                    }
                }
                // The partner's side, from the lane that paired with this
                // lane's electron in this step (all lanes take part).
                const int sl = (has_e ? lane_of(src) : lane) << 2;
                const Real rgx = permute(sl, pgx), rgy = permute(sl, pgy);
                const Real rgz = permute(sl, pgz), rl = permute(sl, pl);
                if (has_e && !last_even) {
                    gx_c += rgx;
                    gy_c += rgy;
                    gz_c += rgz;
                    l_c += rl;
                }
            }
            vreal += (double)v_c;
            g[0] += (double)gx_c;
            g[1] += (double)gy_c;
            g[2] += (double)gz_c;
            lap += (double)l_c;
        }
    }

    // Value of `v` on the lane whose byte address is `addr` (ds_bpermute).
    __device__ static Real permute(int addr, Real v) {
        if constexpr (sizeof(Real) == 4) {
            return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
        } else {
            const long long b = __double_as_longlong(v);
            const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffffLL));
            const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
            return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
        }
    }

    // Jastrow row sum U_i of the lane's electron (jastrow_pade.cpp:7-55 rows).
    template <typename ShT>
    __device__ static double jastrow_row(const RRConst<Real> &c, const ShT &sh,
                                         const RRLane<Real, NMAX> &L, int lane, bool has_e, int h) {
        double U = 0.0;
        for (int j = 0; j < N; ++j) {
            const int lj = lane_of(j);
            const double xj = sh.pos[0][lj], yj = sh.pos[1][lj], zj = sh.pos[2][lj];
            if (!has_e || lj == lane) continue;
            const Real dx = (Real)c.mi(L.x - xj);
            const Real dy = (Real)c.mi(L.y - yj);
            const Real dz = (Real)c.mi(L.z - zj);
            const Real r = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
            const Real a = ((lj >> 5) == h) ? c.a_same : c.a_anti;
            U += (double)(a * r * wrcp<Real>((Real)1 + c.b * r));
        }
        return U;
    }

    // Local energy after a sweep; also refreshes U (drift-free) and returns
    // kinetic / potential parts.
    template <typename ShT>
    __device__ static Measured measure(const DevSys &s, const RRConst<Real> &cst, ShT &sh,
                                       RRLane<Real, NMAX> &L, int lane, bool has_e, int h) {
        double g[3], lap, vreal;
        pair_terms(cst, sh, L, lane, has_e, h, g, lap, vreal);
        // Slater gradient and Laplacian of ln|D_h| at the lane's electron:
        // sum_m invT[r][m] grad phi_m(r) with phi_m from the axis phasors
        // (slater_plane_wave.cpp:425-499).
        double sx = 0.0, sy = 0.0, sz = 0.0, sl = 0.0;
        if (has_e) {
            rr::Cx<Real> cur[3], a2[3];
            cur[0] = rr::axis_phasor<Real>(L.x, s.inv_L);
            cur[1] = rr::axis_phasor<Real>(L.y, s.inv_L);
            cur[2] = rr::axis_phasor<Real>(L.z, s.inv_L);
            squares(cur, a2);
            const Real ku = (Real)s.kunit;
            constexpr int nh = NU;  // ND is 0 or NU
#pragma unroll
            for (int m = 0; m < NMAX; ++m) {
                if (m >= nh) break;
                if (orb[m].t == 1) continue;  // handled with its cos partner below
                __builtin_amdgcn_sched_barrier(0);
                const rr::Cx<Real> ph = rr::orbital_phasor<Real>(orb[m].x, orb[m].y, orb[m].z, cur, a2);
                const Real k2 = ku * ku * (Real)(orb[m].x * orb[m].x + orb[m].y * orb[m].y + orb[m].z * orb[m].z);
                // cos orbital m: grad = -k sin, lap = -k^2 cos
                const Real wc = L.inv[m];
                Real gsc = -wc * ph.im;
                Real lpc = -wc * ph.re;
                if (m + 1 < nh && orb[m + 1].t == 1) {
                    // sin orbital m+1 (same k): grad = k cos, lap = -k^2 sin
                    const Real ws = L.inv[m + 1];
                    gsc += ws * ph.re;
                    lpc -= ws * ph.im;
                }
                sx += (double)(gsc * ku * (Real)orb[m].x);
                sy += (double)(gsc * ku * (Real)orb[m].y);
                sz += (double)(gsc * ku * (Real)orb[m].z);
                sl += (double)(lpc * k2);
            }
        }
        const double slap = sl - (sx * sx + sy * sy + sz * sz);
        const double tx = sx + g[0], ty = sy + g[1], tz = sz + g[2];
        const double kin_lane = has_e ? (slap + lap + tx * tx + ty * ty + tz * tz) : 0.0;
        Measured m;
        m.kinetic = -0.5 * rr::wave_total(kin_lane);
        m.vreal = rr::wave_total(has_e ? vreal : 0.0);
        // Both are complete here: their inputs (the inverse rows among them)
        // are not carried through the structure-factor loop.
        asm volatile("" : "+v"(m.kinetic), "+v"(m.vreal));
        // (an uncontracted product: callers that add it to other terms round
        // it the same way in every kernel they are inlined into)
        m.vrecip = __dmul_rn(s.pref, recip(s, sh, L, lane, has_e));
        m.e_local = m.kinetic + m.vreal + m.vrecip + s.e_const;
        return m;
    }

    // sum_G w_G |S_G|^2 with S_G rebuilt from per-electron axis-power tables
    // staged in LDS (energy_tracking.cpp:72-91,146-191).
    template <typename ShT>
    __device__ static double recip(const DevSys &s, ShT &sh, const RRLane<Real, NMAX> &L, int lane,
                                   bool has_e) {
        static_assert(ShT::kStageBytes >= sizeof(Real) * sf::table_reals<Real>(),
                      "structure-factor tables exceed the stage run");
        auto *tab = reinterpret_cast<sf::Pair<Real> *>(sh.stage());
        // Electron e sits in lane e (up) or 32 + e - NU (down).
        const auto pos_of = [&sh](int ax, int e) { return sh.pos[ax][e < NU ? e : 32 + e - NU]; };
        return rr::wave_total(sf::energy_cols<Real, N, kWave>(s, tab, lane, pos_of));
    }

    // Metropolis test u < R^2 exp(2 dJ) (simulation.cpp:125-131).  The
    // hardware exp2 (relative error < 1e-6 here) decides unless u lies within
    // 1e-5 of the threshold, where the exact FP64 exp decides, so the outcome
    // equals the FP64 comparison.  Inputs are wave-uniform; so is the result.
    __device__ static int accept_test(double R, double dj, double ua) {
        const double r2 = R * R;
        const float e32 = __builtin_amdgcn_exp2f((float)(2.0 * dj * 1.4426950408889634));
        const double approx = r2 * (double)e32;
        int acc;
        if (!(approx == approx)) {
            acc = 0;
        } else if (fabs(ua - approx) > 1e-5 * approx || !(fabs(dj) < 40.0)) {
            acc = ua < approx;
        } else {
            acc = ua < r2 * rr_exp_f64(2.0 * dj);
        }
        return wave_uniform(acc);
    }

    // Pade pair value of the lane's electron with the point (xq, yq, zq).
    __device__ static Real pair_with(const RRConst<Real> &c, const RRLane<Real, NMAX> &L, double xq, double yq,
                                     double zq, Real a) {
        const Real dx = c.mi_r(xq, L.x);
        const Real dy = c.mi_r(yq, L.y);
        const Real dz = c.mi_r(zq, L.z);
        const Real rn = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
        return a * rn * wrcp<Real>((Real)1 + c.b * rn);
    }

    // Sherman-Morrison for an accepted move of electron (hm, i) with ratio R
    // (row s_k of every lane of spin hm), then the Jastrow row sums: every
    // other lane adds its new pair term un minus its old one, recomputed from
    // the mover's old position.  Lane lm takes the proposal (xn, yn, zn) and
    // the row sum jn.
    __device__ static void apply_move(const RRConst<Real> &c, Sh &sh, RRLane<Real, NMAX> &L, int lane,
                                      bool has_e, int h, int r, int hm, int i, double R, Real sk, Real un,
                                      double jn, double xn, double yn, double zn) {
        constexpr int NP = NU;
        const int lm = hm * 32 + i;
        if (lane == lm) {
#pragma unroll
            for (int m = 0; m < NP; ++m) sh.row[m] = L.inv[m];
        }
        __syncthreads();
        {
            const Real inv_ratio = wrcp<Real>((Real)R);
            Real f = (r == i) ? (Real)1 - inv_ratio : sk * inv_ratio;
            f = (h == hm && r < NP) ? f : (Real)0;
#pragma unroll
            for (int m = 0; m < NP; ++m) L.inv[m] -= sh.row[m] * f;
        }
        const double xo = rr::readlane(L.x, lm), yo = rr::readlane(L.y, lm), zo = rr::readlane(L.z, lm);
        if (has_e && lane != lm) {
            const Real ao = (h == hm) ? c.a_same : c.a_anti;
            L.U += (double)un - (double)pair_with(c, L, xo, yo, zo, ao);
        }
        if (lane == lm) {
            L.U = jn;
            L.x = xn;
            L.y = yn;
            L.z = zn;
            sh.pos[0][lane] = xn;
            sh.pos[1][lane] = yn;
            sh.pos[2][lane] = zn;
        }
    }

    // The moves of up electron i and then down electron i (the sweep order
    // of equal spins), evaluated together: the two determinants are
    // independent, so the wavefront's two halves form the orbital rows and
    // ratios of both proposals at once, and every lane forms its pair terms
    // with both.  The down move's Jastrow change uses the up electron where
    // the up move left it: its pair term with the down proposal is formed for
    // both outcomes.  Same decisions and updates as two calls of move.
    // Returns the number of accepted moves.
    __device__ static int move_pair(const RRConst<Real> &c, Sh &sh, RRLane<Real, NMAX> &L, int lane,
                                    bool has_e, int h, int r, int i, LogDetTrack &ldt, double *rec) {
        static_assert(ND == NU, "paired moves need equal spins");
        constexpr int NP = NU;
        const int lu = i, ld = 32 + i;
        const double xu = sh.prop[0][lu], yu = sh.prop[1][lu], zu = sh.prop[2][lu];
        const double xd = sh.prop[0][ld], yd = sh.prop[1][ld], zd = sh.prop[2][ld];
        // Orbital rows: lane (h, j) evaluates orbital j at its spin's proposal.
        {
            const int lp = h ? ld : lu;
            rr::Cx<Real> a[3], a2[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a[k].re = sh.pph[2 * k][lp];
                a[k].im = sh.pph[2 * k + 1][lp];
            }
            squares(a, a2);
            if (r < NP) {
                const rr::Cx<Real> ph = rr::orbital_phasor<Real>(orb_nx(L.orb), orb_ny(L.orb), orb_nz(L.orb), a, a2);
                sh.phi[h][r] = orb_t(L.orb) ? ph.im : ph.re;
            }
        }
        // Jastrow: new pair terms against both proposals; the down proposal's
        // pair with the up proposal (taken when the up move is accepted).
        Real un_u = (Real)0, un_d = (Real)0;
        if (has_e && lane != lu) un_u = pair_with(c, L, xu, yu, zu, h == 0 ? c.a_same : c.a_anti);
        if (has_e && lane != ld) un_d = pair_with(c, L, xd, yd, zd, h == 1 ? c.a_same : c.a_anti);
        Real ud_new;
        {
            const Real dx = c.mi_r(xd, xu), dy = c.mi_r(yd, yu), dz = c.mi_r(zd, zu);
            const Real rn = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
            ud_new = c.a_anti * rn * wrcp<Real>((Real)1 + c.b * rn);
        }
        __syncthreads();
        // s_k = sum_m phi_m invT[k][m] with the lane's spin's orbital row.
        Real sk0 = (Real)0, sk1 = (Real)0;
#pragma unroll
        for (int m = 0; m < NP; m += 2) {
            sk0 += sh.phi[h][m] * L.inv[m];
            if (m + 1 < NP) sk1 += sh.phi[h][m + 1] * L.inv[m + 1];
        }
        const Real sk = sk0 + sk1;
        const double Ru = (double)rr::readlane(sk, lu), Rd = (double)rr::readlane(sk, ld);
        const double jn_u = (double)rr::wave_total(un_u);
        const double jn_d0 = (double)rr::wave_total(un_d);    // up electron at its old position
        const Real ud_old = rr::readlane(un_d, lu);
        // Up move.
        const double dj_u = jn_u - rr::readlane(L.U, lu);
        const int acc_u = accept_test(Ru, dj_u, sh.prop[3][lu]);
        if (acc_u) {
            apply_move(c, sh, L, lane, has_e, h, r, 0, i, Ru, sk, un_u, jn_u, xu, yu, zu);
            if (lane == lu) un_d = ud_new;
            ldt.add(0, Ru);
        }
        // Down move, after the up move's outcome.
        const double jn_d = acc_u ? jn_d0 + ((double)ud_new - (double)ud_old) : jn_d0;
        const double dj_d = jn_d - rr::readlane(L.U, ld);
        const int acc_d = accept_test(Rd, dj_d, sh.prop[3][ld]);
        if (acc_d) {
            __syncthreads();  // the up move's reads of the broadcast row are done
            apply_move(c, sh, L, lane, has_e, h, r, 1, i, Rd, sk, un_d, jn_d, xd, yd, zd);
            ldt.add(1, Rd);
        }
        if (rec && lane == 0) {
            double *ru = rec + kTraceCount * i, *rd = rec + kTraceCount * (NU + i);
            ru[0] = Ru;
            ru[1] = dj_u;
            ru[2] = (double)acc_u;
            rd[0] = Rd;
            rd[1] = dj_d;
            rd[2] = (double)acc_d;
        }
        return acc_u + acc_d;
    }

    // One move of electron (HP, rp).  Returns 1 if accepted.
    template <int HP>
    __device__ static int move(const RRConst<Real> &c, Sh &sh, RRLane<Real, NMAX> &L, int lane,
                               bool has_e, int h, int r, int rp, LogDetTrack &ldt, double *rec) {
        constexpr int NP = HP ? ND : NU;
        const int lp = HP * 32 + rp;
        const double xn = sh.prop[0][lp], yn = sh.prop[1][lp], zn = sh.prop[2][lp];
        rr::Cx<Real> a[3], a2[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a[k].re = sh.pph[2 * k][lp];
            a[k].im = sh.pph[2 * k + 1][lp];
        }
        squares(a, a2);
        // Orbital row of the proposal: lane (HP, j) evaluates orbital j.
        if (h == HP && r < NP) {
            const rr::Cx<Real> ph = rr::orbital_phasor<Real>(orb_nx(L.orb), orb_ny(L.orb), orb_nz(L.orb), a, a2);
            sh.phi[HP][r] = orb_t(L.orb) ? ph.im : ph.re;
        }
        // Jastrow: new pair terms against every other electron.
        const Real apair = (h == HP) ? c.a_same : c.a_anti;
        Real un = (Real)0;
        if (has_e && lane != lp) {
            const Real dx = c.mi_r(xn, L.x);
            const Real dy = c.mi_r(yn, L.y);
            const Real dz = c.mi_r(zn, L.z);
            const Real rn = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
            un = apair * rn * wrcp<Real>((Real)1 + c.b * rn);
        }
        __syncthreads();
        // s_k = sum_m phi_m invT[k][m] for every row k of spin HP (computed
        // on every lane without branching; only spin HP's rows are used).
        Real sk0 = (Real)0, sk1 = (Real)0;
#pragma unroll
        for (int m = 0; m < NP; m += 2) {
            sk0 += sh.phi[HP][m] * L.inv[m];
            if (m + 1 < NP) sk1 += sh.phi[HP][m + 1] * L.inv[m + 1];
        }
        const Real sk = sk0 + sk1;
        const double R = (double)rr::readlane(sk, lp);
        // New Jastrow row sum of the mover: summed in the working precision
        // (its FP32 rounding only moves the accept threshold by ~1e-6).
        const double jn = (double)rr::wave_total(un);
        const double dj = jn - rr::readlane(L.U, lp);
        const double ua = sh.prop[3][lp];
        // Metropolis test u < R^2 exp(2 dJ) (simulation.cpp:125-131).  The
        // hardware exp2 (relative error < 1e-6 here) decides unless u lies
        // within 1e-5 of the threshold, where the exact FP64 exp decides, so
        // the outcome equals the FP64 comparison.
        const double r2 = R * R;
        const float e32 = __builtin_amdgcn_exp2f((float)(2.0 * dj * 1.4426950408889634));
        const double approx = r2 * (double)e32;
        int acc;
        if (!(approx == approx)) {
            acc = 0;
        } else if (fabs(ua - approx) > 1e-5 * approx || !(fabs(dj) < 40.0)) {
            acc = ua < approx;
        } else {
            acc = ua < r2 * rr_exp_f64(2.0 * dj);
        }
        // Every input above is wave-uniform; say so, so the branches are scalar.
        const int accepted = wave_uniform(acc);
        if (rec && lane == 0) {
            double *q = rec + kTraceCount * (HP ? NU + rp : rp);
            q[0] = R;
            q[1] = dj;
            q[2] = (double)accepted;
        }
        // Sherman-Morrison (simulation.cpp:133-145, slater_plane_wave.cpp:
        // 367-423) as one FMA per element and lane: rows k != p of spin HP
        // lose (s_k / R) row_p; row p itself (row_p is its own registers)
        // loses (1 - 1/R) of itself, i.e. becomes row_p / R; every other lane
        // subtracts 0 x (a finite row).  A rejected move skips the update on
        // a wave-uniform branch.
        if (!accepted) return 0;
        if (lane == lp) {
#pragma unroll
            for (int m = 0; m < NP; ++m) sh.row[m] = L.inv[m];
        }
        __syncthreads();
        {
            const Real inv_ratio = wrcp<Real>((Real)R);
            Real f = (r == rp) ? (Real)1 - inv_ratio : sk * inv_ratio;
            f = (accepted && h == HP && r < NP) ? f : (Real)0;
#pragma unroll
            for (int m = 0; m < NP; ++m) L.inv[m] -= sh.row[m] * f;
        }
        if (!accepted) return 0;
        ldt.add(HP, R);
        const double xo = rr::readlane(L.x, lp), yo = rr::readlane(L.y, lp), zo = rr::readlane(L.z, lp);
        if (has_e && lane != lp) {
            const Real dx = c.mi_r(xo, L.x);
            const Real dy = c.mi_r(yo, L.y);
            const Real dz = c.mi_r(zo, L.z);
            const Real ro = wsqrt<Real>(dx * dx + dy * dy + dz * dz);
            L.U += (double)un - (double)(apair * ro * wrcp<Real>((Real)1 + c.b * ro));
        }
        if (lane == lp) {
            L.U = jn;
            L.x = xn;
            L.y = yn;
            L.z = zn;
            sh.pos[0][lane] = xn;
            sh.pos[1][lane] = yn;
            sh.pos[2][lane] = zn;
        }
        return 1;
    }
};

// The proposal q (in [0, L)), its axis phasors and the accept uniform ua
// into the lane's LDS slots.
template <typename Real>
__device__ __forceinline__ void rr_stage_proposal(const DevSys &s, const double (&q)[3], double ua, int lane,
                                                  double *prop, Real *pph) {
    for (int k = 0; k < 3; ++k) {
        prop[k * kWave + lane] = q[k];
        const rr::Cx<Real> ph = rr::axis_phasor<Real>(q[k], s.inv_L);
        pph[(2 * k) * kWave + lane] = ph.re;
        pph[(2 * k + 1) * kWave + lane] = ph.im;
    }
    prop[3 * kWave + lane] = ua;
}

// Philox draws (tags MOVE and ACCP, counter {electron, walker, sweep}),
// Box-Muller or uniform-cube displacement, wrap into [0, L) and the axis
// phasors of the proposal, written to the lane's LDS slots.
template <typename Real>
__device__ __forceinline__ void rr_proposal(const DevSys &s, double sigma, double x0,
                                                      double y0, double z0, int e, int w,
                                                      uint32_t sweep, int lane, double *prop,
                                                      Real *pph) {
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
    rr_stage_proposal<Real>(s, q, u01(ra.x), lane, prop, pph);
}

__device__ __attribute__((noinline)) double rr_exp_f64(double x) { return exp(x); }

// kGiven: the moves come from the caller (SweepGiven) instead of Philox and
// each move's ratio, Jastrow change and decision are recorded.
template <typename Real, int NU, int ND, int GPL, bool kGiven = false>
__device__ __forceinline__ void sweep_rr_body(const DevSys &s, const DevState &st, int w,
                                              unsigned long long sweep0, int flags,
                                              SweepGiven given = SweepGiven{}) {
    using K = RRKernel<Real, NU, ND, GPL>;
    constexpr int NMAX = K::NMAX;
    constexpr int N = NU + ND;
    __shared__ RRShared<Real, NU, ND> sh;
    const RRConst<Real> cst(s);
    const int lane = threadIdx.x;
    const int h = lane >> 5, r = lane & 31;
    const int nh = h ? ND : NU;
    const bool has_e = r < nh;
    const int e = h ? NU + r : r;

    RRLane<Real, NMAX> L;
    const Real *inv_hbm = reinterpret_cast<const Real *>(st.inv) + (size_t)w * s.ninv;
    const double *gp = st.pos + (size_t)w * 3 * N;
    rr_inverse_in<Real, NU, ND, NMAX>(inv_hbm, rr_stage(sh), L.inv, lane, h, r);
    L.x = has_e ? gp[e] : 0.0;
    L.y = has_e ? gp[N + e] : 0.0;
    L.z = has_e ? gp[2 * N + e] : 0.0;
    sh.pos[0][lane] = L.x;
    sh.pos[1][lane] = L.y;
    sh.pos[2][lane] = L.z;
    if (lane < kStatCount) sh.stat[lane] = st.stats[(size_t)w * kStatCount + lane];
    if (lane == 0) {
        sh.misc[0] = st.sigma[w];
        sh.misc[1] = st.logdet[2 * w + 0];
        sh.misc[2] = st.logdet[2 * w + 1];
    }
    L.orb = 0;
#pragma unroll
    for (int m = 0; m < NMAX; ++m)
        if (r == m)
            L.orb = (K::orb[m].x + 4) | ((K::orb[m].y + 4) << 4) | ((K::orb[m].z + 4) << 8) | (K::orb[m].t << 12);
    __syncthreads();
    int since = st.since[w];
    if (s.reinv_k > 0 && since >= s.reinv_k) since = 0;  // re-inverted after the last sweep
    int wflags = st.flags[w];
    const uint32_t sweep = (uint32_t)sweep0;
    // Jastrow row sums: kept from the previous sweep unless stale or due for
    // the periodic recomputation (the decision is walker-uniform).
    if (!(wflags & kFlagJrowValid) || (sweep % kJrowRefresh) == 0)
        L.U = K::jastrow_row(cst, sh, L, lane, has_e, h);
    else
        L.U = has_e ? st.jrow[(size_t)w * N + e] : 0.0;
    // Proposal of the lane's electron for this sweep (electron order is fixed,
    // so its position is current until its own move); staged in LDS for the
    // move that reads it.  One sweep per launch keeps this straight-line code
    // out of the move loop, so its transcendental constants die before it.
    double *rec = nullptr;
    if constexpr (kGiven) {
        if (has_e) {
            const size_t q = (size_t)w * N + e;
            const double p3[3] = {given.prop[3 * q], given.prop[3 * q + 1], given.prop[3 * q + 2]};
            rr_stage_proposal<Real>(s, p3, given.u[q], lane, &sh.prop[0][0], &sh.pph[0][0]);
        }
        rec = given.trace + (size_t)w * N * kTraceCount;
    } else {
        if (has_e) rr_proposal<Real>(s, sh.misc[0], L.x, L.y, L.z, e, st.w0 + w, sweep, lane, &sh.prop[0][0],
                                      &sh.pph[0][0]);
    }
    __syncthreads();
    // Sweep order (sweep_visit): with two determinants (ND == NU) the spins
    // alternate, up p then down p.
    int acc = 0;
    LogDetTrack ldt;
    if constexpr (ND > 0) {
        static_assert(ND == NU, "two determinants of equal size");
#pragma unroll 1
        for (int p = 0; p < NU; ++p) acc += K::move_pair(cst, sh, L, lane, has_e, h, r, p, ldt, rec);
    } else {
#pragma unroll 1
        for (int p = 0; p < NU; ++p) acc += K::template move<0>(cst, sh, L, lane, has_e, h, r, p, ldt, rec);
    }
    const int n_acc = acc;
    since += acc;
    if (flags & VMC_SWEEP_ADAPT) {
        const double f = exp(0.25 * ((double)acc / (double)N - 0.5));
        if (lane == 0) {
            const double sigma = sh.misc[0] * f;
            sh.misc[0] = sigma > 0.5 * s.L ? 0.5 * s.L : sigma;
        }
    }
    Real *inv_out = reinterpret_cast<Real *>(st.inv) + (size_t)w * s.ninv;
    rr_inverse_out<Real, NU, ND, NMAX>(inv_out, rr_stage(sh), L.inv, lane, h, r);
    if (has_e) {
        double *gq = st.pos + (size_t)w * 3 * N;
        gq[e] = L.x;
        gq[N + e] = L.y;
        gq[2 * N + e] = L.z;
        st.jrow[(size_t)w * N + e] = L.U;
    }
    __syncthreads();
    if (lane == 0) {
        sh.stat[6] += (double)n_acc;
        sh.stat[7] += (double)N;

        st.sigma[w] = sh.misc[0];
        st.since[w] = since;
        st.flags[w] = wflags | kFlagJrowValid;
        if ((flags & kSweepListDue) && s.reinv_k > 0 && since >= s.reinv_k) {
            const unsigned i = atomicAdd(&st.due[0], 1u);
            if (i < (unsigned)st.W) st.due_list[i] = w;
        }
        st.logdet[2 * w + 0] = ldt.fold(0, sh.misc[1]);
        if (ND > 0) st.logdet[2 * w + 1] = ldt.fold(1, sh.misc[2]);
    }
    __syncthreads();
    if (lane < kStatCount) st.stats[(size_t)w * kStatCount + lane] = sh.stat[lane];
}

// The sweep kernel, without and with a 4-waves-per-SIMD occupancy bound (the
// bound caps VGPRs at 128 at the price of some spills; the autotuner picks).
// Waves per SIMD the sweep is built for (its registers at that occupancy;
// the paired FP32 move at 27 per spin needs 100 VGPRs, four waves, without
// spilling).
template <typename Real, int NU, int ND>
constexpr int rr_sweep_waves() {
    constexpr bool kLarge = (NU > ND ? NU : ND) >= 19;
    if constexpr (sizeof(Real) == 8) return kLarge ? 3 : 4;
    if constexpr ((NU > ND ? NU : ND) >= 27) return 4;
    return kLarge ? 5 : 6;
}

template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, (rr_sweep_waves<Real, NU, ND>())) void k_sweep_rr(DevSys s, DevState st,
                                                    unsigned long long sweep0, int flags) {
    sweep_rr_body<Real, NU, ND, GPL>(s, st, blockIdx.x, sweep0, flags);
}

template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, 4) void k_sweep_rr_o4(DevSys s, DevState st,
                                                          unsigned long long sweep0, int flags) {
    sweep_rr_body<Real, NU, ND, GPL>(s, st, blockIdx.x, sweep0, flags);
}

// The same sweep with the caller's moves, recording each move (parity checks
// of the timed move code).
template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, (rr_sweep_waves<Real, NU, ND>())) void k_sweep_rr_given(
    DevSys s, DevState st, int flags, SweepGiven given) {
    sweep_rr_body<Real, NU, ND, GPL, true>(s, st, blockIdx.x, 0, flags, given);
}

typedef double rr_v4d __attribute__((ext_vector_type(4)));

// C = A * B for FP64 row-major LDS matrices of row stride kNSStride and
// order 16 TT (TT = 1 or 2), on the matrix cores: TT x TT output tiles of
// v_mfma_f64_16x16x4f64 over 4 TT k-steps.  Operand layout of that
// instruction: lane l supplies A[l & 15][l >> 4] and B[l >> 4][l & 15];
// output register q of lane l is C[4q + (l >> 4)][l & 15].
constexpr int kNSStride = 33;
template <int TT>
__device__ __forceinline__ void ns_gemm(const double *A, const double *B, rr_v4d acc[TT * TT],
                                        int lane) {
    const int lo = lane & 15, hi = lane >> 4;
#pragma unroll
    for (int t = 0; t < TT * TT; ++t) acc[t] = rr_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 4 * TT; ++kk) {
        const int k = 4 * kk + hi;
        double a[TT], b[TT];
#pragma unroll
        for (int u = 0; u < TT; ++u) {
            a[u] = A[(16 * u + lo) * kNSStride + k];
            b[u] = B[k * kNSStride + 16 * u + lo];
        }
#pragma unroll
        for (int u = 0; u < TT; ++u)
#pragma unroll
            for (int v = 0; v < TT; ++v)
                acc[u * TT + v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[v], acc[u * TT + v], 0, 0, 0);
    }
}

// Row / column of accumulator q of tile t (= u TT + v) for this lane.
template <int TT>
__device__ __forceinline__ int ns_row(int t, int q, int lane) { return 16 * (t / TT) + 4 * q + (lane >> 4); }
template <int TT>
__device__ __forceinline__ int ns_col(int t, int lane) { return 16 * (t % TT) + (lane & 15); }

// Shared memory of the re-inversion kernel: two FP64 32 x 33 matrices for
// the Newton-Schulz refinement and the Gauss-Jordan pivot workspace.
struct RRReinvShared {
    double E[32 * kNSStride];
    double Y[32 * kNSStride];
    double gjcol[32];
    int perm[32];
};

// E = D^T of one spin in FP64 into LDS (row stride `stride`): lane r of that
// spin writes column r, E[m][r] = phi_m(r_r); with stride kNSStride the
// 32 x 32 padding is zeroed too.
template <int NP, int NMAX>
__device__ __forceinline__ void ns_build_e(const DevSys &s, double *E, double x, double y,
                                                     double z, bool mine, int r, int stride) {
    constexpr auto orb = rr::shell_orbitals<NMAX>();
    if (stride == kNSStride)
        for (int e = threadIdx.x; e < 32 * kNSStride; e += kWave) {
            const int i = e / kNSStride, j = e - i * kNSStride;
            if (i >= NP || j >= NP) E[e] = 0.0;
        }
    if (mine) {
        rr::Cx<double> a[3], a2[3];
        a[0] = rr::axis_phasor<double>(x, s.inv_L);
        a[1] = rr::axis_phasor<double>(y, s.inv_L);
        a[2] = rr::axis_phasor<double>(z, s.inv_L);
#pragma unroll
        for (int c = 0; c < 3; ++c) a2[c] = rr::cmul(a[c], a[c]);
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const rr::Cx<double> ph = rr::orbital_phasor<double>(orb[m].x, orb[m].y, orb[m].z, a, a2);
            E[m * stride + r] = orb[m].t ? ph.im : ph.re;
        }
    }
}

constexpr int kNSMaxSteps = 6;

// Re-inversion of spin HP in FP64 on the matrix cores (Newton-Schulz):
// with E = D^T (E[m][i] = phi_m(r_i), built from the positions in FP64) and
// Y = invT, Z = E Y measures the drift d = |I - Z| (= |I - D^-1 D|), and
// Y <- Y (2 I - Z) leaves a drift of about d^2.  Steps repeat (E rebuilt each
// time) until d^2 is at the rounding floor of the stored inverse (a few 1e-7
// in max|I - Z| for an FP32 inverse, 1e-16 for FP64): one step for drifts up
// to 5e-4 in FP32, at most kNSMaxSteps from any drift below 0.5.  A drift of 0.5 or more (or a
// non-finite one) falls back to Gauss-Jordan from scratch.
// Returns the drift found before refinement (negative if singular); a
// Gauss-Jordan rebuild also refreshes log|det D| into `logdet`.
template <typename Real, int NU, int ND, int GPL, int HP>
__device__ double ns_reinvert(const DevSys &s, RRReinvShared &ps,
                              RRLane<Real, (NU > ND ? NU : ND)> &L, int lane, int h, int r,
                              double &logdet) {
    constexpr int NP = HP ? ND : NU;
    constexpr int NMAX = NU > ND ? NU : ND;
    // E: lane (HP, i) writes column i (its electron's D row) in FP64.
    for (int e = lane; e < 32 * kNSStride; e += kWave) ps.Y[e] = 0.0;
    ns_build_e<NP, NMAX>(s, ps.E, L.x, L.y, L.z, h == HP && r < NP, r, kNSStride);
    if (h == HP && r < NP) {
#pragma unroll
        for (int m = 0; m < NP; ++m) ps.Y[r * kNSStride + m] = (double)L.inv[m];
    }
    __syncthreads();
    constexpr int TT = NP > 16 ? 2 : 1;
    constexpr double kDone = sizeof(Real) == 4 ? 5e-4 : 1e-8;  // d below this: d^2 is at the rounding floor
    double drift0 = -1.0;
#pragma unroll 1
    for (int it = 0; it < kNSMaxSteps; ++it) {
        rr_v4d z[TT * TT];
        ns_gemm<TT>(ps.E, ps.Y, z, lane);
        double dmax = 0.0;
        bool finite = true;
#pragma unroll
        for (int t = 0; t < TT * TT; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = ns_row<TT>(t, q, lane), j = ns_col<TT>(t, lane);
                const double d = (i == j ? 1.0 : 0.0) - z[t][q];
                if (i < NP && j < NP) {
                    dmax = fabs(d) > dmax ? fabs(d) : dmax;
                    finite = finite && (d == d);
                }
            }
        // wave max of the drift
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off, kWave));
        dmax = wave_uniform(dmax);
        const int all_finite = __syncthreads_and(finite ? 1 : 0);
        if (it == 0) drift0 = all_finite ? dmax : INFINITY;
        if (!all_finite || !(dmax < 0.5)) {
            drift0 = INFINITY;
            break;
        }
        if (dmax < 1e-15) break;  // already at FP64 rounding
        __syncthreads();
        // W = 2 I - Z into E's storage, then Y <- Y W.
#pragma unroll
        for (int t = 0; t < TT * TT; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = ns_row<TT>(t, q, lane), j = ns_col<TT>(t, lane);
                ps.E[i * kNSStride + j] = (i == j ? 2.0 : 0.0) - z[t][q];
            }
        __syncthreads();
        rr_v4d y[TT * TT];
        ns_gemm<TT>(ps.Y, ps.E, y, lane);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < TT * TT; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) ps.Y[ns_row<TT>(t, q, lane) * kNSStride + ns_col<TT>(t, lane)] = y[t][q];
        __syncthreads();
        if (dmax < kDone) break;
        // another step: rebuild E (overwritten by W)
        ns_build_e<NP, NMAX>(s, ps.E, L.x, L.y, L.z, h == HP && r < NP, r, kNSStride);
        __syncthreads();
    }
    if (!(drift0 < 0.5)) {
        // Gauss-Jordan from scratch on A = D^T (row stride NP) in E's storage.
        __syncthreads();
        ns_build_e<NP, NMAX>(s, ps.E, L.x, L.y, L.z, h == HP && r < NP, r, NP);
        __syncthreads();
        const double ld = rr_gauss_jordan(ps.E, ps.gjcol, ps.perm, NP, lane);
        if (!(ld > -INFINITY)) return -1.0;
        logdet = ld;
        if (h == HP && r < NP) {
#pragma unroll
            for (int m = 0; m < NP; ++m) L.inv[m] = (Real)ps.E[r * NP + m];
        }
        __syncthreads();
        return drift0;
    }
    if (h == HP && r < NP) {
#pragma unroll
        for (int m = 0; m < NP; ++m) L.inv[m] = (Real)ps.Y[r * kNSStride + m];
    }
    __syncthreads();
    return drift0;
}

// After a sweep: the periodic FP64 re-inversion of every walker whose
// accepted moves since its last one reached K.  Stat VMC_STAT_REINV counts
// re-inversions; the largest drift |I - D^-1 D| met before one is kept per
// walker in st.drift.  Walkers not due leave at once.
// One spin of one walker: spin HP's FP64 re-inversion when the walker's
// accepted moves since its last one reached K.  The two spins of a walker run
// in two workgroups at once; they share no LDS and write disjoint inverse
// blocks and log|det| entries, and only spin 0 counts the re-inversion.  The
// walker's move count since re-inversion is cleared by its next sweep (both
// workgroups read it here).
template <typename Real, int NU, int ND, int GPL, int HP>
__device__ __forceinline__ void reinv_rr_spin(const DevSys &s, const DevState &st, int w) {
    constexpr int NMAX = NU > ND ? NU : ND;
    constexpr int N = NU + ND;
    if (s.reinv_k <= 0 || st.since[w] < s.reinv_k) return;
    __shared__ RRReinvShared ps;
    const int lane = threadIdx.x;
    const int h = lane >> 5, r = lane & 31;
    const int nh = h ? ND : NU;
    const bool has_e = r < nh;
    const int e = h ? NU + r : r;
    RRLane<Real, NMAX> L;
    Real *inv_hbm = reinterpret_cast<Real *>(st.inv) + (size_t)w * s.ninv;
    const double *gp = st.pos + (size_t)w * 3 * N;
    rr_inverse_in<Real, NU, ND, NMAX, HP>(inv_hbm, reinterpret_cast<Real *>(ps.E), L.inv, lane, h, r);
    L.x = has_e ? gp[e] : 0.0;
    L.y = has_e ? gp[N + e] : 0.0;
    L.z = has_e ? gp[2 * N + e] : 0.0;
    double ld = st.logdet[2 * w + HP];
    const double d = ns_reinvert<Real, NU, ND, GPL, HP>(s, ps, L, lane, h, r, ld);
    rr_inverse_out<Real, NU, ND, NMAX, HP>(inv_hbm, reinterpret_cast<Real *>(ps.E), L.inv, lane, h, r);
    if (lane == 0) {
        st.logdet[2 * w + HP] = ld;
        if (d < 0.0) {
            atomicOr(&st.flags[w], kFlagSingular);
        } else {
            // drift >= 0 (or +inf / NaN): its bit pattern orders as the value
            atomicMax(reinterpret_cast<unsigned long long *>(&st.drift[w]),
                      (unsigned long long)__double_as_longlong(d));
        }
        if (HP == 0) st.stats[(size_t)w * kStatCount + 15] += 1.0;
    }
}

// Workgroup g of a re-inversion launch: walker g / 2, spin g % 2 (walker g
// for one determinant).
template <typename Real, int NU, int ND, int GPL>
__device__ __forceinline__ void reinv_rr_body(const DevSys &s, const DevState &st, int g) {
    if constexpr (ND > 0) {
        if (g & 1) reinv_rr_spin<Real, NU, ND, GPL, 1>(s, st, g >> 1);
        else reinv_rr_spin<Real, NU, ND, GPL, 0>(s, st, g >> 1);
    } else {
        reinv_rr_spin<Real, NU, ND, GPL, 0>(s, st, g);
    }
}

// After a sweep: the local energy of every walker with the same lane layout
// as the sweep, feeding the FP64 accumulators and the blocking state
// (blocking_analysis.cpp:17-39).
// With `parts` the walker's kinetic, V_real, V_recip, E_L, log|det| and J go
// there ([w][kPartCount]) and the statistics are left alone.
template <typename Real, int NU, int ND, int GPL>
__device__ __forceinline__ void measure_rr_body(const DevSys &s, const DevState &st, int w,
                                                int block_size, double *parts = nullptr) {
    using K = RRKernel<Real, NU, ND, GPL>;
    constexpr int NMAX = K::NMAX;
    constexpr int N = NU + ND;
    __shared__ RRShared<Real, NU, ND> sh;
    const RRConst<Real> cst(s);
    const int lane = threadIdx.x;
    const int h = lane >> 5, r = lane & 31;
    const int nh = h ? ND : NU;
    const bool has_e = r < nh;
    const int e = h ? NU + r : r;
    RRLane<Real, NMAX> L;
    const Real *inv_hbm = reinterpret_cast<const Real *>(st.inv) + (size_t)w * s.ninv;
    const double *gp = st.pos + (size_t)w * 3 * N;
    rr_inverse_in<Real, NU, ND, NMAX>(inv_hbm, rr_stage(sh), L.inv, lane, h, r);
    L.x = has_e ? gp[e] : 0.0;
    L.y = has_e ? gp[N + e] : 0.0;
    L.z = has_e ? gp[2 * N + e] : 0.0;
    sh.pos[0][lane] = L.x;
    sh.pos[1][lane] = L.y;
    sh.pos[2][lane] = L.z;
    __syncthreads();
    const Measured m = K::measure(s, cst, sh, L, lane, has_e, h);
    if (parts) {
        const double J = 0.5 * rr::wave_total(has_e ? K::jastrow_row(cst, sh, L, lane, has_e, h) : 0.0);
        if (lane == 0) {
            double *q = parts + (size_t)w * kPartCount;
            q[0] = m.kinetic;
            q[1] = m.vreal;
            q[2] = m.vrecip;
            q[3] = m.e_local;
            q[4] = st.logdet[2 * w + 0];
            q[5] = ND > 0 ? st.logdet[2 * w + 1] : 0.0;
            q[6] = J;
        }
        return;
    }
    // The re-inversion that ran before this launch has read the list.
    if (w == 0 && lane == 0 && st.due) st.due[0] = 0;
    if (lane == 0) accumulate_measure(st.stats + (size_t)w * kStatCount, m, s.e_const, block_size);
}

// ------------------------------------------- FP64 local energy, mixed path
// A mixed-precision walker's local energy is evaluated in FP64 from its FP32
// inverse refined to FP64 accuracy first, in the same wavefront: with
// Y = X^T the FP32 inverse (X = invT, lane k of spin h holding row x_k) and
// D the FP64 Slater matrix (D[i][j] = phi_j(r_i)), a Newton-Schulz step is
//   E = I - D Y,  Y <- Y + Y E,  i.e.  x_k <- x_k + sum_i E[i][k] x_i,
// which leaves a residual of about E^2.  Lane k forms column k of E itself
// (E[i][k] = delta_ik - d_i . x_k, FP64 from the D rows staged in LDS and its
// own row) and then its correction from the X rows staged in LDS, in FP32
// (E is below 1e-4 here, so the correction's FP32 rounding sits under 1e-11
// of the row).  A second step follows when n max|E|^2 exceeds 1e-10; a
// residual of 0.5 or more (or a non-finite one) rebuilds the inverse by FP64
// Gauss-Jordan instead (also refreshing log|det|).
// LDS of the mixed-precision local energy: positions, and one staging area
// that holds the refinement's D rows (FP64) or X rows (FP32) and, after it,
// the structure-factor tables.
template <int NU, int ND>
struct RRMixedShared {
    static constexpr int NMAX = NU > ND ? NU : ND;
    static constexpr int P = (NMAX + 1) / 2 * 2;  // row pitch in doubles (16-byte rows)
    static constexpr int kRefineDoubles = 2 * NMAX * P;
    static constexpr int kTableDoubles = sf::table_reals<double>();
    double pos[3][kWave];
    alignas(16) double buf[kRefineDoubles > kTableDoubles ? kRefineDoubles : kTableDoubles];
    float et[2 * NMAX][NMAX | 1];  // each row lane's column of E (FP32; odd pitch)
    double gjcol[32];
    int perm[32];
    static constexpr size_t kStageBytes = sizeof(buf);
    __device__ void *stage() { return buf; }
    __device__ double *rows(int h) { return buf + h * NMAX * P; }
};

// A refinement step Y' = Y + Y E from the residual E = I - D Y leaves the
// residual E^2, whose entries are at most n max|E|^2: the steps stop once
// that bound is below 1e-8, which keeps the local energy's share of the
// inverse error about two orders under 1e-6 of |E_L|.  After a sweep from
// a refined inverse one step is the rule (max|E| of order 1e-6).
constexpr double kRefineTarget = 1e-8;

// Refines lane (h, r)'s row of the FP32 inverse, held widened in L.inv, to
// FP64 accuracy in place.  Returns the residual max|I - D Y| of the FP32
// inverse (the drift; +inf when it was rebuilt from scratch, -1 when a
// determinant is singular).  ld[h] receives a rebuilt determinant's
// log|det|.
template <int NU, int ND, int GPL>
__device__ double rr_refine(const DevSys &s, RRMixedShared<NU, ND> &sh, RRLane<double, (NU > ND ? NU : ND)> &L,
                            int lane, int h, int r, bool has_e, double (&ld)[2]) {
    using K = RRKernel<double, NU, ND, GPL>;
    using Sh = RRMixedShared<NU, ND>;
    constexpr int NMAX = K::NMAX, P = Sh::P;
    constexpr int NP = NU;  // equal spins or polarised: every row lane has NU rows
    const int nh = h ? ND : NU;
    constexpr auto orb = rr::shell_orbitals<NMAX>();
    // D row of the lane's electron (FP64 axis phasors) into its LDS row.
    const auto put_d_row = [&](double *dst) {
        // The position made opaque: otherwise the row, invariant across the
        // refinement steps, is hoisted out of their loop and held in
        // registers throughout.
        double px = L.x, py = L.y, pz = L.z;
        asm volatile("" : "+v"(px), "+v"(py), "+v"(pz));
        rr::Cx<double> a[3], a2[3];
        a[0] = rr::axis_phasor<double>(px, s.inv_L);
        a[1] = rr::axis_phasor<double>(py, s.inv_L);
        a[2] = rr::axis_phasor<double>(pz, s.inv_L);
        K::squares(a, a2);
#pragma unroll
        for (int m = 0; m < NMAX; ++m) {
            const rr::Cx<double> ph = rr::orbital_phasor<double>(orb[m].x, orb[m].y, orb[m].z, a, a2);
            dst[m] = orb[m].t ? ph.im : ph.re;
        }
    };
    double drift0 = 0.0;
    bool rebuild = false;
#pragma unroll 1
    for (int it = 0; it < 3; ++it) {
        if (has_e) put_d_row(sh.rows(h) + r * P);
        __syncthreads();
        // Column r of E into the lane's LDS row (a rolled loop: unrolled, the
        // rows' reads are interleaved and spill).
        double dm = 0.0;
        bool finite = true;
#pragma unroll 1
        for (int i = 0; i < NMAX; ++i) {
            const double *Di = sh.rows(h) + i * P;
            double a[4] = {0.0, 0.0, 0.0, 0.0};  // four independent chains
#pragma unroll
            for (int j = 0; j < NMAX; ++j) a[j & 3] = fma(Di[j], L.inv[j], a[j & 3]);
            const double ev = (i == r ? 1.0 : 0.0) - ((a[0] + a[1]) + (a[2] + a[3]));
            if (has_e) sh.et[h * NMAX + r][i] = (float)ev;
            if (has_e && i < nh) {
                dm = fabs(ev) > dm ? fabs(ev) : dm;
                finite = finite && (ev == ev);
            }
        }
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) dm = fmax(dm, __shfl_xor(dm, off, kWave));
        dm = wave_uniform(dm);
        const int all_finite = __syncthreads_and(finite ? 1 : 0);
        if (it == 0) drift0 = dm;
        if (!all_finite || !(dm < 0.5)) {
            rebuild = true;
            break;
        }
        // X rows in the D rows' place: the current rows rounded to FP32 (the
        // correction Y E needs Y only to the accuracy of its small factor E).
        float *X = reinterpret_cast<float *>(sh.rows(h));
        if (has_e) {
#pragma unroll
            for (int m = 0; m < NMAX; ++m) X[r * 2 * P + m] = (float)L.inv[m];
        }
        __syncthreads();
        float c[NMAX];
#pragma unroll
        for (int j = 0; j < NMAX; ++j) c[j] = 0.0f;
#pragma unroll 1
        for (int i = 0; i < NMAX; ++i) {
            const float *Xi = X + i * 2 * P;
            const float e = has_e ? sh.et[h * NMAX + r][i] : 0.0f;
#pragma unroll
            for (int j = 0; j < NMAX; ++j) c[j] = fmaf(e, Xi[j], c[j]);
        }
#pragma unroll
        for (int j = 0; j < NMAX; ++j) L.inv[j] += (double)c[j];
        __syncthreads();
        if ((double)NP * dm * dm <= kRefineTarget) break;
    }
    if (rebuild) {
        // FP64 Gauss-Jordan of each determinant from scratch, A = D^T.
#pragma unroll 1
        for (int hp = 0; hp < (ND > 0 ? 2 : 1); ++hp) {
            double *A = sh.buf;
            if (h == hp && has_e) {
                double d[NMAX];
                put_d_row(d);
#pragma unroll
                for (int m = 0; m < NMAX; ++m) A[m * NP + r] = d[m];
            }
            __syncthreads();
            const double l = rr_gauss_jordan(A, sh.gjcol, sh.perm, NP, lane);
            if (!(l > -INFINITY)) return -1.0;
            ld[hp] = l;
            if (h == hp && has_e) {
#pragma unroll
                for (int m = 0; m < NMAX; ++m) L.inv[m] = A[r * NP + m];
            }
            __syncthreads();
        }
        return INFINITY;
    }
    return drift0;
}

// Mixed-precision local energy (k_measure_rr for FP32 inverses): the FP32
// inverse refined to FP64 (rr_refine), then the FP64 measurement.  The
// refined inverse, rounded to FP32, replaces the stored one (this is the
// walker's periodic re-inversion: move count cleared, drift recorded) unless
// only the parts are asked for.
template <int NU, int ND, int GPL>
__device__ __forceinline__ void measure_rr_refined_body(const DevSys &s, const DevState &st, int w,
                                                        int block_size, double *parts = nullptr) {
    using K = RRKernel<double, NU, ND, GPL>;
    constexpr int NMAX = K::NMAX;
    constexpr int N = NU + ND;
    __shared__ RRMixedShared<NU, ND> sh;
    const RRConst<double> cst(s);
    const int lane = threadIdx.x;
    const int h = lane >> 5, r = lane & 31;
    const int nh = h ? ND : NU;
    const bool has_e = r < nh;
    const int e = h ? NU + r : r;
    RRLane<double, NMAX> L;
    {
        float x[NMAX];
        const float *inv_hbm = reinterpret_cast<const float *>(st.inv) + (size_t)w * s.ninv;
        rr_inverse_in<float, NU, ND, NMAX>(inv_hbm, reinterpret_cast<float *>(sh.buf), x, lane, h, r);
#pragma unroll
        for (int m = 0; m < NMAX; ++m) L.inv[m] = (double)x[m];
    }
    const double *gp = st.pos + (size_t)w * 3 * N;
    L.x = has_e ? gp[e] : 0.0;
    L.y = has_e ? gp[N + e] : 0.0;
    L.z = has_e ? gp[2 * N + e] : 0.0;
    sh.pos[0][lane] = L.x;
    sh.pos[1][lane] = L.y;
    sh.pos[2][lane] = L.z;
    double ld[2] = {st.logdet[2 * w + 0], st.logdet[2 * w + 1]};
    const double drift = rr_refine<NU, ND, GPL>(s, sh, L, lane, h, r, has_e, ld);
    if (!parts) {
        float xo[NMAX];
#pragma unroll
        for (int m = 0; m < NMAX; ++m) xo[m] = (float)L.inv[m];
        float *inv_hbm = reinterpret_cast<float *>(st.inv) + (size_t)w * s.ninv;
        rr_inverse_out<float, NU, ND, NMAX>(inv_hbm, reinterpret_cast<float *>(sh.buf), xo, lane, h, r);
        if (lane == 0) {
            st.since[w] = 0;
            st.stats[(size_t)w * kStatCount + 15] += 1.0;
            if (drift < 0.0) {
                st.flags[w] |= kFlagSingular;
            } else {
                st.logdet[2 * w + 0] = ld[0];
                st.logdet[2 * w + 1] = ld[1];
                if (drift > st.drift[w] || drift != drift) st.drift[w] = drift;
            }
        }
    }
    __syncthreads();
    const Measured m = K::measure(s, cst, sh, L, lane, has_e, h);
    if (parts) {
        const double J = 0.5 * rr::wave_total(has_e ? K::jastrow_row(cst, sh, L, lane, has_e, h) : 0.0);
        if (lane == 0) {
            double *q = parts + (size_t)w * kPartCount;
            q[0] = m.kinetic;
            q[1] = m.vreal;
            q[2] = m.vrecip;
            q[3] = m.e_local;
            q[4] = st.logdet[2 * w + 0];
            q[5] = ND > 0 ? st.logdet[2 * w + 1] : 0.0;
            q[6] = J;
        }
        return;
    }
    if (w == 0 && lane == 0 && st.due) st.due[0] = 0;
    if (lane == 0) accumulate_measure(st.stats + (size_t)w * kStatCount, m, s.e_const, block_size);
}

// The re-inversion over the walkers the preceding sweep listed (sweep
// flag kSweepListDue): workgroup g takes items g, g + G, ... (item i: walker
// due_list[i / spins], spin i % spins), so a launch of G workgroups costs the
// due walkers' work rather than one workgroup per walker and spin.  The next
// kernel on the stream empties the list (k_measure_rr, else k_due_reset).
template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave) void k_reinv_rr_list(DevSys s, DevState st) {
    constexpr unsigned kSpins = ND > 0 ? 2 : 1;
    const unsigned n = st.due[0] < (unsigned)st.W ? st.due[0] : (unsigned)st.W;
    for (unsigned i = blockIdx.x; i < n * kSpins; i += gridDim.x)
        reinv_rr_body<Real, NU, ND, GPL>(s, st, (int)((unsigned)st.due_list[i / kSpins] * kSpins + i % kSpins));
}

// Empties the re-inversion list after k_reinv_rr_list when no local energy
// follows.
template <typename Real>
__global__ void k_due_reset(DevState st) {
    st.due[0] = 0;
}

// Waves per SIMD the measurement is built for: the structure-factor loop
// would otherwise hold many staged values in flight at low occupancy.
// FP64 measurements take the registers they need.
template <typename Real, int N, int GPL>
constexpr int rr_measure_waves() {
    if constexpr (sizeof(Real) == 8) return 1;
    return GPL <= 5 ? 5 : (N <= kRRChunk ? 1 : 4);
}

template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, (rr_measure_waves<Real, NU + ND, GPL>())) void k_measure_rr(DevSys s, DevState st,
                                                                               int block_size) {
    measure_rr_body<Real, NU, ND, GPL>(s, st, blockIdx.x, block_size);
}

// The local energy's parts per walker, statistics untouched.
template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave) void k_measure_rr_parts(DevSys s, DevState st, double *parts) {
    measure_rr_body<Real, NU, ND, GPL>(s, st, blockIdx.x, 1, parts);
}

// Mixed precision: the FP64 local energy from the refined FP32 inverse
// (which it also refreshes), its parts-only form, and the batch form.
template <int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, ND > 0 ? 2 : 1) void k_measure_rr_mixed(DevSys s, DevState st, int block_size) {
    measure_rr_refined_body<NU, ND, GPL>(s, st, blockIdx.x, block_size);
}

template <int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, ND > 0 ? 2 : 1) void k_measure_rr_mixed_parts(DevSys s, DevState st, double *parts) {
    measure_rr_refined_body<NU, ND, GPL>(s, st, blockIdx.x, 1, parts);
}

template <int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, ND > 0 ? 2 : 1) void k_measure_rr_mixed_b(const BatchItem *__restrict__ items, int n_items,
                                                              int block_size) {
    const BatchItem &it = items[batch_find(items, n_items, blockIdx.x)];
    measure_rr_refined_body<NU, ND, GPL>(it.s, it.st, blockIdx.x - it.walker_begin, block_size);
}

// Batch forms: one launch covers the walkers of every system of a batch group
// (systems of one size); workgroup g runs walker g - walker_begin of the
// system whose range holds g, with that system's tables and state, and the
// sweep index comes from device memory (so a captured graph continues the
// chains).  A system's chain is the same in a batch and alone.
template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, (rr_sweep_waves<Real, NU, ND>())) void k_sweep_rr_b(
    const BatchItem *__restrict__ items, int n_items,
                                                      const unsigned long long *__restrict__ sweep_ctr,
                                                      int flags) {
    const BatchItem &it = items[batch_find(items, n_items, blockIdx.x)];
    sweep_rr_body<Real, NU, ND, GPL>(it.s, it.st, blockIdx.x - it.walker_begin, *sweep_ctr, flags);
}

template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave) void k_reinv_rr_b(const BatchItem *__restrict__ items, int n_items) {
    constexpr int kSpins = ND > 0 ? 2 : 1;
    const int g = blockIdx.x / kSpins, sp = blockIdx.x % kSpins;
    const BatchItem &it = items[batch_find(items, n_items, g)];
    reinv_rr_body<Real, NU, ND, GPL>(it.s, it.st, (g - it.walker_begin) * kSpins + sp);
}

template <typename Real, int NU, int ND, int GPL>
__global__ __launch_bounds__(kWave, (rr_measure_waves<Real, NU + ND, GPL>())) void k_measure_rr_b(
    const BatchItem *__restrict__ items, int n_items, int block_size) {
    const BatchItem &it = items[batch_find(items, n_items, blockIdx.x)];
    measure_rr_body<Real, NU, ND, GPL>(it.s, it.st, blockIdx.x - it.walker_begin, block_size);
}

}  // namespace vmc
