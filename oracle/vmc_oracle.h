/*
 * vmc_oracle.h -- FP64 CPU oracle for the VMC sweep of the 3D homogeneous
 * electron gas.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * (tests/, __graft_entry__.smoke()) and the CPU baseline leg of bench.py.
 * The product library (libvmc_hip.so) never links, loads or calls it.
 *
 * Every function restates the reference algorithm operation by operation and
 * cites the reference file:line it follows (paths relative to the reference
 * repository root).  Extensions the GPU path needs beyond the reference
 * (two spin determinants, spin-resolved Jastrow, parameterised Ewald,
 * Gaussian moves, Philox streams, per-sweep measurement) are composed from
 * the restated pieces and say so.
 */
#ifndef VMC_ORACLE_H
#define VMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- system */
typedef struct {
    int n_up;          /* electrons in determinant 0 (all of them when n_dn == 0) */
    int n_dn;          /* electrons in determinant 1; 0 = polarised single determinant */
    double L;          /* cubic box length */
    double a_same;     /* Pade a for parallel spins (the only a in polarised mode) */
    double a_anti;     /* Pade a for antiparallel spins */
    double b;          /* Pade b */
    double alpha;      /* Ewald splitting; <= 0 selects the reference rule 6/L */
    int recip_nmax2;   /* > 0: half-space G with |n|^2 <= recip_nmax2;
                          0: reference tolerance rule G^2 <= 4 alpha^2 ln(1e6) */
    double rcut;       /* > 0: real-space spherical cutoff; <= 0: all minimum-image pairs */
    int erfc_kind;     /* 0 = Abramowitz-Stegun 7.1.26 (reference); 1 = libm erfc */
} orc_system;

/* ------------------------------------------------------------ RNG pieces */
typedef struct {
    uint64_t mt[312];
    int idx;
} orc_mt64;

void orc_mt64_seed(orc_mt64 *g, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64 *g);
double orc_std_canonical(orc_mt64 *g);                    /* generate_canonical<double,53> */
double orc_std_uniform_real(orc_mt64 *g, double a, double b);
uint64_t orc_std_uniform_index(orc_mt64 *g, uint64_t n); /* uniform_int_distribution{0,n-1} */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---------------------------------------------------------- shell filling */
/* Returns the number of unique k; nvec gets nk*3 ints, orb_k/orb_type n each. */
int orc_shell_fill(int n, int *nvec, int *orb_k, unsigned char *orb_type);

/* --------------------------------------------------------- Ewald tables */
/* Returns the number of half-space G vectors; if gx != NULL fills gx/gy/gz/w
   (capacity cap).  alpha_out receives the alpha actually used. */
int orc_ewald_gvectors(const orc_system *sys, double *gx, double *gy, double *gz, double *w,
                       int *nvec, int cap, double *alpha_out);
double orc_ewald_self(const orc_system *sys);
double orc_ewald_background(const orc_system *sys);

/* ------------------------------------------------------------- walker */
typedef struct orc_walker orc_walker;

orc_walker *orc_walker_new(const orc_system *sys);
void orc_walker_free(orc_walker *w);
int orc_walker_num_electrons(const orc_walker *w);
int orc_walker_num_g(const orc_walker *w);
int orc_walker_num_k(const orc_walker *w, int spin);

/* pos is AoS [N][3].  Builds D, the inverse, log|det|, J, S_G, V_recip, V_real
   (simulation.cpp:43-76 minus the random draw).  Returns 0, or 1 if either
   determinant is singular (log|det| = -inf, slater_plane_wave.cpp:299-301). */
int orc_walker_set_positions(orc_walker *w, const double *pos);
void orc_walker_get_positions(const orc_walker *w, double *pos);

double orc_walker_logdet(const orc_walker *w, int spin);
void orc_walker_inverse(const orc_walker *w, int spin, double *inv_t); /* n*n, invT[c*n+r]=Dinv[r][c] */
void orc_walker_dmatrix(const orc_walker *w, int spin, double *d);     /* n*n row-major */
double orc_walker_jastrow(const orc_walker *w);                        /* J = sum_{i<j} u */
double orc_walker_vreal(const orc_walker *w);
double orc_walker_vrecip(const orc_walker *w);
void orc_walker_structure_factor(const orc_walker *w, double *s_re, double *s_im);

/* Full derivative refresh (wavefunction.cpp:12-58): grad [N][3], lap [N] of ln psi. */
void orc_walker_derivatives(orc_walker *w, double *grad, double *lap);
/* Kinetic energy from the last derivative refresh (energy_tracking.cpp:253-278). */
double orc_walker_kinetic(const orc_walker *w);
/* Derivative refresh + kinetic + potential (energy_tracking.cpp:367-382). */
double orc_walker_local_energy(orc_walker *w, double *kinetic, double *potential);

/* Trial move of electron p to newpos (wrapped into [0,L) by the caller).
   ratio: slater_plane_wave.cpp:154-200,320-365; delta_j: jastrow_pade.cpp:140-200.
   Nothing is committed. */
double orc_walker_ratio(orc_walker *w, int p, const double newpos[3]);
double orc_walker_delta_j(const orc_walker *w, int p, const double newpos[3]);
/* Commit the move (simulation.cpp:133-145): SM update, S_G and V_real updates,
   Jastrow derivative cache update. */
void orc_walker_accept(orc_walker *w, int p, const double newpos[3]);
/* Incremental derivative path (wavefunction.cpp:60-114). */
void orc_walker_derivatives_after_step(orc_walker *w, int accepted, int p, const double oldpos[3],
                                       double *grad, double *lap);

/* ------------------------------------------------------ blocking analysis */
typedef struct {
    size_t num_blocks;
    double running_mean;
    double running_m2;
    size_t block_size;
    size_t in_block;
    double block_sum;
} orc_blocking;

void orc_blocking_init(orc_blocking *b, size_t block_size);
void orc_blocking_add(orc_blocking *b, double x);
int orc_blocking_ready(const orc_blocking *b);
/* Returns 0 and fills mean/se, or 1 when fewer than two blocks exist. */
int orc_blocking_mean_se(const orc_blocking *b, double *mean, double *se);

/* --------------------------------------------- reference simulation driver */
typedef struct {
    size_t num_particles;
    double box_length;
    size_t warmup_steps;
    size_t measure_steps;
    double step_size;
    uint64_t seed;
    size_t block_size;
    double jastrow_a;
    double jastrow_b;
} orc_ref_config;

typedef struct {
    double mean_energy;
    double standard_error;     /* 0 when unavailable */
    int has_standard_error;
    double acceptance_rate;
    size_t accepted;
    size_t proposed;
    double final_step_size;
} orc_ref_summary;

/* Simulation::run (simulation.cpp:265-290) with mt19937_64 + libstdc++
   distributions.  If frames_e != NULL it receives measure_steps local
   energies; frames_pos (optional) receives measure_steps*N*3 positions. */
int orc_ref_simulation_run(const orc_ref_config *cfg, orc_ref_summary *out, double *frames_e,
                           double *frames_mean, double *frames_se, double *frames_acc,
                           double *frames_pos);

/* ------------------------------------------- ensemble sweep driver (GPU twin) */
enum { ORC_MOVE_UNIFORM = 0, ORC_MOVE_GAUSSIAN = 1 };

typedef struct {
    uint64_t seed;
    int move_kind;
    double step0;            /* initial step (sigma for Gaussian, half-width for uniform) */
} orc_sweep_config;

/* Draws the Philox initial configuration of walker `walker` (retrying up to
   100 attempts while a determinant is singular).  Returns the attempt used or
   -1.  pos receives AoS [N][3]. */
int orc_init_positions(const orc_system *sys, uint64_t seed, uint64_t walker, double *pos);

/* One sequential sweep of walker `walker` at global sweep index `sweep`:
   electrons 0..N-1 in order, or, with equal spins, the spins alternating
   (up 0, down 0, up 1, down 1, ...).  Returns the number of accepted moves; *step is adapted when
   adapt != 0.  When e_local != NULL the local energy after the sweep is
   written there. */
int orc_sweep(orc_walker *w, const orc_sweep_config *cfg, uint64_t walker, uint64_t sweep,
              double *step, int adapt, double *e_local);

/* Threaded CPU baseline: one independent chain per thread (main.cpp:50-68),
   `sweeps` sweeps each, measuring every sweep when measure != 0.  Returns
   aggregate proposed moves per second; *elapsed gets the wall time. */
double orc_bench_threads(const orc_system *sys, const orc_sweep_config *cfg, int threads,
                         int sweeps, int measure, double *elapsed);

#ifdef __cplusplus
}
#endif

#endif
