/*
 * vmc_hip.h -- C ABI of the MI355X-native VMC sweep for the 3D homogeneous
 * electron gas (libvmc_hip.so, gfx950).
 *
 * The reference has no FFI: its hot path is in-process C++ classes
 * (SlaterPlaneWave, JastrowPade, EnergyTracker, WaveFunction, Simulation in
 * src/).  This header is the drop-in boundary that replaces them.  Each entry
 * point names the reference interface it stands in for (file:line, paths
 * relative to the reference repository root).  The C++ facade in
 * variational-monte-carlo_amd/host/ restores the reference's class
 * signatures above it.
 *
 * Conventions
 *  - Every function returns VMC_OK (0) or a VMC_ERR_* code; the message of
 *    the last failure on the calling thread is vmc_last_error().  No C++
 *    exception crosses this boundary.
 *  - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *    Functions that take or return host buffers synchronise that stream
 *    before returning; the others are stream-ordered.
 *  - Positions are AoS [walker][electron][3] doubles in [0, L).
 *  - An ensemble is not thread-safe: one host thread per handle, or
 *    external synchronisation.
 */
#ifndef VMC_HIP_H
#define VMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VMC_OK = 0,
    VMC_ERR_INVALID = 1,  /* bad argument (reference: std::invalid_argument, config.hpp:64-75) */
    VMC_ERR_DEVICE = 2,   /* HIP runtime failure */
    VMC_ERR_SINGULAR = 3, /* no non-singular start found (reference: std::runtime_error,
                             simulation.cpp:68-70) */
    VMC_ERR_MEMORY = 4,   /* device allocation failed */
    VMC_ERR_STATE = 5,    /* call out of order */
    VMC_ERR_CANCELLED = 6 /* a served request ended by vmc_server_cancel */
};

enum { VMC_PREC_FP64 = 0, VMC_PREC_MIXED = 1 };
enum { VMC_MOVE_UNIFORM = 0, VMC_MOVE_GAUSSIAN = 1 };
enum { VMC_ERFC_AS7126 = 0, VMC_ERFC_EXACT = 1 };

/* Sweep kernel selection (vmc_ensemble_set_kernel). */
enum { VMC_KERNEL_AUTO = 0, VMC_KERNEL_GENERIC = 1, VMC_KERNEL_REGISTER = 2, VMC_KERNEL_DELAYED = 3 };

/* Sweep flags. */
enum { VMC_SWEEP_MEASURE = 1, VMC_SWEEP_ADAPT = 2 };

/* One HEG system.  Zero/negative optional fields select the reference's own
 * rules, so {n_up=N, n_dn=0, L, a, a, b, 0, 0, 0, AS7126, UNIFORM, FP64, 0}
 * is exactly the reference's physics (slater_plane_wave.cpp:30-38,
 * jastrow_pade.hpp:10-18, energy_tracking.cpp:6-70). */
typedef struct vmc_system {
    int32_t n_up;              /* electrons of determinant 0 (all, when n_dn == 0) */
    int32_t n_dn;              /* electrons of determinant 1; 0 = polarised, one determinant */
    double box_length;         /* L (reference Config::box_length, config.hpp:20) */
    double jastrow_a_same;     /* Pade a for parallel spins (reference Jastrow_A) */
    double jastrow_a_anti;     /* Pade a for antiparallel spins */
    double jastrow_b;          /* Pade b (reference Jastrow_B) */
    double ewald_alpha;        /* <= 0: reference 6/L (energy_tracking.cpp:7) */
    int32_t recip_nmax2;       /* > 0: G = 2 pi n / L with 0 < |n|^2 <= recip_nmax2, half space;
                                  0: reference rule G^2 <= 4 alpha^2 ln(1e6) (energy_tracking.cpp:13-21) */
    double real_cutoff;        /* > 0: spherical real-space cutoff; <= 0: every minimum-image pair */
    int32_t erfc_kind;         /* VMC_ERFC_*; AS7126 is energy_tracking.cpp:101-139 */
    int32_t move_kind;         /* VMC_MOVE_*; UNIFORM is simulation.cpp:101-103 */
    int32_t precision;         /* VMC_PREC_* */
    int32_t reinvert_interval; /* K: accepted moves per determinant between FP64
                                  re-inversions (a walker is due once its accepted moves
                                  over all determinants reach K x determinants); 0 = never
                                  (the reference never re-inverts); at most 2^24 */
} vmc_system;

/* Per-walker FP64 statistics, vmc_ensemble_read_stats() layout (doubles). */
enum {
    VMC_STAT_NMEAS = 0,    /* measured sweeps */
    VMC_STAT_SUM_E = 1,    /* sum of local energies */
    VMC_STAT_SUM_E2 = 2,   /* sum of squared local energies */
    VMC_STAT_SUM_KIN = 3,  /* sum of kinetic parts */
    VMC_STAT_SUM_POT = 4,  /* sum of potential parts */
    VMC_STAT_LAST_E = 5,   /* last local energy */
    VMC_STAT_ACCEPTED = 6, /* accepted moves since the last reset */
    VMC_STAT_PROPOSED = 7, /* proposed moves since the last reset */
    VMC_STAT_BLK_SUM = 8,  /* blocking (blocking_analysis.cpp:17-39): open block sum */
    VMC_STAT_BLK_IN = 9,   /* samples in the open block */
    VMC_STAT_BLK_N = 10,   /* completed blocks */
    VMC_STAT_BLK_MEAN = 11,/* Welford mean of block means */
    VMC_STAT_BLK_M2 = 12,  /* Welford M2 of block means */
    VMC_STAT_LAST_KIN = 13,
    VMC_STAT_LAST_POT = 14,
    VMC_STAT_REINV = 15,   /* FP64 re-inversions performed */
    VMC_STAT_COUNT = 16
};

/* Scalars returned per configuration by vmc_frozen_eval(). */
enum {
    VMC_FZ_LOGDET0 = 0,     /* log|det D0|   (slater_plane_wave.cpp:202-318) */
    VMC_FZ_LOGDET1 = 1,     /* log|det D1|   (0 when polarised) */
    VMC_FZ_JASTROW = 2,     /* J             (jastrow_pade.cpp:7-55) */
    VMC_FZ_KINETIC = 3,     /* -1/2 sum(lap + |grad|^2) (energy_tracking.cpp:253-278) */
    VMC_FZ_VREAL = 4,       /* real-space Ewald  (energy_tracking.cpp:93-144) */
    VMC_FZ_VRECIP = 5,      /* reciprocal Ewald  (energy_tracking.cpp:72-91,146-191) */
    VMC_FZ_ELOCAL = 6,      /* kinetic + potential (energy_tracking.cpp:367-382) */
    VMC_FZ_RATIO = 7,       /* determinant ratio of the move (slater_plane_wave.cpp:351-365) */
    VMC_FZ_DELTA_J = 8,     /* Jastrow change of the move (jastrow_pade.cpp:140-200) */
    VMC_FZ_VREAL_INC = 9,   /* V_real after the move, incremental (energy_tracking.cpp:280-365) */
    VMC_FZ_VRECIP_INC = 10, /* V_recip after the move, incremental (energy_tracking.cpp:193-251) */
    VMC_FZ_ELOCAL_AFTER = 11, /* local energy after the move, full re-evaluation */
    VMC_FZ_KINETIC_AFTER = 12,
    VMC_FZ_JASTROW_AFTER = 13, /* J + delta_J tracked incrementally */
    VMC_FZ_SINGULAR = 14,   /* 1 if a determinant was singular */
    VMC_FZ_COUNT = 15
};

typedef struct vmc_ensemble vmc_ensemble;

/* Message of the last failure on this thread ("" if none). */
const char *vmc_last_error(void);

/* Library and device sanity: returns VMC_OK when a gfx950 device is usable. */
int vmc_device_check(int32_t *device_count);

/* Derived sizes of a system: unique k per determinant (n_k[2]), half-space G
 * count, inverse elements per walker (n_up^2 + n_dn^2). */
int vmc_system_sizes(const vmc_system *sys, int32_t *n_k, int32_t *n_g, int32_t *n_inv);

/* Ensemble of independent walkers resident on the current device.  Replaces
 * one Simulation per std::async thread (simulation.hpp:77-85, main.cpp:50-68)
 * and its Particles/SlaterPlaneWave/JastrowPade/EnergyTracker state
 * (particles.hpp:10-75, slater_plane_wave.hpp:147-169, jastrow_pade.hpp:5-44,
 * energy_tracking.hpp:10-93).  `step0` is the initial proposal width (uniform
 * half-width or Gaussian sigma; reference default L/10, config.hpp:78). */
int vmc_ensemble_create(const vmc_system *sys, int64_t n_walkers, uint64_t seed, double step0,
                        vmc_ensemble **out);
int vmc_ensemble_destroy(vmc_ensemble *e);

/* Uniform random start with the singular-retry loop of
 * Simulation::initialize_positions (simulation.cpp:43-76): up to 100 Philox
 * draws per walker; VMC_ERR_SINGULAR if a walker found none. */
int vmc_ensemble_init_positions(vmc_ensemble *e, void *stream);

/* Host positions in/out ([W][N][3]).  set_positions rebuilds every inverse in
 * FP64 (the log_abs_det path, slater_plane_wave.cpp:202-318). */
int vmc_ensemble_set_positions(vmc_ensemble *e, const double *pos, void *stream);
int vmc_ensemble_get_positions(vmc_ensemble *e, double *pos, void *stream);

/* n_sweeps sequential sweeps (every electron proposed once per sweep, in
 * index order, or with equal spins alternating up 0, down 0, up 1, ...) of
 * every walker in one launch.  Replaces the warmup/measure
 * loops over Simulation::metropolis_step (simulation.cpp:80-263).
 * VMC_SWEEP_MEASURE evaluates the local energy after each sweep and feeds the
 * per-walker accumulators and blocking (block_size measured sweeps per
 * block); VMC_SWEEP_ADAPT applies the warmup step rule
 * s *= exp(0.25 (acc - 0.5)), s <= L/2 after each sweep
 * (simulation.cpp:178-194).  Other flag bits are rejected (VMC_ERR_INVALID). */
int vmc_ensemble_sweep(vmc_ensemble *e, int32_t n_sweeps, int32_t flags, int32_t block_size,
                       void *stream);

/* Values recorded per move by vmc_ensemble_sweep_given() and per walker by
 * vmc_ensemble_measure(). */
enum {
    VMC_MV_RATIO = 0,    /* determinant ratio R (slater_plane_wave.cpp:351-365) */
    VMC_MV_DELTA_J = 1,  /* Jastrow change dJ (jastrow_pade.cpp:140-200) */
    VMC_MV_ACCEPTED = 2, /* 1 if u < R^2 exp(2 dJ) (simulation.cpp:125-131), else 0 */
    VMC_MV_COUNT = 3
};
enum {
    VMC_PART_KINETIC = 0, /* -1/2 sum(lap + |grad|^2) of ln psi (energy_tracking.cpp:253-278) */
    VMC_PART_VREAL = 1,   /* real-space Ewald (energy_tracking.cpp:93-144) */
    VMC_PART_VRECIP = 2,  /* reciprocal Ewald (energy_tracking.cpp:72-91) */
    VMC_PART_ELOCAL = 3,  /* kinetic + potential + Ewald constants (energy_tracking.cpp:367-382) */
    VMC_PART_LOGDET0 = 4, /* log|det D0| (slater_plane_wave.cpp:202-318): FP64 at the last
                             (re)build, then tracked through the accepted moves' ratios */
    VMC_PART_LOGDET1 = 5, /* log|det D1| (0 when polarised) */
    VMC_PART_JASTROW = 6, /* J = sum_{i<j} u(r_ij) (jastrow_pade.cpp:7-55) */
    VMC_PART_COUNT = 7
};

/* One sweep of every walker on the kernel the ensemble's sweeps use, with
 * the caller's moves in place of the Philox draws: electron p of walker w is
 * proposed at proposals[w][p][0..2] (in [0, L)) and accepted iff
 * uniforms[w][p] < R^2 exp(2 dJ) -- the reference's
 * Simulation::metropolis_step (simulation.cpp:80-156) with its proposal and
 * uniform supplied.  The electron order is vmc_ensemble_sweep's; each
 * move's ratio, Jastrow change and decision go to trace[w][p][VMC_MV_*].
 * No measurement or step adaptation; the Philox stream position does not
 * advance; the accept counters count the moves and the walkers that came due
 * are re-inverted as after a sweep.  A uniform of 0 accepts every move with
 * R != 0.  For parity checks of the timed kernels against a reference
 * implementation on the same moves. */
int vmc_ensemble_sweep_given(vmc_ensemble *e, const double *proposals, const double *uniforms, double *trace,
                             void *stream);

/* One sweep of every walker (as vmc_ensemble_sweep with n_sweeps = 1, on the
 * generic kernel) with walker `walker`'s local energy evaluated after each of
 * its moves: the reference's measurement cadence, E_L after every step
 * (Simulation::measure, src/simulation/simulation.cpp:203-243), for the walker
 * whose frames the host writes.  Per move t in sweep order (host buffers of
 * N entries): e_local[t], moved[t] (the electron), accepted[t] (0 or 1), and
 * when pos is not NULL the walker's positions after the move, pos[t][N][3].
 * FP64 systems only (VMC_ERR_INVALID otherwise). */
int vmc_ensemble_sweep_traced(vmc_ensemble *e, int32_t flags, int32_t block_size, int64_t walker,
                              double *e_local, int32_t *moved, int32_t *accepted, double *pos, void *stream);

/* The local energy and its parts, log|det| and J of every walker at its
 * current positions, evaluated by the local-energy kernel the ensemble's
 * measured sweeps use (WaveFunction::evaluate_derivatives +
 * EnergyTracker::eval_total_energy, wavefunction.cpp:12-58,
 * energy_tracking.cpp:367-382).  parts [W][VMC_PART_COUNT]; the statistics
 * are left alone. */
int vmc_ensemble_measure(vmc_ensemble *e, double *parts, void *stream);

/* vmc_ensemble_measure without any allocation: no hipMalloc, hipFree or
 * allocating library call.  No counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it).  The per-walker parts are staged in a
 * device array of the library, one per GPU, held until the process ends:
 * at most VMC_NOALLOC_MAX_WALKERS walkers (VMC_ERR_INVALID above that).
 * Calls that overlap on one GPU share that array and race; serialise them.
 * parts [W][VMC_PART_COUNT] host buffer, as vmc_ensemble_measure. */
#define VMC_NOALLOC_MAX_WALKERS 65536
int vmc_ensemble_measure_noalloc(vmc_ensemble *e, double *parts, void *stream);

/* The walkers' transposed inverses as doubles, [W][n_up^2 + n_dn^2]
 * (invT[k][j] = (D^-1)[j][k], slater_plane_wave.cpp:312-314; spin up
 * first); mixed-precision inverses are widened from FP32. */
int vmc_ensemble_get_inverse(vmc_ensemble *e, double *inv, void *stream);

/* Choose the sweep kernel: VMC_KERNEL_GENERIC keeps the inverses in LDS (or
 * HBM) for any N; VMC_KERNEL_REGISTER keeps each inverse row in a lane's
 * registers (7, 19, 27, 33, 57 or 81 electrons per spin, FP32 also 123 and
 * 147; equal spins or polarised); VMC_KERNEL_DELAYED (FP64, at most 160
 * electrons per spin) keeps the inverse in HBM and applies accepted moves
 * as delayed rank-8 Woodbury updates on the matrix cores.
 * VMC_KERNEL_AUTO (default) takes the register kernel whenever it applies,
 * else, in FP64, the delayed-update kernel when the inverses do not fit the
 * generic kernel's LDS.  No reference counterpart: the reference has a
 * single CPU code path (its rank-1 update, slater_plane_wave.cpp:367-423);
 * this is the autotuner's knob.  VMC_ERR_INVALID if the system does not fit
 * the requested kernel. */
int vmc_ensemble_set_kernel(vmc_ensemble *e, int32_t kernel);
/* The kernel the next sweep launch will use (VMC_KERNEL_GENERIC, _REGISTER
 * or _DELAYED). */
int vmc_ensemble_get_kernel(const vmc_ensemble *e, int32_t *kernel);
/* Register-kernel build variant: 0 = unbounded registers, 1 = at most 128
 * VGPRs (4 waves per SIMD, some spills).  No reference counterpart; an
 * autotuner knob. */
int vmc_ensemble_set_rr_variant(vmc_ensemble *e, int32_t variant);
/* Walker groups of the register kernels' measured sweeps: the walkers are
 * split into `groups` contiguous ranges (1 to 8; default 1), each swept and
 * measured on a library stream of its own, forked from and joined back into
 * the caller's stream, so that one group's sweep overlaps another's local
 * energy.  Each library stream is made with a CU mask of every CU, which
 * gives it a hardware queue of its own.  The chains are the same for any
 * group count, because the draws are keyed by the ensemble's walker index.
 * Sweeps timed per kernel (vmc_ensemble_set_timing) and sweeps followed by a
 * listed FP64 re-inversion run as one group.  Measured at N = 54 with 4096
 * walkers and 3 groups: the FP64 step is about 8 % faster, the mixed step
 * within a few percent either way.  No reference counterpart: the reference
 * runs one chain per thread (src/main.cpp:50-68). */
int vmc_ensemble_set_walker_groups(vmc_ensemble *e, int32_t groups);

/* Zero the per-walker statistics (start of a measurement phase). */
int vmc_ensemble_reset_stats(vmc_ensemble *e, void *stream);
/* Copy per-walker statistics to host: [W][VMC_STAT_COUNT]. */
int vmc_ensemble_read_stats(vmc_ensemble *e, double *stats, void *stream);
/* Copy per-walker inverse drift to host: [W], the largest max-norm of
 * I - D^-1 D met when a walker's mixed-precision inverse came due for its
 * FP64 re-inversion (0 until one happens).  No reference
 * counterpart: the reference works in FP64 throughout and never re-inverts. */
int vmc_ensemble_read_drift(vmc_ensemble *e, double *drift, void *stream);
/* Per-kernel device time: while enabled, every sweep launch records HIP
 * events on its own stream around each kernel.  read_timing waits for them
 * and returns (then forgets) the summed milliseconds and launch counts per
 * kernel kind, arrays of VMC_TIMING_KINDS: [0] sweep kernel (the generic
 * kernel's fused sweep, re-inversion and measurement included), [1] FP64
 * re-inversion kernel, [2] local-energy kernel.  No reference counterpart
 * (instrumentation). */
#define VMC_TIMING_KINDS 3
int vmc_ensemble_set_timing(vmc_ensemble *e, int32_t enable);
int vmc_ensemble_read_timing(vmc_ensemble *e, double *ms, int64_t *launches);
/* Copy per-walker step sizes to host: [W]. */
int vmc_ensemble_read_steps(vmc_ensemble *e, double *steps, void *stream);
/* One walker's output frame: its VMC_STAT_COUNT statistics and its positions
 * ([N][3], AoS) -- the data behind Simulation::measure's FrameData
 * (src/simulation/simulation.cpp:226-235) for the walker the host streams
 * out.  Either pointer may be NULL. */
int vmc_ensemble_read_frame(vmc_ensemble *e, int64_t walker, double *stats, double *pos,
                            void *stream);
/* Global sweep counter (the Philox stream position) and walker count. */
int vmc_ensemble_info(const vmc_ensemble *e, int64_t *n_walkers, uint64_t *sweeps_done);

/* Frozen-configuration evaluation for parity: for each configuration c,
 * builds D and its FP64 inverse, evaluates every energy term, proposes the
 * move of electron moved[c] to newpos[c] (ratio, delta J), applies it
 * unconditionally (Sherman-Morrison, incremental S_G and V_real) and
 * re-evaluates.  Uses the generic kernel's device routines: the per-move
 * values (ratio, delta J, inverses, incremental Ewald updates) in the
 * system's precision, log|det|, J, the derivatives, S_G and the energies in
 * FP64 (as the mixed-precision sweeps measure).  Host buffers:
 *   pos [C][N][3], moved [C], newpos [C][3]        (inputs)
 *   scalars [C][VMC_FZ_COUNT]                       (required)
 *   grad_lap [C][N][4]   grad x,y,z and lap of ln psi before the move (optional)
 *   inv_before, inv_after [C][n_inv] transposed inverses
 *                        (invT[k][j] = (D^-1)[j][k], slater_plane_wave.cpp:312-314) (optional)
 *   s_k [C][n_g][2]      structure factor before the move (optional)        */
int vmc_frozen_eval(const vmc_system *sys, int32_t n_configs, const double *pos,
                    const int32_t *moved, const double *newpos, double *scalars, double *grad_lap,
                    double *inv_before, double *inv_after, double *s_k, void *stream);

/* ------------------------------------------------------------------ batch
 * Systems of uneven size, r_s, Jastrow parameters and walker counts swept
 * together.  The reference runs one system per process and one chain per
 * thread (src/main.cpp:50-68); a batch runs many systems' walker ensembles --
 * each an independent vmc_ensemble with its own seed -- with one launch per
 * kernel per sweep for each group of systems: the register-resident systems
 * of one size form a group (one workgroup per walker across all of them),
 * and the systems on the generic kernel share one persistent launch whose
 * workgroups take walker indices from a device queue.  Groups run on streams
 * of their own, forked from and joined back into the caller's stream around
 * each sweep (parallel branches of a captured graph).  The sweep index lives
 * in device memory.  A walker's chain depends only on its system, seed,
 * walker index, sweep index and kernel, so every system gives the same
 * result in a batch as alone with the same kernel.  All systems of a batch
 * share one working precision.
 *
 * inv_placement[i] (NULL = all 0): generic kernel's inverse, 0 in LDS when
 * it fits, 1 LDS (error if it does not fit), 2 HBM.  walker_threads (64, 128
 * or 256) is the generic persistent sweep's workgroup per walker (1, 2 or 4
 * wavefronts). */
typedef struct vmc_batch vmc_batch;
int vmc_batch_create(const vmc_system *systems, const int64_t *n_walkers, const uint64_t *seeds,
                     const double *step0, const int32_t *inv_placement, int32_t n_systems,
                     int32_t walker_threads, vmc_batch **out);
int vmc_batch_destroy(vmc_batch *b);
int vmc_batch_init_positions(vmc_batch *b, void *stream);
/* n_sweeps batch sweeps (flags as for vmc_ensemble_sweep, block_size per
 * system in measured sweeps). */
int vmc_batch_sweep(vmc_batch *b, int32_t n_sweeps, int32_t flags, int32_t block_size, void *stream);
/* Captures n_sweeps batch sweeps into a HIP graph (on a stream the batch
 * owns); vmc_batch_replay launches it on `stream`.  Each replay continues
 * the chains.  Changing kernels or overlap drops the graph. */
int vmc_batch_capture(vmc_batch *b, int32_t n_sweeps, int32_t flags, int32_t block_size);
int vmc_batch_replay(vmc_batch *b, void *stream);
/* Borrowed handle of system i for the vmc_ensemble_read_* / get_positions /
 * reset_stats calls; sweeping or destroying it directly is refused. */
int vmc_batch_member(vmc_batch *b, int32_t system, vmc_ensemble **out);
/* grid: workgroups of one sweep's sweep launches over all groups; lds_bytes:
 * the generic group's dynamic LDS per workgroup. */
int vmc_batch_info(const vmc_batch *b, int32_t *n_systems, int64_t *total_walkers, int32_t *grid,
                   int64_t *lds_bytes);
/* Kernel of system `system` (-1: every system) as for vmc_ensemble_set_kernel;
 * regroups the batch (synchronises the device).  No reference counterpart. */
int vmc_batch_set_kernel(vmc_batch *b, int32_t system, int32_t kernel);
/* overlap != 0 (default): groups on their own streams; 0: all on the
 * caller's stream, one after another.  No reference counterpart. */
int vmc_batch_set_overlap(vmc_batch *b, int32_t overlap);
/* Groups: their count, family[g] (0 generic, 1 register, 2 multi-wavefront
 * register) and group_of[i] for each system, up to `cap` entries each. */
int vmc_batch_groups(const vmc_batch *b, int32_t *n_groups, int32_t *family, int32_t *group_of, int32_t cap);

/* ---------------------------------------------------- placement, footprint */
/* Generic kernel's inverse placement: 0 LDS when it fits (default), 1 LDS
 * (error if it does not fit), 2 HBM.  No reference counterpart (a tuning
 * knob of the device layout). */
int vmc_ensemble_set_inverse_placement(vmc_ensemble *e, int32_t placement);
/* Generic kernel's workgroup per walker: 64, 128 or 256 threads (1, 2 or 4
 * wavefronts share one walker's orbital row, Jastrow pairs, Sherman-Morrison
 * rows and energy sums; default 64).  No reference counterpart. */
int vmc_ensemble_set_walker_threads(vmc_ensemble *e, int32_t threads);
/* Device bytes an ensemble of this system and walker count holds (state and
 * tables), the admission measure of the request server: per walker the
 * positions, inverses, FP64 Gauss-Jordan scratch, statistics, step state,
 * drift, Jastrow row sums and the re-inversion list entry, and for mixed
 * ensembles the FP64 inverses a measured sweep re-inverts into and, above
 * 80 electrons per spin, the Newton-Schulz residual.  No reference
 * counterpart. */
int vmc_ensemble_footprint(const vmc_system *sys, int64_t n_walkers, int64_t *bytes);

/* --------------------------------------------------------------- autotuner
 * Picks, for one system and walker count, the kernel family (register
 * resident or generic), its build variant, the wavefronts per walker and the
 * inverse placement (LDS or HBM) of the generic kernel and -- in mixed
 * precision -- the longest FP64 re-inversion interval K
 * in {256, 64, 16} whose measured inverse drift stays below drift_tol, by
 * timing `sweeps` measured sweeps of each candidate with HIP events.  No
 * reference counterpart (the reference has no device layout to tune). */
typedef struct vmc_tuning {
    int32_t kernel;            /* VMC_KERNEL_GENERIC, _REGISTER or _DELAYED */
    int32_t walker_threads;    /* generic kernel: 64, 128 or 256 threads per walker */
    int32_t rr_variant;        /* register kernel build variant */
    int32_t inv_placement;     /* generic kernel: 1 LDS, 2 HBM */
    int32_t reinvert_interval; /* chosen K (mixed); the system's own value in FP64 */
    double moves_per_s;        /* electron-moves/s of the chosen settings */
    double drift;              /* largest inverse drift met at the chosen K */
} vmc_tuning;
int vmc_autotune(const vmc_system *sys, int64_t n_walkers, double drift_tol, int32_t sweeps,
                 vmc_tuning *out, void *stream);
/* The walker workgroup the request server and the C++ facade use for a
 * system: 64 threads below 64 electrons, 128 below 128, else 256. */
int vmc_default_walker_threads(const vmc_system *sys);

/* ------------------------------------------------------------------ server
 * Continuous batching of energy evaluations (the runs the reference's
 * Jastrow optimiser issues, src/optimizer/jastrow_optimizer.cpp:21-171, and
 * main's chains, src/main.cpp:50-68).  A worker thread admits requests in
 * arrival order while their device footprint fits hbm_budget_bytes and at
 * most max_active run; each admitted request owns a stream and advances
 * sweeps_per_turn sweeps per scheduler turn (at most two turns in flight),
 * so requests that arrive while others run start at once and their kernels
 * overlap.  Walker state stays on the device from admission to result.  A
 * request's result depends only on the request: the same alone and in any
 * mix. */
typedef struct vmc_request {
    vmc_system system;
    int64_t n_walkers;
    int32_t warmup_sweeps;    /* adaptive (simulation.cpp:159-196) */
    int32_t measure_sweeps;   /* measured, E_L per sweep */
    int32_t block_size;       /* measured sweeps per blocking block; below 1 the first
                                 measured sweep fails and ends the request */
    uint64_t seed;
    double step0;
} vmc_request;

typedef struct vmc_result {
    int64_t id;
    int32_t status;           /* VMC_OK or the error that ended the request (its
                                 admission or one of its sweeps) */
    int64_t walkers;
    double mean_energy;       /* walker aggregate as src/main.cpp:79-104 */
    double standard_error;    /* < 0 when no walker completed two blocks */
    double acceptance_rate;
    double moves;             /* electron moves of the sweeps that ran */
    double queued_s;          /* submit -> admission */
    double latency_s;         /* submit -> result */
} vmc_result;

typedef struct vmc_server vmc_server;
int vmc_server_create(int64_t hbm_budget_bytes, int32_t max_active, int32_t sweeps_per_turn,
                      vmc_server **out);
/* Thread-safe.  VMC_ERR_MEMORY when the request alone exceeds the budget. */
int vmc_server_submit(vmc_server *s, const vmc_request *req, int64_t *id);
/* Up to `capacity` finished results; with wait != 0 blocks until one is
 * ready or nothing is in flight. */
int vmc_server_poll(vmc_server *s, vmc_result *out, int32_t capacity, int32_t *n_out, int32_t wait);
/* Cancels request `id`: a request still waiting for admission ends at once
 * (no sweeps), a running one after the turn in flight; either reports a
 * result with status VMC_ERR_CANCELLED and the moves its sweeps made.
 * VMC_ERR_STATE when the request already finished or the id is unknown.
 * No reference counterpart: the reference's chains always run to the end
 * (src/main.cpp:50-68, futures joined at :75-76). */
int vmc_server_cancel(vmc_server *s, int64_t id);
/* Finishes every submitted request, then stops the worker. */
int vmc_server_destroy(vmc_server *s);

/* ------------------------------------------------------------------ per-move objects
 * One configuration's determinant, Jastrow and Ewald pieces with the
 * reference's per-move calls.  The state the reference keeps in class members
 * is resident on the current device.  Each call is one small launch on the
 * object's own stream and returns when its result is on the host.  FP64, the
 * reference's physics: one plane-wave determinant over all n particles, one
 * Pade a, alpha = 6/L, the A&S 7.1.26 erfc.  They back the host facade's
 * SlaterPlaneWave, JastrowPade, EnergyTracker and WaveFunction classes for
 * one walker; the ensemble API above is the throughput path.  Positions are
 * caller-owned SoA arrays x[n], y[n], z[n] as in Particles
 * (particles.hpp:10-75), n in [1, VMC_SINGLE_MAX_N].  Objects are not
 * thread-safe (as the reference's). */
#define VMC_SINGLE_MAX_N 1024

typedef struct vmc_slater vmc_slater;
/* SlaterPlaneWave(particles, L) (slater_plane_wave.cpp:30-152): closed-shell
 * orbitals, k = 2 pi n / L, zeroed trig cache and matrices. */
int vmc_slater_create(int32_t n, double box_length, vmc_slater **out);
int vmc_slater_destroy(vmc_slater *s);
/* Unique k count and the row strides of the n x n matrices and of the trig
 * cache: n and n_unique_k padded to 8 doubles, as the reference pads
 * (aligned_soa.hpp:63-65, slater_plane_wave.cpp:31-33,138). */
int vmc_slater_shape(const vmc_slater *s, int32_t *n_unique_k, int32_t *matrix_stride, int32_t *trig_stride);
/* n vectors and k vectors [n_unique_k][3], orbital -> k index [n], orbital
 * type [n] (0 cos, 1 sin) (slater_plane_wave.hpp:83-95,108-138). */
int vmc_slater_tables(const vmc_slater *s, int32_t *n_vec, double *k_vec, int32_t *orbital_k,
                      uint8_t *orbital_type);
/* log_abs_det (slater_plane_wave.cpp:202-318): trig cache of every particle,
 * D, partial-pivot LU with pivot tolerance 1e-12 (matrix.hpp:13-59),
 * log|det| = sum log|U_ii|, and the transposed inverse by n LU solves
 * (matrix.hpp:61-96).  -inf for a singular D, the inverse then unchanged. */
int vmc_slater_log_abs_det(vmc_slater *s, const double *x, const double *y, const double *z,
                           double *log_abs_det);
/* update_trig_cache, save_trig_row, restore_trig_row
 * (slater_plane_wave.cpp:154-200); p in [0, n). */
int vmc_slater_update_trig(vmc_slater *s, int32_t p, double x, double y, double z);
int vmc_slater_save_trig(vmc_slater *s, int32_t p);
int vmc_slater_restore_trig(vmc_slater *s, int32_t p);
/* build_row (slater_plane_wave.cpp:320-349): the n orbitals at particle p's
 * trig cache row, copied to row. */
int vmc_slater_build_row(vmc_slater *s, int32_t p, double *row);
/* determinant_ratio (slater_plane_wave.cpp:351-365). */
int vmc_slater_ratio(vmc_slater *s, int32_t p, const double *row, double *ratio);
/* accept_move (slater_plane_wave.cpp:367-423): Sherman-Morrison update of the
 * transposed inverse and row p of D replaced by row. */
int vmc_slater_accept(vmc_slater *s, int32_t p, const double *row, double ratio);
/* add_derivatives (slater_plane_wave.cpp:425-499): adds grad log|D| and
 * lap log|D| of every particle into the arrays (in/out, n each). */
int vmc_slater_add_derivatives(vmc_slater *s, double *grad_x, double *grad_y, double *grad_z, double *lap);
/* Copies of the device state: D, the transposed inverse and the LU factors
 * [n][matrix_stride]; the sin and cos caches [n][trig_stride]; pivots [n]. */
enum { VMC_SLATER_D = 0, VMC_SLATER_INV = 1, VMC_SLATER_LU = 2, VMC_SLATER_SIN = 3, VMC_SLATER_COS = 4 };
int vmc_slater_read(const vmc_slater *s, int32_t which, double *out);
int vmc_slater_pivots(const vmc_slater *s, int32_t *pivots);

typedef struct vmc_pade vmc_pade;
/* JastrowPade(L, a, b) (jastrow_pade.hpp:5-44): u(r) = a r / (1 + b r) over
 * minimum-image pairs.  Positions are passed with every call. */
int vmc_pade_create(double box_length, double a, double b, vmc_pade **out);
int vmc_pade_destroy(vmc_pade *j);
/* value (jastrow_pade.cpp:7-55): J = sum over pairs of u(r_ij). */
int vmc_pade_value(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z, double *value);
/* add_derivatives (jastrow_pade.cpp:57-138): adds grad_i J and lap_i J
 * (in/out arrays). */
int vmc_pade_add_derivatives(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z,
                             double *grad_x, double *grad_y, double *grad_z, double *lap);
/* delta_value (jastrow_pade.cpp:140-200): the positions hold particle `moved`
 * at its new place, old_x/y/z is where it was. */
int vmc_pade_delta_value(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z, int32_t moved,
                         double old_x, double old_y, double old_z, double *delta);
/* update_derivatives_for_move (jastrow_pade.cpp:203-317): updates Jastrow
 * grad/lap arrays (in/out) for one move, positions as for delta_value. */
int vmc_pade_update_for_move(vmc_pade *j, int32_t n, const double *x, const double *y, const double *z,
                             int32_t moved, double old_x, double old_y, double old_z, double *grad_x,
                             double *grad_y, double *grad_z, double *lap);

typedef struct vmc_ewald vmc_ewald;
/* EnergyTracker(L, N) (energy_tracking.cpp:6-70): alpha = 6/L, half-space G
 * with G^2 <= 4 alpha^2 ln 1e6 and their weights, self and background terms;
 * S_G, V_real and V_recip are 0 until initialised. */
int vmc_ewald_create(double box_length, double num_particles, vmc_ewald **out);
int vmc_ewald_destroy(vmc_ewald *t);
/* initialize_structure_factors (energy_tracking.cpp:146-191). */
int vmc_ewald_init_structure_factors(vmc_ewald *t, int32_t n, const double *x, const double *y, const double *z);
/* initialize_reciprocal_energy from the cached S_G (energy_tracking.cpp:72-91). */
int vmc_ewald_init_reciprocal(vmc_ewald *t);
/* initialize_real_energy (energy_tracking.cpp:93-144). */
int vmc_ewald_init_real(vmc_ewald *t, int32_t n, const double *x, const double *y, const double *z);
/* update_structure_factors (energy_tracking.cpp:193-251): S_G and V_recip
 * for one particle moved from old to new. */
int vmc_ewald_update_structure_factors(vmc_ewald *t, double old_x, double old_y, double old_z, double new_x,
                                       double new_y, double new_z);
/* update_real_energy (energy_tracking.cpp:280-365): positions with particle
 * `moved` at its new place. */
int vmc_ewald_update_real(vmc_ewald *t, int32_t n, const double *x, const double *y, const double *z, int32_t moved,
                          double old_x, double old_y, double old_z);
/* eval_total_energy (energy_tracking.cpp:253-278,367-382): -1/2 sum(lap +
 * |grad|^2) of the arrays plus the cached V_real + V_recip + self +
 * background. */
int vmc_ewald_total_energy(vmc_ewald *t, int32_t n, const double *grad_x, const double *grad_y,
                           const double *grad_z, const double *lap, double *energy);
/* Cached state; s_re and s_im ([n_g] each) may be NULL. */
int vmc_ewald_state(const vmc_ewald *t, double *v_real, double *v_recip, double *self_term, double *background,
                    int32_t *n_g, double *s_re, double *s_im);

#ifdef __cplusplus
}
#endif

#endif
