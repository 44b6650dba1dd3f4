/*
 * mcgb.h — C ABI of the MI355X-native Gay-Berne checkerboard Monte Carlo library (libmcgb.so).
 *
 * Plain C: no C++ or torch types cross this boundary.  All device memory is owned by the
 * opaque context; callers pass host arrays that are copied in and out.
 *
 * Each entry point names the reference interface it replaces (MCnptColloid.c file:line).
 * Return convention: 0 on success, a negative MCGB_ERR_* code on failure; the text of the last
 * error is available from mcgb_last_error().  The reference prints and calls exit(1) instead
 * (e.g. MCnptColloid.c:3862-3866, :4100-4127); the C driver maps any error to that behaviour.
 *
 * Threading: one context is driven by one host thread; calls on the same context are not
 * thread-safe.  Each context owns one HIP stream at the default priority.
 */
#ifndef MCGB_H
#define MCGB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCGB_OK            0
#define MCGB_ERR_ARG      -1   /* invalid argument */
#define MCGB_ERR_IO       -2   /* file could not be opened / read */
#define MCGB_ERR_PARAM    -3   /* parameter file: missing key or malformed value */
#define MCGB_ERR_HIP      -4   /* HIP runtime error (including "no GPU") */
#define MCGB_ERR_NOMEM    -5   /* allocation failed */
#define MCGB_ERR_GEOMETRY -6   /* box too small for the 8-sublattice cell grid, bad move size */
#define MCGB_ERR_STATE    -7   /* state not set, overlap with the colloid core, NaN energy */
#define MCGB_ERR_CAPACITY -8   /* a fixed-size buffer limit was exceeded */

/* Number of keys the reference's loader requires (MCnptColloid.c:3937, :4100). */
#define MCGB_NUM_REQUIRED_KEYS 23

/*
 * Run parameters.  The first 23 fields are the reference's parametersFile.txt keys with the
 * reference's names and C types (globals MCnptColloid.c:96-145, parsed at :3952-4089).
 * The remaining fields are optional keys that the reference does not have; they take the
 * defaults set by mcgb_params_default() when absent from the file.
 */
typedef struct mcgb_params {
    int      mc_ensemble_type;              /* 1 = NVT, 2 = NPT (ln V moves)              */
    int      num_particles;
    double   system_density;
    double   system_temperature;
    double   system_pressure;
    double   gb_sigma_sidetoside;           /* sigma_0                                    */
    double   gb_epsilon0;
    double   gb_kappa;
    double   gb_kappa_prime;
    double   gb_mu_exponent;
    double   gb_nu_exponent;
    float    max_displacement;              /* full width: step = d*(U-1/2) per axis       */
    float    max_rotation;                  /* |d_rot * v| with v on the unit sphere      */
    int      initial_configuration_type;
    double   cutoff_radius;
    double   verlet_radius;                 /* read and kept; the cell list replaces lists */
    int      total_mc_cycles;               /* driver: number of sweeps                    */
    int      configuration_save_interval;
    int      statistics_print_interval;
    int      colloid_wall_option;           /* 1 = colloid at the box centre               */
    double   colloid_interaction_epsilon;   /* read, unused (as in the reference)          */
    double   colloid_interaction_amplitude; /* read, unused (as in the reference)          */
    double   colloid_radius;

    /* ---- optional keys (no counterpart in the reference) ---- */
    uint64_t seed;                          /* Philox key; "seed"                          */
    int      precision;                     /* 32: FP32 pair math, 64: FP64; "precision"   */
    double   npt_max_dlnv;                  /* max |d ln V| per volume move; "npt_max_dlnv" */
    int      energy_check_interval;         /* sweeps between FP64 re-summations, 0 = off  */
    double   lattice_jitter;                /* init: uniform jitter amplitude of sites     */
    double   colloid_exclusion_radius;      /* init: sites closer than this are removed    */
    int      stage_capacity;                /* particles staged in LDS per stencil, 0: auto */
    int      waves_per_cell;                /* wavefronts per active cell: 1, 2, 3, 4 or 8  */
    int      pass_schedule;                 /* 1: speculative (moves of a cell evaluated in
                                               parallel, decided in order), 0: sequential,
                                               2: sequential built for four wavefronts per
                                               SIMD (FP32; FP64 runs 0), 3: two moves of a
                                               cell in flight, one per wavefront (two
                                               wavefronts per cell; waves_per_cell ignored) */
    int      adapt_moves;                   /* step-size adaptation every >= adapt_moves
                                               attempted moves (the reference: 10000), 0 = off */
    int      sublattice_colors;             /* 2: 8 sublattices of cells >= r_c wide (default);
                                               3: 27 sublattices of cells >= r_c/2 wide      */

    uint32_t keys_read;                     /* bit i set: required key i was read          */
} mcgb_params;

/* Outcome of one trial move (trace records). */
#define MCGB_MOVE_EMPTY   0   /* active cell had no particle: no move made                 */
#define MCGB_MOVE_LEFT    1   /* trial position left the active cell: rejected             */
#define MCGB_MOVE_HARD    2   /* trial position inside the colloid core: rejected          */
#define MCGB_MOVE_REJECT  3   /* Metropolis rejection                                      */
#define MCGB_MOVE_ACCEPT  4   /* accepted                                                  */

/* One trial move, as emitted by mcgb_sweep_traced() (32 bytes). */
typedef struct mcgb_move_trace {
    int32_t pid;      /* particle id moved (-1 for an empty cell)                              */
    int32_t outcome;  /* MCGB_MOVE_*                                                           */
    int32_t dcap;     /* change of the number of capped GB pairs (U = 1e9, MCnptColloid.c:2670) */
    int32_t dccap;    /* change of the number of capped colloid terms (U = 1e8, :1778, :1833)   */
    double  dfin;     /* finite part of dE                                                     */
    double  scale;    /* sum of |finite pair terms| of both states                              */
} mcgb_move_trace;

/* Per-replica running statistics. */
typedef struct mcgb_stats {
    uint64_t sweeps;
    uint64_t attempted;            /* every trial move, including the two early rejections  */
    uint64_t accepted;
    uint64_t rejected_left_cell;
    uint64_t rejected_hard_core;
    uint64_t candidates;           /* stencil candidates tested (both states)              */
    uint64_t pairs;                /* GB pair energies inside the cutoff (both states)      */
    uint64_t colloid_evals;        /* colloid-rod energies evaluated (both states)          */
    uint64_t volume_attempted;
    uint64_t volume_accepted;
    int64_t  cap_pairs;            /* running number of capped GB pairs                     */
    int64_t  cap_colloid;          /* running number of capped colloid terms                */
    double   energy_finite;        /* running finite part of the energy (FP64)              */
    double   energy;               /* cap_pairs*1e9 + cap_colloid*1e8 + energy_finite       */
    double   box[3];
    double   volume;
    double   density;              /* N / (V - V_colloid), MCnptColloid.c:2799-2804          */
    double   max_drift;            /* largest |running - recomputed| at a periodic
                                      re-summation (energy_check_interval, mcgb_resum_energy) */
    uint64_t resummations;         /* FP64 re-summations of the running energy (all kinds)   */
    uint64_t da_stage1_rejected;   /* delayed acceptance: moves rejected by the surrogate    */
    uint64_t da_stage2_rejected;   /* delayed acceptance: rejected by the exact second stage */
    double   max_drift_forced;     /* the same at re-summations forced by a sweep whose
                                      accepted moves moved > 1e10 of finite energy (near-cap
                                      pairs), where the running sum has lost its precision    */
    uint64_t resummations_forced;  /* how many of the re-summations were forced               */
} mcgb_stats;

typedef struct mcgb_ctx mcgb_ctx;

/* ---------------------------------------------------------------- parameters (host only) */

/* Defaults for every field (required keys zeroed, keys_read = 0). */
void mcgb_params_default(mcgb_params *p);

/* Replaces load_initial_conditions() (MCnptColloid.c:3933-4133) and read_next_parameter()
 * (:3892-3906): same 23 keys, same "key value # comment" format, lines starting with '#' or
 * empty are skipped, unknown keys print "Warning: Unknown parameter '<k>' in configuration
 * file", a missing key or a malformed value is an error (MCGB_ERR_PARAM; the message lists the
 * missing keys).  Optional keys are recognised in addition. */
int mcgb_params_load(const char *path, mcgb_params *out);

/* Derived constants of get_system_const() (MCnptColloid.c:2752-2783), computed in dependency
 * order: out = {beta, 4 eps0, rc^2, kappa^2, chi, chi^2, kappa'^(1/mu), chi', sigma_ee,
 * colloid cutoff 4 sigma_ee, R^2}. */
int mcgb_params_derived(const mcgb_params *p, double out[11]);

/* ---------------------------------------------------------------- context */

/* Creates a context holding n_replicas independent systems; p points to n_replicas parameter
 * blocks (each replica has its own N, T*, P*, density, colloid).  All replicas must share
 * p[0].precision.  Replaces the allocation/broadcast set-up of main() (MCnptColloid.c:379-530)
 * and block_init()/cell_init() (:3283-3388).  Fails with MCGB_ERR_HIP when no GPU is present. */
int mcgb_create(const mcgb_params *p, int n_replicas, mcgb_ctx **out);
void mcgb_destroy(mcgb_ctx *ctx);

/* Text of the last error on this thread ("" if none). */
const char *mcgb_last_error(void);

/* Number of GPUs visible to HIP (0 when none); never fails. */
int mcgb_device_count(void);

/* Sets replica r's configuration: n particles (must equal its num_particles), positions in
 * [0, box) (wrapped otherwise), orientations (normalised here), box lengths.  Positions are
 * stored as 32-bit fixed-point fractions of the box (x = q * L / 2^32).  Recomputes the FP64
 * total energy.  Replaces the scatter of main() (MCnptColloid.c:528-609).
 * The box may be rectangular; every edge must hold 4 cells of the cutoff.  The cell arrays are
 * laid out at mcgb_create for the cubic box of num_particles / system_density (NPT: 1.3x its
 * edge); a box whose own cell grid needs more cells than that coarsens its longest axis (wider
 * cells, still exact, but not the reference's grid).  Give system_density the density of the
 * box you will set to keep its own grid. */
int mcgb_set_state(mcgb_ctx *ctx, int replica, int n,
                   const double *x, const double *y, const double *z,
                   const double *ux, const double *uy, const double *uz,
                   const double box[3]);

/* Same with the fixed-point positions given directly (bit-exact parity tests). */
int mcgb_set_state_fixed(mcgb_ctx *ctx, int replica, int n,
                         const uint32_t *qx, const uint32_t *qy, const uint32_t *qz,
                         const double *ux, const double *uy, const double *uz,
                         const double box[3]);

/* Copies replica r's configuration back, in particle-id order: converted to FP64 on the device
 * (x = q L / 2^32, the host's product) and copied through a pinned record of its own. */
int mcgb_get_state(mcgb_ctx *ctx, int replica, int n,
                   double *x, double *y, double *z,
                   double *ux, double *uy, double *uz, double box[3]);
int mcgb_get_state_fixed(mcgb_ctx *ctx, int replica, int n,
                         uint32_t *qx, uint32_t *qy, uint32_t *qz,
                         double *ux, double *uy, double *uz, double box[3]);

/* ---------------------------------------------------------------- hot path */

/* Runs n_sweeps checkerboard sweeps on every replica.  One sweep = random grid shift, cell-list
 * rebuild, P sublattice passes of ceil(N/n_cells) trial moves per active cell, then (NPT) one
 * ln V move; P = 8 with sublattice_colors 2 (cells >= r_c wide), 27 with 3 (cells >= r_c/2
 * wide, partners within +-2 cells).  Replaces the sub-cell loop of main() (MCnptColloid.c:643-706) with its TODO body
 * filled by the move of attempt_monte_carlo_move() (:2445-2604), and build_neighbor_cells()
 * (:5048-5264) / clean_neighbor_cells() (:5035-5046) by the device cell list. */
int mcgb_sweep(mcgb_ctx *ctx, int n_sweeps);

/* Runs ONE plain Metropolis sweep of a single-replica context and records every trial move
 * (MCGB_ERR_ARG with more than one replica or with delayed acceptance enabled).
 * Record index = pass*(n_cells/P)*k + active_cell*k + move; capacity must be >= n_cells*k.
 * info receives {shift_x, shift_y, shift_z, k, n_cells, P, order[0..P-1]}, the rest 0. */
int mcgb_sweep_traced(mcgb_ctx *ctx, mcgb_move_trace *trace, int capacity, uint32_t info[36]);

/* Running statistics of replica r. */
int mcgb_get_stats(mcgb_ctx *ctx, int replica, mcgb_stats *out);

/* Resets the running counters of replica r (state and energy are kept). */
int mcgb_reset_counters(mcgb_ctx *ctx, int replica);

/* FP64 full recompute of the energy of replica r over the cell list (half stencil,
 * deterministic reduction).  out = {total, cap_pairs, cap_colloid, finite}.  Replaces
 * tot_ene() (MCnptColloid.c:1598-1622) / tot_ene_lstnb() (:1857-1882). */
int mcgb_total_energy(mcgb_ctx *ctx, int replica, double out[4]);

/* Re-sums the FP64 energy and replaces the running total with it (drift bound). */
int mcgb_resum_energy(mcgb_ctx *ctx, int replica, double *drift);

/* Nematic order parameter of replica r: P2 = 1.5 * largest eigenvalue of Q = <uu> - I/3
 * (FP64 sums on the GPU, closed-form 3x3 eigen solve on the host) and the director.
 * Replaces compute_liquid_crystal_order_parameters() (MCnptColloid.c:2117-2218). */
int mcgb_order_parameter(mcgb_ctx *ctx, int replica, double *p2, double director[3]);

/* Per-sweep series recorded since the last mcgb_series_clear(): E/N, P2, density.
 * Returns the number of entries (<= max) or a negative error code. */
int mcgb_series(mcgb_ctx *ctx, int replica, int max,
                double *energy_per_particle, double *p2, double *density);
int mcgb_series_clear(mcgb_ctx *ctx);

/* Autotuner: times `sweeps` sweeps with each pass schedule (0 sequential, 1 speculative,
 * 2 sequential built for four wavefronts per SIMD, 3 two moves in flight) and each supported
 * number of wavefronts per active cell (1, 2, 3, 4, 8; schedule 3: 2 only; the LDS staging
 * capacity follows) and keeps the fastest; with schedule 0 kept it then re-checks the pacing
 * (mcgb_set_pace, one warm-up and `sweeps` timed sweeps per setting).  The timed sweeps are ordinary sweeps of the
 * simulation (they advance it).  No counterpart in the reference.  ms_per_pass (may be NULL)
 * receives the measured pass times, index 5*schedule + {0,1,2,3,4} for 1, 2, 3, 4, 8
 * wavefronts (0 for a combination it does not run: the speculative schedule with
 * sublattice_colors 3, schedule 2 in FP64, schedule 3 with other than 2 wavefronts). */
int mcgb_autotune(mcgb_ctx *ctx, int sweeps, int *best_waves, int *best_schedule,
                  double ms_per_pass[20]);

/* Pacing of the sequential pass schedule (schedule 0): the cells of a pass run side by side and
 * the pass lasts as long as its slowest cell, so a cell that falls behind `percent` of the last
 * sweep's mean time per move (timed from the start of its workgroup) raises its wavefronts'
 * issue priority over the cells sharing their SIMDs (in steps of half a move).  Scheduling
 * only: the chain and every result are the same at any setting.  0 turns it off; the default is
 * 109, and mcgb_autotune, when it keeps schedule 0, re-checks off, 109 and 112 and keeps the
 * fastest.  No counterpart in the
 * reference (its cells run one after another on a CPU core). */
int mcgb_set_pace(mcgb_ctx *ctx, int percent);
int mcgb_get_pace(mcgb_ctx *ctx, int *percent);

/* Average device time per launch of the sweep-pass kernel, the cell-list rebuild and the whole
 * sweep, measured with HIP events on the context's stream over the last mcgb_sweep() call. */
int mcgb_kernel_times(mcgb_ctx *ctx, double *ms_pass, double *ms_rebuild, double *ms_sweep);

/* Kernel timing on (default) or off.  On: the last two sweeps of every mcgb_sweep() call are
 * launched kernel by kernel between HIP events (for mcgb_kernel_times); off: every sweep is a
 * graph replay.  The sweeps' results do not depend on it. */
int mcgb_set_kernel_timing(mcgb_ctx *ctx, int on);

/* Cell list for sweep index t of replica r without moving anything: grid shift, sublattice
 * order, cell_start[n_cells+1] and the particle ids in cell order.  Replaces
 * get_owner_block_index()/get_owner_cell_index() (MCnptColloid.c:4371-4399) and
 * dist_batch_to_cells() (:4942-4964). */
int mcgb_cell_list(mcgb_ctx *ctx, int replica, uint32_t t,
                   int *cell_start, int cell_start_len, int *pid_sorted, int n,
                   uint32_t shift[3], uint32_t order[27], int dims[3]);

/* ---------------------------------------------------------------- kernels on batches */

/* GB pair energy of n pairs on the GPU: U(e_i, e_j, dr) exactly as pij_GB()
 * (MCnptColloid.c:2618-2681) with inverse norms and e_i.e_j formed as in GBij() (:1524-1535).
 * precision 32 or 64.  capped[k] = 1 where the reference returns its 1e9 cap; u[k] then
 * holds 1e9.  ei, ej, dr are n*3 interleaved. */
int mcgb_pair_energy_batch(const mcgb_params *p, int n, const double *ei, const double *ej,
                           const double *dr, int precision, double *u, int *capped);

/* The same as mcgb_pair_energy_batch without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The device copies of the inputs and outputs live in a library-owned
 * static array of MCGB_NOALLOC_MAX_PAIRS pairs (84 bytes each, ~5.5 MB) on every GPU the library
 * runs on, held until the process ends; larger n fails with MCGB_ERR_CAPACITY.  Every call on
 * one GPU shares that GPU's array, so calls that overlap on one GPU (several host threads) race. */
#define MCGB_NOALLOC_MAX_PAIRS 65536
int mcgb_pair_energy_batch_noalloc(const mcgb_params *p, int n, const double *ei,
                                   const double *ej, const double *dr, int precision, double *u,
                                   int *capped);

/* Colloid-rod energy of n rods on the GPU, as GBCollRod() (MCnptColloid.c:1758-1854);
 * rel = rod centre minus colloid centre (n*3), e = rod orientation (n*3). */
int mcgb_colloid_energy_batch(const mcgb_params *p, int n, const double *rel, const double *e,
                              int precision, double *u, int *capped);

/* Philox4x32-10 on the GPU: ctr n*4, key n*2, out n*4 (known-answer tests). */
int mcgb_philox_batch(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* Parity hook for the move proposal: copies the draws the device made for the last sweep of
 * replica r, in record order pass*(nact*k) + active_cell*k + move: j = raw particle selector
 * (the move picks cell member (j * n_c) >> 32), dq = fixed-point step (n*3; the reference's
 * d_max (U - 1/2) per axis, MCnptColloid.c:2502-2509), v = rotation vector on the unit sphere
 * (n*3; replaces the Marsaglia table of pto_ransphr/Marsaglia4MC, :3036-3078, read at
 * :2526-2550), thr = Metropolis threshold -log(U) T (:2566-2574).  Any array may be NULL.
 * info = {sweep index of the draws, k, active cells per pass, passes}.  Returns the number of
 * records (MCGB_ERR_CAPACITY when above capacity). */
int mcgb_sweep_draws(mcgb_ctx *ctx, int replica, int capacity, uint32_t *j, int32_t *dq,
                     double *v, double *thr, uint32_t info[4]);

/* ---------------------------------------------------------------- host utilities */

/* Seeded lattice start: n_s^3 sites, n_s the smallest side that leaves >= N sites outside the
 * exclusion sphere, at (i+1/2)*L/n_s plus uniform jitter in [-jitter, jitter]; N of the
 * remaining sites chosen by a seeded shuffle; orientations uniform on the sphere.  Box volume
 * V = N/rho + 4/3 pi R^3 (MCnptColloid.c:2843-2846), cubic.  Host only. */
int mcgb_init_lattice(int n, double density, double colloid_radius, double exclusion_radius,
                      double jitter, uint64_t seed,
                      double *x, double *y, double *z, double *ux, double *uy, double *uz,
                      double box[3]);

/* initial_configuration_type 1 (initialize_random_positions, MCnptColloid.c:2835-2920): random
 * sequential addition in the box V = N/rho + 4/3 pi R^3 (L = V^(1/3)), rejecting points within
 * sigma_ee of an earlier particle (minimum image) or inside the colloid; orientations uniform on
 * the sphere.  Philox(seed) replaces rand(); MCGB_ERR_GEOMETRY after max_tries rejections for
 * one particle (the reference loops forever).  Host only. */
int mcgb_init_random(int n, double density, double colloid_radius, double sigma_ee, uint64_t seed,
                     int max_tries, double *x, double *y, double *z, double *ux, double *uy,
                     double *uz, double box[3]);

/* initial_configuration_type 2 (initialize_cubic_lattice_positions, MCnptColloid.c:2924-3033):
 * n_s^3 sites (n_s >= 2, n_s^3 >= N) at (i, j, k) L/n_s filled x-fastest, sites within the
 * colloid skipped; orientations uniform on the sphere (Philox).  The spacing stays L/n_s (the
 * reference's reset to sigma_ee, defect D20, would put sites outside the box).  Host only. */
int mcgb_init_cubic(int n, double density, double colloid_radius, uint64_t seed,
                    double *x, double *y, double *z, double *ux, double *uy, double *uz,
                    double box[3]);

/* Trajectory writers, same text layouts as the reference's (append = 0 starts the file):
 * print_Trayectoria (MCnptColloid.c:3081-3107, "Configuraciones.dat"), print_Qmga (:3110-3157,
 * "qmga_mcnvtGB.cnf", positions relative to the box centre), print_pos_xyz (:3160-3180,
 * "DiscosEsfera.xyz": append = 0 writes the two header lines, append = 1 the positions).
 * Host only. */
int mcgb_write_trajectory(const char *path, int append, int n, const double *x, const double *y,
                          const double *z, const double *ux, const double *uy, const double *uz,
                          const double box[3]);
int mcgb_write_qmga(const char *path, int append, int n, const double *x, const double *y,
                    const double *z, const double *ux, const double *uy, const double *uz,
                    const double box[3]);
int mcgb_write_xyz(const char *path, int append, int n, const double *x, const double *y,
                   const double *z, const double box[3]);

/* Asynchronous sweeps and snapshots (no counterpart in the reference, whose output is
 * synchronous: print_Trayectoria / backup_restart_file at the save interval,
 * MCnptColloid.c:958-960).  mcgb_sweep_async enqueues n sweeps of every replica on the
 * context's stream and returns at once; mcgb_sync waits for them and reports what mcgb_sweep
 * would (a non-finite running energy).  Every other entry point waits for enqueued work first.
 * mcgb_snapshot_begin enqueues, after the work already enqueued, a device-side conversion of
 * replica r's configuration to FP64 in particle order and its copy into pinned host memory;
 * mcgb_snapshot_wait waits for that copy only (sweeps enqueued after the snapshot keep
 * running) and returns the same arrays as mcgb_get_state would have at that point.  One
 * snapshot per replica is in flight at a time; a second begin replaces the first. */
int mcgb_sweep_async(mcgb_ctx *ctx, int n_sweeps);
int mcgb_sync(mcgb_ctx *ctx);
int mcgb_snapshot_begin(mcgb_ctx *ctx, int replica);
int mcgb_snapshot_wait(mcgb_ctx *ctx, int replica, int n, double *x, double *y, double *z,
                       double *ux, double *uy, double *uz, double box[3]);

/* Step sizes of a replica (float, as the reference's max_displacement / max_rotation globals,
 * MCnptColloid.c:108-109).  With adapt_moves > 0 the device adapts them after every sweep that
 * brings the attempted moves since the last adaptation to >= adapt_moves, from the acceptance
 * of those moves: the controller of MCnptColloid.c:987-1031 (ratio < 0.30: d_max -= 5 %, at
 * d_max < 1e-4 min_thick reset to 0.01 min_thick and d_rot halved while > 1e-6; ratio > 0.55:
 * d_max += 5 %, above 1.5 min_thick reset to 0.75 min_thick and d_rot x 1.5 while < 0.7;
 * min_thick = min(sigma_ee, sigma_ss)).  *adaptations (may be NULL) counts the controller's
 * actions since mcgb_create.  mcgb_set_step_sizes overrides both and restarts the count of
 * moves towards the next adaptation; it requires 0 <= d_max < box/2 on every axis and
 * d_rot >= 0. */
int mcgb_get_step_sizes(mcgb_ctx *ctx, int replica, float *max_displacement, float *max_rotation,
                        uint64_t *adaptations);
int mcgb_set_step_sizes(mcgb_ctx *ctx, int replica, float max_displacement, float max_rotation);

/* Integrated autocorrelation time with Sokal's automatic window (smallest M >= c*tau(M)),
 * tau = 1 + 2 sum_{t=1..M} rho(t).  ESS = n / tau.  Host only. */
int mcgb_tau_int(const double *x, int n, double c, double *tau, int *window);

/* Eigen-decomposition of a symmetric 3x3 matrix q = {xx, yy, zz, xy, xz, yz}: eigenvalues in
 * ascending order, eigenvectors as columns of evec (row-major 3x3).  Host only. */
int mcgb_eigen_sym3(const double q[6], double evals[3], double evec[9]);

/* fortSphere.1 restart file (MCnptColloid.c:3193-3210 writer, :2786-2831 reader):
 * header "N Lx Ly Lz R", then per particle 4 lines: position, 0 0 0, orientation, 0 0 0. */
int mcgb_write_fort1(const char *path, int n, const double *x, const double *y, const double *z,
                     const double *ux, const double *uy, const double *uz,
                     const double box[3], double colloid_radius);
int mcgb_read_fort1(const char *path, int n, double *x, double *y, double *z,
                    double *ux, double *uy, double *uz, double box[3], double *colloid_radius);

/* Bit-reproducible resume (no counterpart in the reference, whose rand() stream cannot be
 * resumed; fortSphere.1 above stays the reference's format).  The record holds what a run
 * needs to continue bit for bit besides its state: the Philox key (seed), the sweep index (the
 * Philox counter of the next sweep's draws), the step sizes, the step-size adaptation counters
 * and the running energy (integer counts of capped pairs and colloid overlaps and the finite
 * part, which the NPT volume test compares its trial sum against).  mcgb_resume_state uploads a state verbatim (fixed-point positions; orientations as
 * stored, not renormalised) together with the record, whose seed must be the replica's.
 * mcgb_write_resume / mcgb_read_resume keep both in a text file (positions as integers,
 * floating-point values as exact hexadecimal literals). */
typedef struct {
    uint64_t seed;
    uint64_t sweep_index;
    float max_displacement, max_rotation;
    uint64_t adapt_attempted, adapt_accepted, adaptations;
    int64_t capped_pairs, capped_colloid;
    double energy_finite;
} mcgb_resume;
int mcgb_get_resume(mcgb_ctx *ctx, int replica, mcgb_resume *out);
int mcgb_resume_state(mcgb_ctx *ctx, int replica, int n, const uint32_t *qx, const uint32_t *qy,
                      const uint32_t *qz, const double *ux, const double *uy, const double *uz,
                      const double box[3], const mcgb_resume *record);
int mcgb_write_resume(const char *path, int n, const uint32_t *qx, const uint32_t *qy,
                      const uint32_t *qz, const double *ux, const double *uy, const double *uz,
                      const double box[3], const mcgb_resume *record);
int mcgb_read_resume(const char *path, int n, uint32_t *qx, uint32_t *qy, uint32_t *qz,
                     double *ux, double *uy, double *uz, double box[3], mcgb_resume *record);

/* ---------------------------------------------------------------- learned first stage
 *
 * No counterpart in the reference (its MC has no surrogate).  A per-pair MLP 4 -> 64 -> 64 -> 1
 * (SiLU) on the features (r, u_i.r^, u_j.r^, u_i.u_j) of GB pairs inside the cutoff, fitted to
 * the exact pair energy clipped at `clip`, trained on the GPU by the library's own forward,
 * backward and Adam kernels, and evaluated inside the sweep with BF16 matrix-core tiles as the
 * first stage of delayed acceptance (Christen & Fox 2005); the second stage is the exact GB
 * energy, so the chain keeps the exact Boltzmann distribution. */
#define MCGB_SUR_NPARAM 4545   /* W1[64][4] b1[64] W2[64][64] b2[64] W3[64] b3, row-major [out][in] */
#define MCGB_DA_BINS    66     /* histogram of (dE_surrogate - dE_exact)/kT: 64 bins over [-4, 4], under, over */

/* FP32 master parameters (host copies in and out). */
int mcgb_surrogate_set(mcgb_ctx *ctx, const float *params);
int mcgb_surrogate_get(mcgb_ctx *ctx, float *params);
/* Uniform(+-1/sqrt(fan_in)) initialisation from Philox(seed). */
int mcgb_surrogate_init(mcgb_ctx *ctx, uint32_t seed);
/* y[i] for n feature rows x[i] = (r, a, b, c): use_mfma = 1 the BF16 matrix-core path of the
 * sweep, 0 the FP32 path. */
int mcgb_surrogate_eval(mcgb_ctx *ctx, int n, const float *x, int use_mfma, float *y);
/* Gradient of the mean loss (1/n) sum (y - min(t, clip))^2 with respect to every parameter,
 * by the training kernels; params == NULL uses the context's parameters. */
int mcgb_surrogate_grad(mcgb_ctx *ctx, const float *params, int n, const float *x, const float *t,
                        float clip, float *grad, double *loss);
/* steps Adam steps (beta 0.9/0.999, eps 1e-8; learning rate cosine-decayed from lr to lr/20)
 * on batches of `batch` pairs drawn with replacement (Philox(seed)) from the emitted pairs, each
 * mapped by a random symmetry of the pair energy (head/tail flips, i <-> j).  wfloor > 0
 * weights each squared error by (wfloor + 1)/(wfloor + |min(t, clip)|) (0: unweighted);
 * losses[k] = mean (weighted) loss of batch k (or NULL). */
int mcgb_surrogate_train(mcgb_ctx *ctx, int steps, int batch, float lr, float clip, float wfloor,
                         uint32_t seed, float *losses);
/* Device storage of exact pair records (r, a, b, c, U; U = 1e30 for a capped pair). */
int mcgb_pairs_reserve(mcgb_ctx *ctx, long long capacity);
int mcgb_pairs_clear(mcgb_ctx *ctx);
int mcgb_pairs_count(mcgb_ctx *ctx, long long *n);
int mcgb_pairs_get(mcgb_ctx *ctx, long long first, int n, float *out);
/* Appends n host records (e.g. pairs emitted by another context). */
int mcgb_pairs_put(mcgb_ctx *ctx, int n, const float *rec);
/* Appends every in-cutoff ordered pair (i, j) of the list positions i = offset + m*stride of
 * replica's current state, with its exact FP64 GB energy (pij_GB, MCnptColloid.c:2618-2681). */
int mcgb_emit_pairs(mcgb_ctx *ctx, int replica, int stride, int offset, long long *added);
/* Delayed acceptance on/off for the following sweeps; judge = 1 also evaluates the exact
 * stage for moves the surrogate rejected (decisions unchanged; fills the histogram). */
int mcgb_da_enable(mcgb_ctx *ctx, int on, int judge);
int mcgb_da_histogram(mcgb_ctx *ctx, int replica, uint64_t counts[MCGB_DA_BINS], int clear);
/* The plain kernel as automatic fallback: with window_sweeps > 0, mcgb_sweep runs delayed
 * acceptance in windows of that many sweeps and, after a window in which the exact stage
 * rejected more than max_stage2_rejection of the moves the surrogate let through, or the
 * surrogate let through less than min_stage1_acceptance of the moves (a broken or poisoned
 * network: NaN weights reject everything), switches delayed acceptance off and runs the rest
 * with the plain kernel (counted in *fallbacks; mcgb_da_enable turns it on again).  The check
 * sums the counters of all replicas; it applies to mcgb_sweep, not to mcgb_sweep_async.
 * window_sweeps = 0 (the default) turns the guard off.  mcgb_da_status: whether delayed
 * acceptance is on, the fallbacks so far, and the last window's {stage-2 rejection, stage-1
 * acceptance}. */
int mcgb_da_guard(mcgb_ctx *ctx, int window_sweeps, double max_stage2_rejection,
                  double min_stage1_acceptance);
int mcgb_da_status(mcgb_ctx *ctx, int *on, uint64_t *fallbacks, double rates[2]);

#ifdef __cplusplus
}
#endif
#endif /* MCGB_H */
