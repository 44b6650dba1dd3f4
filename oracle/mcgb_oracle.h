/*
 * mcgb_oracle.h — FP64 CPU oracle of the Gay-Berne checkerboard Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference algorithm
 * (MCnptColloid.c) used as the checker: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (libmcgb.so) never links or calls it.
 *
 * Each function cites the reference lines it restates.  What pins each function (golden
 * vectors from the reference's own compiled physics, closed forms, brute force) is listed in
 * the header comment of oracle/mcgb_oracle.c.
 */
#ifndef MCGB_ORACLE_H
#define MCGB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters read by the hot path and the constants get_system_const() derives from them
 * (MCnptColloid.c:2752-2783), computed in dependency order (the reference computes chi from
 * kappa^2 before assigning kappa^2, :2772-2773). */
typedef struct orc_consts {
    double sigma0, epsilon0, kappa, kappa_prime, mu, nu;
    double temperature, pressure, cutoff;
    double max_displacement, max_rotation;
    double colloid_radius;
    int    colloid_on;                 /* colloid_wall_option == 1 */
    int    colors;                     /* sublattice colours per axis: 2 (default, 0 means 2)
                                          or 3 */
    /* derived */
    double beta, four_eps0, rc2, kappa2, chi, chi2, k_imu, chi_prime, sigma_ee, colloid_cut, colloid_R2;
} orc_consts;

/* Trace record of one trial move; same layout as mcgb_move_trace in include/mcgb.h. */
typedef struct orc_move_trace {
    int32_t pid, outcome, dcap, dccap;
    double  dfin, scale;
} orc_move_trace;

void   orc_consts_init(orc_consts *c);

/* pij_GB + g_omega (MCnptColloid.c:2607-2681) with iei, iej and e_i.e_j formed as in GBij
 * (:1524-1535).  Returns the reference value (1e9 when capped) and sets *capped. */
double orc_gb_pair(const orc_consts *c, const double ei[3], const double ej[3], const double dr[3],
                   int *capped);

/* GBCollRod (MCnptColloid.c:1758-1854).  rel = rod centre - colloid centre. */
double orc_colloid(const orc_consts *c, const double rel[3], const double e[3], int *capped);

/* Philox4x32-10 (Salmon et al., SC'11). */
void   orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Largest even n with L/n >= rc*(1+1e-6); 0 if that n is below 4. */
int    orc_cell_dim(double L, double rc);

/* Sweep grid for `colors` sublattice colours per axis: 2 -> orc_cell_dim; 3 -> the largest
 * multiple of 3 with L/n >= (rc/2)(1+1e-6), 0 if below 6 (cells of one colour are then two
 * cells, >= rc, apart, and a move's partners lie within +-2 cells). */
int    orc_grid_dim(double L, double rc, int colors);

/* Cell list of one configuration for grid shift g: cell_start[n_cells+1] and particle
 * indices sorted by (cell, index). */
void   orc_cell_list(int N, const uint32_t *qx, const uint32_t *qy, const uint32_t *qz,
                     const int n[3], const uint32_t g[3], int *cell_start, int *sorted);

/* Per-sweep set-up draws: grid shift, sublattice order (colors^3 entries), NPT draws
 * (vol[0]: d lnV, vol[1]: accept). */
void   orc_sweep_setup(uint64_t seed, uint32_t replica, uint32_t t, int colors, uint32_t g[3],
                       uint32_t order[27], uint32_t vol[2]);

/* One checkerboard sweep (grid shift, cell list, colors^3 passes, and for mc_npt != 0 one ln V
 * move).
 * energy = {cap_pairs, cap_colloid, finite} running total, updated in place.
 * counters: [0] attempted [1] accepted [2] left cell [3] hard core [4] candidates [5] pairs
 *           [6] colloid evals [7] volume attempted [8] volume accepted.
 * trace may be NULL; else it must hold n_cells*k records.  info (may be NULL) receives
 * {g[3], k, n_cells, passes, order[passes]} (36 words, the rest 0).  nthreads > 1 runs the independent cells of a pass with OpenMP
 * (results do not depend on it).  After the passes the running energy is replaced by an FP64
 * recompute when resum != 0 or when the accepted moves carried sum |finite dE| > 1e10.
 * Returns 0, or -1 on a geometry error. */
int    orc_sweep(const orc_consts *c, int N, uint32_t *qx, uint32_t *qy, uint32_t *qz,
                 double *ux, double *uy, double *uz, double L[3],
                 uint64_t seed, uint32_t replica, uint32_t t, int mc_npt, double npt_max_dlnv,
                 double energy[3], int64_t counters[9], orc_move_trace *trace, uint32_t info[36],
                 int resum, int nthreads);

/* Step-size controller of the reference's serial driver (MCnptColloid.c:987-1031), applied to
 * the acceptance accepted/attempted (a float quotient there) of the moves since its last
 * action.  The step sizes are floats (:108-109) updated by double expressions.  An increase
 * that would reach half_box is not made (the reference has no such bound). */
void   orc_adapt_steps(float *max_displacement, float *max_rotation, int64_t accepted,
                       int64_t attempted, float min_thick, double half_box);

/* FP64 total energy over the cell list: out = {cap_pairs, cap_colloid, finite, hard_core_count}.
 * Each unordered pair counted once (tot_ene, MCnptColloid.c:1598-1622, counts both orders and
 * halves). */
int    orc_total_energy(const orc_consts *c, int N, const uint32_t *qx, const uint32_t *qy,
                        const uint32_t *qz, const double *ux, const double *uy, const double *uz,
                        const double L[3], double out[4], int nthreads);

/* Same by brute force over all pairs with the reference's minimum image (small N only). */
int    orc_total_energy_brute(const orc_consts *c, int N, const double *x, const double *y,
                              const double *z, const double *ux, const double *uy, const double *uz,
                              const double L[3], double out[4]);

/* Order parameter: FP64 Q tensor, cyclic Jacobi; P2 = 1.5 lambda_max, director = its vector
 * (compute_liquid_crystal_order_parameters, MCnptColloid.c:2117-2218, with the true maximum). */
void   orc_order_parameter(int N, const double *ux, const double *uy, const double *uz,
                           double *p2, double director[3]);

/* Draws of moves m0 .. m0+n-1 of active cell `cell` in pass `pass` of sweep t, by the code the
 * sweep uses: j = raw particle selector (member (j * n_c) >> 32 of the cell), dq = fixed-point
 * step (n*3), v = rotation vector on the unit sphere (n*3), thr = -log(U) T.  Any output may be
 * NULL. */
void   orc_move_draws(uint64_t seed, uint32_t replica, uint32_t t, int pass, uint32_t cell, int m0,
                      int n, const double L[3], double max_displacement, double temperature,
                      uint32_t *j, int32_t *dq, double *v, double *thr);

/* orc_gb_pair / orc_colloid over n rows (vectors n*3 interleaved; capped may be NULL). */
void   orc_gb_pair_batch(const orc_consts *c, int n, const double *ei, const double *ej,
                         const double *dr, double *u, int *capped);
void   orc_colloid_batch(const orc_consts *c, int n, const double *rel, const double *e,
                         double *u, int *capped);

/* nchains independent NVT chains of N particles each (chain ch = replica ch of `seed`), run
 * for nsweeps sweeps from sweep index t0, OpenMP over the chains.  q (nchains*3*N, axis-major
 * per chain), u likewise, L (nchains*3) are updated in place.  After every `stride` sweeps the
 * state and the energy value (caps*1e9 + colloid caps*1e8 + finite) are recorded:
 * out_q/out_u [chain][record][3N], out_e [chain][record]; any may be NULL.  Returns 0 or -1. */
int    orc_run_chains(const orc_consts *c, int N, int nchains, uint32_t *q, double *u, double *L,
                      uint64_t seed, uint32_t t0, int nsweeps, int stride, uint32_t *out_q,
                      double *out_u, double *out_e, int nthreads);

/* Planted mistakes, used only by the tests that show the statistical checks of the sweep fail
 * on them (global, 0 = none):
 *   ORC_MUT_PHI_HALF         rotation azimuth drawn on [0, pi) instead of [0, 2 pi)
 *   ORC_MUT_Z_HALF           rotation z drawn on (0, 1) instead of (-1, 1)
 *   ORC_MUT_BETA             Metropolis threshold -log(U)/T instead of -log(U) T
 *   ORC_MUT_NO_TRIAL_COLLOID trial state's colloid term dropped from dE
 *   ORC_MUT_STEP_DRIFT       fixed-point step shifted by K/8 (an asymmetric proposal)
 *   ORC_MUT_DROP_OLD         old state's GB pair terms dropped from dE */
#define ORC_MUT_PHI_HALF          1
#define ORC_MUT_Z_HALF            2
#define ORC_MUT_BETA              3
#define ORC_MUT_NO_TRIAL_COLLOID  4
#define ORC_MUT_STEP_DRIFT        5
#define ORC_MUT_DROP_OLD          6
void   orc_set_mutation(int m);

#ifdef __cplusplus
}
#endif
#endif
