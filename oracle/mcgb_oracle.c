/*
 * mcgb_oracle.c — FP64 CPU oracle of the Gay-Berne checkerboard Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline; never part of libmcgb.so).
 *
 * What pins each function (tests in tests/test_oracle*.py):
 *   orc_consts_init   closed forms chi = (k^2-1)/(k^2+1), chi' = (k'^(1/mu)-1)/(k'^(1/mu)+1);
 *                     the reference's own get_system_const() called twice (its ordering bug
 *                     makes the first call give chi = -1) via oracle/_ref; golden vectors.
 *   orc_gb_pair       golden vectors from the reference's compiled pij_GB (oracle/_ref, fixtures
 *                     under tests/golden/); side-by-side / end-to-end / T / X closed forms; the
 *                     kappa = kappa' = 1 Lennard-Jones limit; i<->j and head-tail symmetry; the cap.
 *   orc_colloid       golden vectors from the reference's compiled GBCollRod; theta = 0 and 90
 *                     degree closed forms; cutoff and caps.
 *   orc_philox        Random123 known-answer vectors.
 *   orc_cell_list     brute force (every particle lies in its cell, sorted by index, counts).
 *   orc_sweep         per-move dE against an all-pairs brute-force recomputation with the
 *                     reference's minimum image; running energy against orc_total_energy_brute;
 *                     exact two-particle Boltzmann statistics; results independent of nthreads.
 *   orc_total_energy  orc_total_energy_brute (reference-style O(N^2) all pairs).
 *   orc_order_parameter  numpy.linalg.eigh on the same Q tensor; P2 = 1 for parallel rods.
 *
 * Sweep semantics (the reference's MC body is a TODO, MCnptColloid.c:683-700, so they are
 * defined here and mirrored exactly by the GPU kernels):
 *   key = {lo32(seed), hi32(seed) + replica}; sweep index t.
 *   set-up draws S_j = philox({j, 0, t, 0xFFFF0000}, key), j = 0, 1, 2:
 *     grid shift g = S_0[0..2] (fixed-point), NPT d lnV draw S_0[3], NPT accept draw S_2[3];
 *     sublattice order = Fisher-Yates of 0..P-1, i = P-1..1 using S_1[0..3], S_2[0..2], then
 *     S_3..S_7 in turn, j = (r*(i+1)) >> 32.
 *   positions are 32-bit fixed-point fractions of the box: x = q*L/2^32; minimum image is the
 *     wrap-around int32 difference; the cell of q is ((q - g)*n) >> 32 per axis.
 *   c colours per axis (consts.colors), P = c^3 sublattices:
 *     c = 2: n the largest even count with L/n >= r_c (n >= 4); sublattice (cx&1) + 2(cy&1) +
 *            4(cz&1) as in cell_coords_to_index (MCnptColloid.c:3692-3695); stencil +-1 cell;
 *     c = 3: n the largest multiple of 3 with L/n >= r_c/2 (n >= 6); sublattice
 *            (cx%3) + 3(cy%3) + 9(cz%3); stencil +-2 cells.
 *     Cells of one sublattice are separated by c-1 cells (>= r_c) and moves are confined to
 *     their cell, so the cells of a pass are independent.
 *   pass p = 0..P-1 activates the cells of sublattice order[p]; every active cell makes
 *     k = ceil(N/n_cells) moves, move m drawing A = philox({m, cell, t, p<<8}, key) and
 *     B = philox({m, cell, t, (p<<8)|1}, key):
 *       particle  = sorted cell member (A0*n_c) >> 32
 *       step      = trunc((2 A_d - 2^32 + 1) * K_d / 2^33), K_d = round(d_max 2^32 / L_d)
 *                   (uniform in (-d_max/2, d_max/2), MCnptColloid.c:2502-2509)
 *       rejected if the trial position leaves the active cell, or (colloid on) lies within R of
 *       the box centre (:2516-2520)
 *       rotation  = normalise(u + d_rot v), v = (s cos phi, s sin phi, z), z = 1 - 2 u(B0),
 *                   phi = 2 pi u(B1), s = sqrt(1 - z^2); u(r) = (r + 1/2) 2^-32 (:2526-2550)
 *       accept    iff dE <= -log(u(B2)) * T   (log U <= -beta dE, :2566-2574)
 *   dE = (change of capped GB pairs)*1e9 + (change of capped colloid terms)*1e8 + finite part.
 *   after the passes the running energy is replaced by an FP64 recompute when requested or when
 *   the accepted moves of the sweep carried sum |finite dE| > 1e10; then (NPT) the ln V move.
 */
#include "mcgb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO32 4294967296.0
#define MCGB_RESUM_THRESHOLD 1.0e10

/* ------------------------------------------------------------------ constants */

void orc_consts_init(orc_consts *c)
{
    /* the move sizes are float globals in the reference (MCnptColloid.c:108-109, read with %f) */
    c->max_displacement = (double)(float)c->max_displacement;
    c->max_rotation     = (double)(float)c->max_rotation;
    /* get_system_const, MCnptColloid.c:2752-2783, in dependency order */
    c->beta        = 1.0 / c->temperature;
    c->four_eps0   = 4.0 * c->epsilon0;
    c->rc2         = c->cutoff * c->cutoff;
    c->k_imu       = pow(c->kappa_prime, 1.0 / c->mu);
    c->kappa2      = c->kappa * c->kappa;
    c->chi         = (c->kappa2 - 1.0) / (c->kappa2 + 1.0);
    c->chi2        = c->chi * c->chi;
    c->chi_prime   = (c->k_imu - 1.0) / (c->k_imu + 1.0);
    c->sigma_ee    = c->sigma0 * c->kappa;            /* :4130 */
    c->colloid_cut = 4.0 * c->sigma_ee;               /* :2777 */
    c->colloid_R2  = c->colloid_radius * c->colloid_radius;
}

/* ------------------------------------------------------------------ pair physics */

static inline double dot3(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* g_omega, MCnptColloid.c:2607-2615 */
static inline double g_omega(double eiej, double amasb2, double amenb2, double omega)
{
    double oe = omega * eiej;
    return 1.0 - 0.5 * omega * (amasb2 / (1.0 + oe) + amenb2 / (1.0 - oe));
}

double orc_gb_pair(const orc_consts *c, const double ei[3], const double ej[3], const double dr[3],
                   int *capped)
{
    /* GBij, MCnptColloid.c:1524-1535: r, 1/r, inverse norms, cosine */
    double r2   = dot3(dr, dr);
    double rij  = sqrt(r2);
    double irij = 1.0 / rij;
    double iei  = 1.0 / sqrt(dot3(ei, ei));
    double iej  = 1.0 / sqrt(dot3(ej, ej));
    double eiej = iei * iej * dot3(ei, ej);

    /* pij_GB, :2656-2680 */
    double a  = irij * iei * dot3(dr, ei);
    double b  = irij * iej * dot3(dr, ej);
    double ap = a + b, am = a - b;
    double ap2 = ap * ap, am2 = am * am;
    double eps1  = 1.0 / sqrt(1.0 - c->chi2 * eiej * eiej);
    double ctep1 = pow(eps1, c->nu);
    double eps2  = g_omega(eiej, ap2, am2, c->chi_prime);
    double ctep2 = pow(eps2, c->mu);
    double schi  = g_omega(eiej, ap2, am2, c->chi);
    double sigma = c->sigma0 / sqrt(schi);
    double distreal = rij - sigma + c->sigma0;
    if (distreal < 0.001) {            /* :2670 */
        *capped = 1;
        return 1.0e9;
    }
    *capped = 0;
    double erre  = c->sigma0 / distreal;
    double erre2 = erre * erre;
    double erre6 = erre2 * erre2 * erre2;
    return c->four_eps0 * (ctep1 * ctep2) * erre6 * (erre6 - 1.0);
}

double orc_colloid(const orc_consts *c, const double rel[3], const double e[3], int *capped)
{
    /* GBCollRod, MCnptColloid.c:1758-1854 */
    double d = sqrt(dot3(rel, rel));
    double dij = d - c->colloid_radius;
    *capped = 0;
    if (d <= 1e-12) { *capped = 1; return 1.0e8; }
    if (dij > c->colloid_cut) return 0.0;
    double cth = (rel[0] / d) * e[0] + (rel[1] / d) * e[1] + (rel[2] / d) * e[2];
    if (cth > 1.0) cth = 1.0;
    if (cth < -1.0) cth = -1.0;
    double sin2 = 1.0 - cth * cth;
    double s0 = c->sigma0;
    double den = 1.0 - c->chi, num = 1.0 - c->chi * sin2;
    if (den <= 0.0) den = 1e-12;
    if (num <= 1e-12) num = 1e-12;
    double sigma_cr = s0 * sqrt(num / den);
    double den_e = 1.0 - c->chi_prime * sin2, num_e = 1.0 - c->chi_prime;
    if (den_e <= 1e-12) den_e = 1e-12;
    double eps_cr = c->kappa_prime * pow(num_e / den_e, c->mu);
    double x = dij - sigma_cr + s0;
    if (x <= 1e-10) { *capped = 1; return 1.0e8; }
    double s3 = s0 * s0 * s0, s9 = s3 * s3 * s3;
    double x2 = x * x, x3 = x2 * x, x8 = x3 * x3 * x2, x9 = x8 * x;
    return eps_cr * (s9 / (45.0 * x9) - s3 / (6.0 * x3) - s9 / (40.0 * d * x8) + s3 / (4.0 * d * x2));
}

/* ------------------------------------------------------------------ Philox4x32-10 */

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void make_key(uint64_t seed, uint32_t replica, uint32_t key[2])
{
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32) + replica;
}

static inline double u01(uint32_t r) { return ((double)r + 0.5) * (1.0 / TWO32); }

/* ------------------------------------------------------------------ cells */

int orc_cell_dim(double L, double rc)
{
    int n = 2 * (int)floor(L / (2.0 * rc * (1.0 + 1e-6)));
    return n >= 4 ? n : 0;
}

int orc_grid_dim(double L, double rc, int colors)
{
    if (colors != 3) return orc_cell_dim(L, rc);
    int n = 3 * (int)floor(L / (1.5 * rc * (1.0 + 1e-6)));
    return n >= 6 ? n : 0;
}

static inline uint32_t cell_coord(uint32_t q, uint32_t g, uint32_t n)
{
    return (uint32_t)(((uint64_t)(uint32_t)(q - g) * n) >> 32);
}

void orc_cell_list(int N, const uint32_t *qx, const uint32_t *qy, const uint32_t *qz,
                   const int n[3], const uint32_t g[3], int *cell_start, int *sorted)
{
    int nc = n[0] * n[1] * n[2];
    int *cid = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    memset(cell_start, 0, sizeof(int) * (size_t)(nc + 1));
    for (int i = 0; i < N; ++i) {
        uint32_t cx = cell_coord(qx[i], g[0], (uint32_t)n[0]);
        uint32_t cy = cell_coord(qy[i], g[1], (uint32_t)n[1]);
        uint32_t cz = cell_coord(qz[i], g[2], (uint32_t)n[2]);
        cid[i] = (int)(cx + (uint32_t)n[0] * (cy + (uint32_t)n[1] * cz));
        cell_start[cid[i] + 1]++;
    }
    for (int c = 0; c < nc; ++c) cell_start[c + 1] += cell_start[c];
    int *fill = (int *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(int));
    for (int i = 0; i < N; ++i) {          /* index order inside each cell: a stable counting sort */
        int c = cid[i];
        sorted[cell_start[c] + fill[c]++] = i;
    }
    free(fill);
    free(cid);
}

void orc_sweep_setup(uint64_t seed, uint32_t replica, uint32_t t, int colors, uint32_t g[3],
                     uint32_t order[27], uint32_t vol[2])
{
    uint32_t key[2], s[8][4];
    make_key(seed, replica, key);
    for (uint32_t j = 0; j < 8; ++j) {
        uint32_t cj[4] = {j, 0u, t, 0xFFFF0000u};
        orc_philox(cj, key, s[j]);
    }
    g[0] = s[0][0]; g[1] = s[0][1]; g[2] = s[0][2];
    vol[0] = s[0][3]; vol[1] = s[2][3];
    /* shuffle_int_array (MCnptColloid.c:3721-3737) with Philox draws: S_1[0..3], S_2[0..2],
     * then (27 colours) S_3..S_7 in order, S_7[3] unused */
    uint32_t draws[26];
    int nd = 0;
    for (int w = 0; w < 4; ++w) draws[nd++] = s[1][w];
    for (int w = 0; w < 3; ++w) draws[nd++] = s[2][w];
    for (int j = 3; j < 8; ++j)
        for (int w = 0; w < 4 && nd < 26; ++w) draws[nd++] = s[j][w];
    const int P = colors == 3 ? 27 : 8;
    for (int i = 0; i < P; ++i) order[i] = (uint32_t)i;
    for (int i = P - 1, d = 0; i >= 1; --i, ++d) {
        uint32_t j = (uint32_t)(((uint64_t)draws[d] * (uint64_t)(i + 1)) >> 32);
        uint32_t tmp = order[i]; order[i] = order[j]; order[j] = tmp;
    }
}

/* ------------------------------------------------------------------ moves */

static inline int32_t step_q(uint32_t r, int64_t K)
{
    int64_t m = 2 * (int64_t)r - 4294967295LL;        /* odd, symmetric in [-(2^32-1), 2^32-1] */
    int64_t p = m * K;
    return (int32_t)(p >= 0 ? (p >> 33) : -((-p) >> 33));
}

/* Planted mistakes for the tests that show the statistical checks detect them (0: none). */
static int g_mutation = 0;

void orc_set_mutation(int m) { g_mutation = m; }

/* The draws of move m of active cell `cell` in pass p of sweep t: particle selector, fixed-point
 * step (MCnptColloid.c:2502-2509), rotation vector on the unit sphere (the distribution of
 * pto_ransphr/Marsaglia4MC, :3036-3078, used at :2526-2550) and Metropolis threshold
 * -log(U) T (:2566-2574). */
typedef struct {
    uint32_t j;
    int32_t dq[3];
    double v[3];
    double thr;
} orc_draw;

static void move_draw(const uint32_t key[2], uint32_t t, int p, uint32_t cell, int m,
                      const int64_t K[3], double temperature, orc_draw *d)
{
    uint32_t ca[4] = {(uint32_t)m, cell, t, ((uint32_t)p << 8)};
    uint32_t cb[4] = {(uint32_t)m, cell, t, ((uint32_t)p << 8) | 1u};
    uint32_t A[4], B[4];
    orc_philox(ca, key, A);
    orc_philox(cb, key, B);
    d->j = A[0];
    for (int k = 0; k < 3; ++k) {
        d->dq[k] = step_q(A[1 + k], K[k]);
        if (g_mutation == ORC_MUT_STEP_DRIFT) d->dq[k] += (int32_t)(K[k] / 8);
    }
    double zz = 1.0 - 2.0 * u01(B[0]);
    if (g_mutation == ORC_MUT_Z_HALF) zz = 1.0 - u01(B[0]);
    double phi = 2.0 * M_PI * u01(B[1]);
    if (g_mutation == ORC_MUT_PHI_HALF) phi = M_PI * u01(B[1]);
    double sxy = sqrt(fmax(0.0, 1.0 - zz * zz));
    d->v[0] = sxy * cos(phi);
    d->v[1] = sxy * sin(phi);
    d->v[2] = zz;
    d->thr = g_mutation == ORC_MUT_BETA ? -log(u01(B[2])) / temperature
                                        : -log(u01(B[2])) * temperature;
}

static void step_amplitudes(double max_displacement, const double L[3], int64_t K[3])
{
    for (int d = 0; d < 3; ++d) K[d] = llround(max_displacement * TWO32 / L[d]);
}

void orc_move_draws(uint64_t seed, uint32_t replica, uint32_t t, int pass, uint32_t cell, int m0,
                    int n, const double L[3], double max_displacement, double temperature,
                    uint32_t *j, int32_t *dq, double *v, double *thr)
{
    uint32_t key[2];
    int64_t K[3];
    make_key(seed, replica, key);
    step_amplitudes((double)(float)max_displacement, L, K);
    for (int k = 0; k < n; ++k) {
        orc_draw d;
        move_draw(key, t, pass, cell, m0 + k, K, temperature, &d);
        if (j) j[k] = d.j;
        for (int a = 0; a < 3; ++a) {
            if (dq) dq[3 * (size_t)k + a] = d.dq[a];
            if (v) v[3 * (size_t)k + a] = d.v[a];
        }
        if (thr) thr[k] = d.thr;
    }
}

typedef struct {
    const orc_consts *c;
    const uint32_t *qx, *qy, *qz;
    const double *ux, *uy, *uz;
    double s[3];
    int n[3];
    const int *cell_start, *sorted;
} energy_view;

/* finite energy difference and cap changes of moving particle i from (q, u) to (qn, un),
 * summed over the (2R+1)^3-cell stencil of cell (cx, cy, cz). */
static void move_delta(const energy_view *v, int i, const uint32_t qn[3], const double un[3],
                       int cx, int cy, int cz, int R, int *dcap, double *dfin, double *scale,
                       int64_t *ncand, int64_t *npair)
{
    const orc_consts *c = v->c;
    const double ui[3] = {v->ux[i], v->uy[i], v->uz[i]};
    int capo_n = 0, capn_n = 0;
    double acc = 0.0, sc = 0.0;
    for (int dz = -R; dz <= R; ++dz)
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                int ex = (cx + dx + v->n[0]) % v->n[0];
                int ey = (cy + dy + v->n[1]) % v->n[1];
                int ez = (cz + dz + v->n[2]) % v->n[2];
                int cell = ex + v->n[0] * (ey + v->n[1] * ez);
                for (int s = v->cell_start[cell]; s < v->cell_start[cell + 1]; ++s) {
                    int j = v->sorted[s];
                    if (j == i) continue;
                    const double uj[3] = {v->ux[j], v->uy[j], v->uz[j]};
                    double dro[3] = {(double)(int32_t)(v->qx[j] - v->qx[i]) * v->s[0],
                                     (double)(int32_t)(v->qy[j] - v->qy[i]) * v->s[1],
                                     (double)(int32_t)(v->qz[j] - v->qz[i]) * v->s[2]};
                    double drn[3] = {(double)(int32_t)(v->qx[j] - qn[0]) * v->s[0],
                                     (double)(int32_t)(v->qy[j] - qn[1]) * v->s[1],
                                     (double)(int32_t)(v->qz[j] - qn[2]) * v->s[2]};
                    *ncand += 2;
                    double uo = 0.0, unw = 0.0;
                    int cap;
                    if (dot3(dro, dro) < c->rc2) {           /* strict <, MCnptColloid.c:1682 */
                        double e = orc_gb_pair(c, ui, uj, dro, &cap);
                        *npair += 1;
                        if (cap) capo_n++; else uo = e;
                    }
                    if (dot3(drn, drn) < c->rc2) {
                        double e = orc_gb_pair(c, un, uj, drn, &cap);
                        *npair += 1;
                        if (cap) capn_n++; else unw = e;
                    }
                    acc += g_mutation == ORC_MUT_DROP_OLD ? unw : unw - uo;
                    sc += fabs(unw) + fabs(uo);
                }
            }
    *dcap = capn_n - capo_n;
    *dfin = acc;
    *scale = sc;
}

int orc_total_energy(const orc_consts *c, int N, const uint32_t *qx, const uint32_t *qy,
                     const uint32_t *qz, const double *ux, const double *uy, const double *uz,
                     const double L[3], double out[4], int nthreads)
{
    int n[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = orc_cell_dim(L[d], c->cutoff);
        if (n[d] == 0) return -1;
    }
    int nc = n[0] * n[1] * n[2];
    int *cs = (int *)malloc(sizeof(int) * (size_t)(nc + 1));
    int *so = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    uint32_t g0[3] = {0u, 0u, 0u};
    orc_cell_list(N, qx, qy, qz, n, g0, cs, so);
    double s[3] = {L[0] / TWO32, L[1] / TWO32, L[2] / TWO32};
    double *fin = (double *)calloc((size_t)(N > 0 ? N : 1), sizeof(double));
    int64_t *cap = (int64_t *)calloc((size_t)(N > 0 ? N : 1), sizeof(int64_t));
    int64_t *ccap = (int64_t *)calloc((size_t)(N > 0 ? N : 1), sizeof(int64_t));
    int64_t *hard = (int64_t *)calloc((size_t)(N > 0 ? N : 1), sizeof(int64_t));
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < N; ++i) {
        uint32_t cx = cell_coord(qx[i], 0u, (uint32_t)n[0]);
        uint32_t cy = cell_coord(qy[i], 0u, (uint32_t)n[1]);
        uint32_t cz = cell_coord(qz[i], 0u, (uint32_t)n[2]);
        const double ui[3] = {ux[i], uy[i], uz[i]};
        double f = 0.0;
        int64_t cp = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int ex = ((int)cx + dx + n[0]) % n[0];
                    int ey = ((int)cy + dy + n[1]) % n[1];
                    int ez = ((int)cz + dz + n[2]) % n[2];
                    int cell = ex + n[0] * (ey + n[1] * ez);
                    for (int k = cs[cell]; k < cs[cell + 1]; ++k) {
                        int j = so[k];
                        if (j <= i) continue;
                        double dr[3] = {(double)(int32_t)(qx[j] - qx[i]) * s[0],
                                        (double)(int32_t)(qy[j] - qy[i]) * s[1],
                                        (double)(int32_t)(qz[j] - qz[i]) * s[2]};
                        if (dot3(dr, dr) < c->rc2) {
                            const double uj[3] = {ux[j], uy[j], uz[j]};
                            int capped;
                            double e = orc_gb_pair(c, ui, uj, dr, &capped);
                            if (capped) cp++; else f += e;
                        }
                    }
                }
        if (c->colloid_on) {
            double rel[3] = {(double)(int32_t)(qx[i] - 0x80000000u) * s[0],
                             (double)(int32_t)(qy[i] - 0x80000000u) * s[1],
                             (double)(int32_t)(qz[i] - 0x80000000u) * s[2]};
            int capped;
            double e = orc_colloid(c, rel, ui, &capped);
            if (capped) ccap[i] = 1; else f += e;
            if (dot3(rel, rel) <= c->colloid_R2) hard[i] = 1;
        }
        fin[i] = f;
        cap[i] = cp;
    }
    double F = 0.0;
    int64_t CP = 0, CC = 0, H = 0;
    for (int i = 0; i < N; ++i) { F += fin[i]; CP += cap[i]; CC += ccap[i]; H += hard[i]; }
    out[0] = (double)CP; out[1] = (double)CC; out[2] = F; out[3] = (double)H;
    free(fin); free(cap); free(ccap); free(hard); free(cs); free(so);
    return 0;
}

/* reference-style minimum image (compute_distance_sq, MCnptColloid.c:2690-2721) */
static inline double min_image_ref(double d, double L)
{
    if (d > 0.5 * L) d -= L;
    else if (d < -0.5 * L) d += L;
    return d;
}

int orc_total_energy_brute(const orc_consts *c, int N, const double *x, const double *y,
                           const double *z, const double *ux, const double *uy, const double *uz,
                           const double L[3], double out[4])
{
    double F = 0.0;
    int64_t CP = 0, CC = 0, H = 0;
    for (int i = 0; i < N; ++i) {
        const double ui[3] = {ux[i], uy[i], uz[i]};
        for (int j = i + 1; j < N; ++j) {
            double dr[3] = {min_image_ref(x[j] - x[i], L[0]), min_image_ref(y[j] - y[i], L[1]),
                            min_image_ref(z[j] - z[i], L[2])};
            if (dot3(dr, dr) < c->rc2) {
                const double uj[3] = {ux[j], uy[j], uz[j]};
                int capped;
                double e = orc_gb_pair(c, ui, uj, dr, &capped);
                if (capped) CP++; else F += e;
            }
        }
        if (c->colloid_on) {
            double rel[3] = {x[i] - 0.5 * L[0], y[i] - 0.5 * L[1], z[i] - 0.5 * L[2]};
            int capped;
            double e = orc_colloid(c, rel, ui, &capped);
            if (capped) CC++; else F += e;
            if (dot3(rel, rel) <= c->colloid_R2) H++;
        }
    }
    out[0] = (double)CP; out[1] = (double)CC; out[2] = F; out[3] = (double)H;
    return 0;
}

/* ------------------------------------------------------------------ sweep */

int orc_sweep(const orc_consts *c, int N, uint32_t *qx, uint32_t *qy, uint32_t *qz,
              double *ux, double *uy, double *uz, double L[3],
              uint64_t seed, uint32_t replica, uint32_t t, int mc_npt, double npt_max_dlnv,
              double energy[3], int64_t counters[9], orc_move_trace *trace, uint32_t info[36],
              int resum, int nthreads)
{
    const int col = c->colors == 3 ? 3 : 2;        /* sublattice colours per axis */
    const int P = col * col * col;                 /* passes per sweep */
    const int R = col - 1;                         /* stencil reach in cells */
    int n[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = orc_grid_dim(L[d], c->cutoff, col);
        if (n[d] == 0) return -1;
    }
    const int nc = n[0] * n[1] * n[2];
    const int hx = n[0] / col, hy = n[1] / col, hz = n[2] / col;
    const int nact = hx * hy * hz;
    const int k = (N + nc - 1) / nc;
    uint32_t g[3], order[27], vol[2], key[2];
    orc_sweep_setup(seed, replica, t, col, g, order, vol);
    make_key(seed, replica, key);
    if (info) {
        memset(info, 0, 36 * sizeof(uint32_t));
        for (int d = 0; d < 3; ++d) info[d] = g[d];
        info[3] = (uint32_t)k;
        info[4] = (uint32_t)nc;
        info[5] = (uint32_t)P;
        for (int p = 0; p < P; ++p) info[6 + p] = order[p];
    }
    int *cs = (int *)malloc(sizeof(int) * (size_t)(nc + 1));
    int *so = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    orc_cell_list(N, qx, qy, qz, n, g, cs, so);

    energy_view v;
    v.c = c; v.qx = qx; v.qy = qy; v.qz = qz; v.ux = ux; v.uy = uy; v.uz = uz;
    for (int d = 0; d < 3; ++d) { v.s[d] = L[d] / TWO32; v.n[d] = n[d]; }
    v.cell_start = cs; v.sorted = so;
    int64_t K[3];
    step_amplitudes(c->max_displacement, L, K);

    /* per-active-cell accumulators: deterministic totals whatever the thread count */
    int64_t *acc_i = (int64_t *)calloc((size_t)nact * 9, sizeof(int64_t));
    double *acc_f = (double *)calloc((size_t)nact * 2, sizeof(double));
    double moved_abs = 0.0;     /* sum of |finite dE| of accepted moves in this sweep */

    for (int p = 0; p < P; ++p) {
        const uint32_t sig = order[p];
        const int sx = (int)(sig % (uint32_t)col), sy = (int)((sig / (uint32_t)col) % (uint32_t)col),
                  sz = (int)(sig / (uint32_t)(col * col));
        memset(acc_i, 0, sizeof(int64_t) * (size_t)nact * 9);
        memset(acc_f, 0, sizeof(double) * (size_t)nact * 2);
        (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
        for (int a = 0; a < nact; ++a) {
            int ax = a % hx, ay = (a / hx) % hy, az = a / (hx * hy);
            int cx = col * ax + sx, cy = col * ay + sy, cz = col * az + sz;
            int cell = cx + n[0] * (cy + n[1] * cz);
            int c0 = cs[cell], ncnt = cs[cell + 1] - cs[cell];
            int64_t *ai = acc_i + (size_t)a * 9;
            for (int m = 0; m < k; ++m) {
                orc_move_trace rec = {-1, 0, 0, 0, 0.0, 0.0};
                if (ncnt == 0) {
                    if (trace) trace[(size_t)p * nact * k + (size_t)a * k + m] = rec;
                    continue;
                }
                orc_draw dr;
                move_draw(key, t, p, (uint32_t)cell, m, K, c->temperature, &dr);
                int jsel = (int)(((uint64_t)dr.j * (uint64_t)ncnt) >> 32);
                int i = so[c0 + jsel];
                rec.pid = i;
                ai[0]++;
                uint32_t qn[3] = {qx[i] + (uint32_t)dr.dq[0], qy[i] + (uint32_t)dr.dq[1],
                                  qz[i] + (uint32_t)dr.dq[2]};
                if (cell_coord(qn[0], g[0], (uint32_t)n[0]) != (uint32_t)cx ||
                    cell_coord(qn[1], g[1], (uint32_t)n[1]) != (uint32_t)cy ||
                    cell_coord(qn[2], g[2], (uint32_t)n[2]) != (uint32_t)cz) {
                    rec.outcome = 1;
                    ai[2]++;
                    if (trace) trace[(size_t)p * nact * k + (size_t)a * k + m] = rec;
                    continue;
                }
                double reln[3] = {(double)(int32_t)(qn[0] - 0x80000000u) * v.s[0],
                                  (double)(int32_t)(qn[1] - 0x80000000u) * v.s[1],
                                  (double)(int32_t)(qn[2] - 0x80000000u) * v.s[2]};
                if (c->colloid_on && dot3(reln, reln) <= c->colloid_R2) {
                    rec.outcome = 2;
                    ai[3]++;
                    if (trace) trace[(size_t)p * nact * k + (size_t)a * k + m] = rec;
                    continue;
                }
                double un[3] = {ux[i] + c->max_rotation * dr.v[0],
                                uy[i] + c->max_rotation * dr.v[1],
                                uz[i] + c->max_rotation * dr.v[2]};
                double n2 = dot3(un, un);
                if (n2 <= 1e-12) { un[0] = ux[i]; un[1] = uy[i]; un[2] = uz[i]; }
                else {
                    double nr = sqrt(n2);
                    un[0] /= nr; un[1] /= nr; un[2] /= nr;
                }
                int dcap = 0, dccap = 0;
                double dfin = 0.0, scale = 0.0;
                int64_t ncand = 0, npair = 0;
                move_delta(&v, i, qn, un, cx, cy, cz, R, &dcap, &dfin, &scale, &ncand, &npair);
                ai[4] += ncand;
                ai[5] += npair;
                if (c->colloid_on) {
                    double relo[3] = {(double)(int32_t)(qx[i] - 0x80000000u) * v.s[0],
                                      (double)(int32_t)(qy[i] - 0x80000000u) * v.s[1],
                                      (double)(int32_t)(qz[i] - 0x80000000u) * v.s[2]};
                    const double uo[3] = {ux[i], uy[i], uz[i]};
                    int co, cn;
                    double eo = orc_colloid(c, relo, uo, &co);
                    double en = orc_colloid(c, reln, un, &cn);
                    if (g_mutation == ORC_MUT_NO_TRIAL_COLLOID) { en = 0.0; cn = 0; }
                    dccap = cn - co;
                    dfin += (cn ? 0.0 : en) - (co ? 0.0 : eo);
                    scale += fabs(cn ? 0.0 : en) + fabs(co ? 0.0 : eo);
                    ai[6] += 2;
                }
                double dE = (double)dcap * 1.0e9 + (double)dccap * 1.0e8 + dfin;
                double thr = dr.thr;
                rec.dcap = dcap; rec.dccap = dccap; rec.dfin = dfin; rec.scale = scale;
                if (dE <= thr) {
                    qx[i] = qn[0]; qy[i] = qn[1]; qz[i] = qn[2];
                    ux[i] = un[0]; uy[i] = un[1]; uz[i] = un[2];
                    ai[1]++;
                    ai[7] += dcap;
                    ai[8] += dccap;
                    acc_f[2 * (size_t)a] += dfin;
                    acc_f[2 * (size_t)a + 1] += fabs(dfin);
                    rec.outcome = 4;
                } else {
                    rec.outcome = 3;
                }
                if (trace) trace[(size_t)p * nact * k + (size_t)a * k + m] = rec;
            }
        }
        for (int a = 0; a < nact; ++a) {
            const int64_t *ai = acc_i + (size_t)a * 9;
            for (int q = 0; q < 7; ++q) counters[q] += ai[q];
            energy[0] += (double)ai[7];
            energy[1] += (double)ai[8];
            energy[2] += acc_f[2 * (size_t)a];
            moved_abs += acc_f[2 * (size_t)a + 1];
        }
    }
    free(acc_i); free(acc_f); free(cs); free(so);

    /* FP64 re-summation: on request (periodic drift bound) or when this sweep moved energies so
     * large (near-cap pairs reach 1e36) that the running sum has lost its precision */
    if (resum || moved_abs > MCGB_RESUM_THRESHOLD) {
        double Ef[4];
        orc_total_energy(c, N, qx, qy, qz, ux, uy, uz, L, Ef, nthreads);
        energy[0] = Ef[0]; energy[1] = Ef[1]; energy[2] = Ef[2];
    }

    if (mc_npt) {
        /* isotropic ln V move with an FP64 full recompute (absent from the reference, whose
         * NPT is a label only: MCnptColloid.c:1197, :1291) */
        counters[7]++;
        double dlnv = npt_max_dlnv * (2.0 * u01(vol[0]) - 1.0);
        double f = exp(dlnv / 3.0);
        double Ln[3] = {L[0] * f, L[1] * f, L[2] * f};
        int ok = 1;
        for (int d = 0; d < 3; ++d)
            if (orc_cell_dim(Ln[d], c->cutoff) == 0) ok = 0;
        if (ok) {
            double En[4];
            orc_total_energy(c, N, qx, qy, qz, ux, uy, uz, Ln, En, nthreads);
            if (En[3] == 0.0) {
                double V = L[0] * L[1] * L[2], Vn = Ln[0] * Ln[1] * Ln[2];
                double dU = (En[0] - energy[0]) * 1.0e9 + (En[1] - energy[1]) * 1.0e8 + (En[2] - energy[2]);
                double arg = -c->beta * (dU + c->pressure * (Vn - V)) + (double)(N + 1) * dlnv;
                if (log(u01(vol[1])) <= arg) {
                    counters[8]++;
                    for (int d = 0; d < 3; ++d) L[d] = Ln[d];
                    energy[0] = En[0]; energy[1] = En[1]; energy[2] = En[2];
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ order parameter */

void orc_order_parameter(int N, const double *ux, const double *uy, const double *uz,
                         double *p2, double director[3])
{
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < N; ++i) {
        s[0] += ux[i] * ux[i]; s[1] += uy[i] * uy[i]; s[2] += uz[i] * uz[i];
        s[3] += ux[i] * uy[i]; s[4] += ux[i] * uz[i]; s[5] += uy[i] * uz[i];
    }
    double a[3][3], V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double inv = 1.0 / (double)N;
    a[0][0] = s[0] * inv - 1.0 / 3.0; a[1][1] = s[1] * inv - 1.0 / 3.0; a[2][2] = s[2] * inv - 1.0 / 3.0;
    a[0][1] = a[1][0] = s[3] * inv; a[0][2] = a[2][0] = s[4] * inv; a[1][2] = a[2][1] = s[5] * inv;
    /* cyclic Jacobi rotations */
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off < 1e-300) break;
        for (int pp = 0; pp < 2; ++pp)
            for (int qq = pp + 1; qq < 3; ++qq) {
                if (a[pp][qq] == 0.0) continue;
                double theta = (a[qq][qq] - a[pp][pp]) / (2.0 * a[pp][qq]);
                double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cc = 1.0 / sqrt(tt * tt + 1.0), ss = tt * cc;
                for (int r = 0; r < 3; ++r) {     /* A <- A J */
                    double arp = a[r][pp], arq = a[r][qq];
                    a[r][pp] = cc * arp - ss * arq;
                    a[r][qq] = ss * arp + cc * arq;
                }
                for (int r = 0; r < 3; ++r) {     /* A <- J^T A */
                    double apr = a[pp][r], aqr = a[qq][r];
                    a[pp][r] = cc * apr - ss * aqr;
                    a[qq][r] = ss * apr + cc * aqr;
                }
                for (int r = 0; r < 3; ++r) {
                    double vrp = V[r][pp], vrq = V[r][qq];
                    V[r][pp] = cc * vrp - ss * vrq;
                    V[r][qq] = ss * vrp + cc * vrq;
                }
            }
    }
    int im = 0;
    for (int d = 1; d < 3; ++d)
        if (a[d][d] > a[im][im]) im = d;
    *p2 = 1.5 * a[im][im];
    director[0] = V[0][im]; director[1] = V[1][im]; director[2] = V[2][im];
}

/* ------------------------------------------------------------------ step-size controller */

void orc_adapt_steps(float *max_displacement, float *max_rotation, int64_t accepted,
                     int64_t attempted, float min_thick, double half_box)
{
    const float ratio = (float)accepted / (float)attempted;       /* :988 */
    float d = *max_displacement, r = *max_rotation;
    if (ratio < 0.3) {                                             /* :999-1014 */
        d = (float)((double)d - (double)d * 0.05);
        if ((double)d < 0.0001 * (double)min_thick) {
            d = (float)(0.01 * (double)min_thick);
            if ((double)r > 0.000001) r = (float)((double)r - 0.5 * (double)r);
        }
    } else if (ratio > 0.55) {                                     /* :1016-1030 */
        d = (float)((double)d + (double)d * 0.05);
        if ((double)d > 1.5 * (double)min_thick) {
            d = (float)(0.75 * (double)min_thick);
            if ((double)r < 0.7) r = (float)((double)r + 0.5 * (double)r);
        }
        if ((double)d >= half_box) return;
    }
    *max_displacement = d;
    *max_rotation = r;
}

/* ------------------------------------------------------------------ batches and chains */

void orc_gb_pair_batch(const orc_consts *c, int n, const double *ei, const double *ej,
                       const double *dr, double *u, int *capped)
{
    for (int k = 0; k < n; ++k) {
        int cp;
        u[k] = orc_gb_pair(c, ei + 3 * (size_t)k, ej + 3 * (size_t)k, dr + 3 * (size_t)k, &cp);
        if (capped) capped[k] = cp;
    }
}

void orc_colloid_batch(const orc_consts *c, int n, const double *rel, const double *e, double *u,
                       int *capped)
{
    for (int k = 0; k < n; ++k) {
        int cp;
        u[k] = orc_colloid(c, rel + 3 * (size_t)k, e + 3 * (size_t)k, &cp);
        if (capped) capped[k] = cp;
    }
}

int orc_run_chains(const orc_consts *c, int N, int nchains, uint32_t *q, double *u, double *L,
                   uint64_t seed, uint32_t t0, int nsweeps, int stride, uint32_t *out_q,
                   double *out_u, double *out_e, int nthreads)
{
    if (N <= 0 || nchains <= 0 || nsweeps < 0 || stride <= 0) return -1;
    const int nrec = nsweeps / stride;
    const size_t n3 = 3 * (size_t)N;
    int bad = 0;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) reduction(|:bad)
    for (int ch = 0; ch < nchains; ++ch) {
        uint32_t *qc = q + (size_t)ch * n3;
        double *uc = u + (size_t)ch * n3;
        double *Lc = L + 3 * (size_t)ch;
        double E[4], energy[3];
        int64_t cnt[9];
        if (orc_total_energy(c, N, qc, qc + N, qc + 2 * N, uc, uc + N, uc + 2 * N, Lc, E, 1)) {
            bad = 1;
            continue;
        }
        energy[0] = E[0]; energy[1] = E[1]; energy[2] = E[2];
        for (int s = 0; s < nsweeps; ++s) {
            memset(cnt, 0, sizeof(cnt));
            if (orc_sweep(c, N, qc, qc + N, qc + 2 * N, uc, uc + N, uc + 2 * N, Lc, seed,
                          (uint32_t)ch, t0 + (uint32_t)s, 0, 0.0, energy, cnt, NULL, NULL, 0, 1)) {
                bad = 1;
                break;
            }
            if ((s + 1) % stride == 0) {
                const size_t rix = (size_t)ch * nrec + (size_t)((s + 1) / stride - 1);
                if (out_q) memcpy(out_q + rix * n3, qc, n3 * sizeof(uint32_t));
                if (out_u) memcpy(out_u + rix * n3, uc, n3 * sizeof(double));
                if (out_e) out_e[rix] = energy[0] * 1.0e9 + energy[1] * 1.0e8 + energy[2];
            }
        }
    }
    return bad ? -1 : 0;
}
