/*
 * oracle_check.c — drives the FP64 oracle (oracle/mcgb_oracle.c) through its functions on small
 * systems, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_sanitizers.py compiles and runs it, with OpenMP on two threads).  Exits 0 when
 * the running energies agree with the recomputations; the sanitizers abort on the first
 * invalid access or undefined operation.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mcgb_oracle.h"

static int failures = 0;
#define EXPECT(c)                                                                   \
    do {                                                                            \
        if (!(c)) {                                                                 \
            fprintf(stderr, "oracle_check: %s:%d: %s\n", __FILE__, __LINE__, #c);   \
            ++failures;                                                             \
        }                                                                           \
    } while (0)

static uint64_t rs = 0x9E3779B97F4A7C15ull;
static double urand(void)
{
    rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17;
    return (double)(rs >> 11) / 9007199254740992.0;
}

/* layered start: rods along z on square layers, colloid sites removed; returns the count */
static void layered(int N, double L, double R, uint32_t *q, double *u)
{
    int nz = (int)(L / 3.25), nx = (int)ceil(sqrt((double)N / nz)), k = 0;
    for (;;) {
        k = 0;
        for (int iz = 0; iz < nz && k < N; ++iz)
            for (int iy = 0; iy < nx && k < N; ++iy)
                for (int ix = 0; ix < nx && k < N; ++ix) {
                    double p[3] = {(ix + 0.5) * L / nx, (iy + 0.5) * L / nx, (iz + 0.5) * L / nz};
                    double d2 = 0.0;
                    for (int d = 0; d < 3; ++d) d2 += (p[d] - 0.5 * L) * (p[d] - 0.5 * L);
                    if (R > 0.0 && d2 < (R + 2.0) * (R + 2.0)) continue;
                    for (int d = 0; d < 3; ++d) q[d * N + k] = (uint32_t)(p[d] / L * 4294967296.0);
                    double v[3] = {0.1 * (urand() - 0.5), 0.1 * (urand() - 0.5), 1.0};
                    double nn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    for (int d = 0; d < 3; ++d) u[d * N + k] = v[d] / nn;
                    ++k;
                }
        if (k == N) return;
        ++nx;
    }
}

static void run(int colors, double R, int npt, int nthreads)
{
    orc_consts c;
    memset(&c, 0, sizeof(c));
    c.sigma0 = 1.0; c.epsilon0 = 1.0; c.kappa = 3.0; c.kappa_prime = 5.0; c.mu = 2.0; c.nu = 1.0;
    c.temperature = 1.0; c.pressure = 2.0; c.cutoff = 4.0; c.max_displacement = 0.1;
    c.max_rotation = 0.1; c.colloid_radius = R; c.colloid_on = R > 0.0; c.colors = colors;
    orc_consts_init(&c);
    const int N = 1300;
    const double L0 = cbrt(N / 0.25 + 4.0 / 3.0 * M_PI * R * R * R);
    double L[3] = {L0, L0, L0};
    uint32_t *q = malloc(3 * sizeof(uint32_t) * N);
    double *u = malloc(3 * sizeof(double) * N);
    layered(N, L0, R, q, u);
    double E[4], energy[3];
    EXPECT(orc_total_energy(&c, N, q, q + N, q + 2 * N, u, u + N, u + 2 * N, L, E, nthreads) == 0);
    energy[0] = E[0]; energy[1] = E[1]; energy[2] = E[2];
    int64_t cnt[9] = {0};
    const int ncell = 8 * 8 * 8 * 8;
    orc_move_trace *tr = malloc(sizeof(orc_move_trace) * (size_t)(N + ncell));
    uint32_t info[36];
    for (uint32_t t = 0; t < 4; ++t)
        EXPECT(orc_sweep(&c, N, q, q + N, q + 2 * N, u, u + N, u + 2 * N, L, 99, 0, t, npt, 0.01,
                         energy, cnt, t == 1 ? tr : NULL, info, t == 3, nthreads) == 0);
    EXPECT(cnt[0] > 0 && cnt[1] > 0);
    EXPECT(orc_total_energy(&c, N, q, q + N, q + 2 * N, u, u + N, u + 2 * N, L, E, nthreads) == 0);
    EXPECT(energy[0] == E[0] && energy[1] == E[1]);
    EXPECT(fabs(energy[2] - E[2]) <= 1e-8 * (fabs(E[2]) + N));
    double x[3 * 1300], B[4];
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < N; ++i) x[d * N + i] = q[d * N + i] * (L[d] / 4294967296.0);
    EXPECT(orc_total_energy_brute(&c, N, x, x + N, x + 2 * N, u, u + N, u + 2 * N, L, B) == 0);
    EXPECT(B[0] == E[0] && fabs(B[2] - E[2]) <= 1e-9 * (fabs(E[2]) + N));
    double p2, dir[3];
    orc_order_parameter(N, u, u + N, u + 2 * N, &p2, dir);
    EXPECT(p2 > 0.5 && p2 <= 1.0);
    free(tr); free(u); free(q);
}

int main(void)
{
    run(2, 0.0, 0, 1);
    run(2, 3.0, 0, 2);
    run(3, 3.0, 0, 2);
    run(2, 0.0, 1, 2);

    /* batches, draws, the step controller, tiny chains */
    orc_consts c;
    memset(&c, 0, sizeof(c));
    c.sigma0 = 1.0; c.epsilon0 = 1.0; c.kappa = 3.0; c.kappa_prime = 5.0; c.mu = 2.0; c.nu = 1.0;
    c.temperature = 0.6; c.cutoff = 4.0; c.max_displacement = 3.0; c.max_rotation = 1.0;
    c.colloid_radius = 3.0; c.colloid_on = 1;
    orc_consts_init(&c);
    double ei[300], ej[300], dr[300], ub[100];
    int cp[100];
    for (int k = 0; k < 300; ++k) { ei[k] = urand() - 0.5; ej[k] = urand() - 0.5; dr[k] = 6.0 * (urand() - 0.5); }
    orc_gb_pair_batch(&c, 100, ei, ej, dr, ub, cp);
    orc_colloid_batch(&c, 100, dr, ei, ub, cp);
    double Ld[3] = {16.5, 16.5, 16.5}, v[30], thr[10];
    uint32_t j[10];
    int32_t dq[30];
    orc_move_draws(7, 1, 3, 2, 11, 0, 10, Ld, 3.0, 0.6, j, dq, v, thr);
    for (int k = 0; k < 10; ++k)
        EXPECT(fabs(v[3 * k] * v[3 * k] + v[3 * k + 1] * v[3 * k + 1] + v[3 * k + 2] * v[3 * k + 2] - 1.0) < 1e-14);
    float dm = 0.1f, drt = 0.1f;
    orc_adapt_steps(&dm, &drt, 1, 100, 1.0f, 8.0);
    orc_adapt_steps(&dm, &drt, 90, 100, 1.0f, 8.0);
    uint32_t qc[2 * 3 * 2] = {0};
    double uc[2 * 3 * 2], Lc[6] = {16.5, 16.5, 16.5, 16.5, 16.5, 16.5};
    for (int k = 0; k < 12; ++k) uc[k] = (k % 3 == 2) ? 1.0 : 0.0;
    qc[1] = 0x80000000u; qc[7] = 0x80000000u;
    double oe[2 * 10];
    EXPECT(orc_run_chains(&c, 2, 2, qc, uc, Lc, 5, 0, 200, 20, NULL, NULL, oe, 2) == 0);
    if (failures) fprintf(stderr, "oracle_check: %d failure(s)\n", failures);
    else printf("oracle_check: ok\n");
    return failures ? 1 : 0;
}
