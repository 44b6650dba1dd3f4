/*
 * host_check.c — drives the C host part of libmcgb (csrc/mcgb_host.c) through every entry point
 * that needs no GPU, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_sanitizers.py compiles and runs it).  Exits 0 when every call returns what it
 * should; the sanitizers abort on the first invalid access or undefined operation.
 *
 * Usage: host_check <directory of parameter fixtures> <scratch directory>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mcgb.h"

static int failures = 0;
#define EXPECT(c)                                                                   \
    do {                                                                            \
        if (!(c)) {                                                                 \
            fprintf(stderr, "host_check: %s:%d: %s (%s)\n", __FILE__, __LINE__, #c, \
                    mcgb_last_error());                                             \
            ++failures;                                                             \
        }                                                                           \
    } while (0)

static char *join(const char *dir, const char *name)
{
    size_t n = strlen(dir) + strlen(name) + 2;
    char *s = malloc(n);
    snprintf(s, n, "%s/%s", dir, name);
    return s;
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: host_check <fixtures> <scratch>\n");
        return 2;
    }
    const char *fix = argv[1], *tmp = argv[2];
    mcgb_params p;
    double dv[11];

    /* loader: good files, unknown key (a warning), missing keys, malformed value, no file */
    const char *good[] = {"params_n8000.txt", "params_n256000_colloid.txt", "params_unknown.txt"};
    for (int k = 0; k < 3; ++k) {
        char *f = join(fix, good[k]);
        EXPECT(mcgb_params_load(f, &p) == MCGB_OK);
        EXPECT(mcgb_params_derived(&p, dv) == MCGB_OK && dv[4] > 0.0);
        free(f);
    }
    const char *bad[] = {"params_missing.txt", "params_malformed.txt", "no_such_file.txt"};
    for (int k = 0; k < 3; ++k) {
        char *f = join(fix, bad[k]);
        EXPECT(mcgb_params_load(f, &p) < 0 && strlen(mcgb_last_error()) > 0);
        free(f);
    }

    /* start generators */
    const int n = 1500;
    double *a = calloc(6 * (size_t)n, sizeof(double));
    double *x = a, *y = a + n, *z = a + 2 * n, *ux = a + 3 * n, *uy = a + 4 * n, *uz = a + 5 * n;
    double box[3], box2[3], R;
    EXPECT(mcgb_init_lattice(n, 0.3, 3.0, 3.5, 0.05, 7, x, y, z, ux, uy, uz, box) == MCGB_OK);
    EXPECT(mcgb_init_cubic(n, 0.3, 3.0, 8, x, y, z, ux, uy, uz, box) == MCGB_OK);
    EXPECT(mcgb_init_random(200, 0.001, 0.0, 3.0, 9, 100000, x, y, z, ux, uy, uz, box) == MCGB_OK);
    EXPECT(mcgb_init_random(n, 0.3, 0.0, 3.0, 9, 1000, x, y, z, ux, uy, uz, box) ==
           MCGB_ERR_GEOMETRY);                                  /* jammed: stops, no hang */
    EXPECT(mcgb_init_lattice(n, 0.3, 3.0, 3.5, 0.05, 7, x, y, z, ux, uy, uz, box) == MCGB_OK);

    /* restart and resume files round trip */
    double *b = calloc(6 * (size_t)n, sizeof(double));
    char *f1 = join(tmp, "fortSphere.1");
    EXPECT(mcgb_write_fort1(f1, n, x, y, z, ux, uy, uz, box, 3.0) == MCGB_OK);
    EXPECT(mcgb_read_fort1(f1, n, b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, box2,
                           &R) == MCGB_OK && R == 3.0);
    EXPECT(mcgb_read_fort1(f1, n + 1, b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n,
                           box2, &R) < 0);                      /* wrong particle count */
    uint32_t *q = calloc(6 * (size_t)n, sizeof(uint32_t));
    for (int i = 0; i < 3 * n; ++i) q[i] = (uint32_t)(2654435761u * (unsigned)i);
    mcgb_resume rec = {42, 17, 0.1f, 0.1f, 3, 1, 2, 5, 0, -1234.5};
    mcgb_resume back;
    char *fr = join(tmp, "state.mcgb");
    EXPECT(mcgb_write_resume(fr, n, q, q + n, q + 2 * n, ux, uy, uz, box, &rec) == MCGB_OK);
    EXPECT(mcgb_read_resume(fr, n, q + 3 * n, q + 4 * n, q + 5 * n, b, b + n, b + 2 * n, box2,
                            &back) == MCGB_OK);
    EXPECT(memcmp(q, q + 3 * n, 3 * (size_t)n * sizeof(uint32_t)) == 0);
    EXPECT(back.seed == 42 && back.sweep_index == 17 && back.energy_finite == -1234.5);

    /* trajectory writers, fresh and appended */
    char *ft = join(tmp, "Configuraciones.dat"), *fq = join(tmp, "qmga.cnf"), *fx = join(tmp, "x.xyz");
    for (int app = 0; app < 2; ++app) {
        EXPECT(mcgb_write_trajectory(ft, app, n, x, y, z, ux, uy, uz, box) == MCGB_OK);
        EXPECT(mcgb_write_qmga(fq, app, n, x, y, z, ux, uy, uz, box) == MCGB_OK);
        EXPECT(mcgb_write_xyz(fx, app, n, x, y, z, box) == MCGB_OK);
    }

    /* integrated autocorrelation time of an AR(1) series, eigen solver */
    const int ns = 20000;
    double *s = malloc(sizeof(double) * ns), tau;
    int w;
    uint64_t st = 88172645463325252ull;
    s[0] = 0.0;
    for (int i = 1; i < ns; ++i) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        s[i] = 0.8 * s[i - 1] + ((double)(st >> 11) / 9007199254740992.0 - 0.5);
    }
    EXPECT(mcgb_tau_int(s, ns, 5.0, &tau, &w) == MCGB_OK && fabs(tau - 9.0) < 2.0);
    EXPECT(mcgb_tau_int(s, 1, 5.0, &tau, &w) < 0);
    double qm[6] = {0.3, -0.1, -0.2, 0.05, -0.02, 0.01}, ev[3], V[9];
    EXPECT(mcgb_eigen_sym3(qm, ev, V) == MCGB_OK && ev[0] <= ev[1] && ev[1] <= ev[2]);
    double qd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    EXPECT(mcgb_eigen_sym3(qd, ev, V) == MCGB_OK);

    free(s); free(ft); free(fq); free(fx); free(fr); free(f1); free(q); free(b); free(a);
    if (failures) fprintf(stderr, "host_check: %d failure(s)\n", failures);
    else printf("host_check: ok\n");
    return failures ? 1 : 0;
}
