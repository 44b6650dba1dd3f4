/*
 * mcgb — C driver of the GPU Gay-Berne Monte Carlo (the role of main(), MCnptColloid.c:379-726,
 * with the serial statistics cadence of :887-1032).
 *
 *   mcgb [parametersFile.txt]
 *
 * Reads the reference's 23 keys (plus optional ones), builds the start configuration
 * (initial_configuration_type 0: fortSphere.1 as at MCnptColloid.c:2786-2831; 1: random
 * sequential addition, :2835-2920; 2: cubic lattice, :2924-3033; 3: the seeded jittered
 * lattice of the benchmark configurations), runs total_mc_cycles checkerboard sweeps on the GPU,
 * prints statistics every statistics_print_interval sweeps, and writes the reference's outputs:
 * Configuraciones.dat at the start and every configuration_save_interval sweeps together with
 * fortSphere.1.BAK (:877, :958-960), and qmga_mcnvtGB.cnf, fortSphere.1 and DiscosEsfera.xyz at
 * the end (:1048-1061).  Every error prints its message and exits with status 1, as the
 * reference does.
 *
 * Beside the final fortSphere.1 it writes fortSphere.1.mcgb, the exact state with the resume
 * record (sweep index, step sizes, adaptation counters, running energy).  A later run with
 * initial_configuration_type 0 and the same seed continues from it bit for bit when it matches
 * fortSphere.1 (same box, orientations and positions to the 8 printed decimals); otherwise it
 * starts from fortSphere.1 alone, as the reference does.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "mcgb.h"

static void die(const char *what)
{
    fprintf(stderr, "%s\n%s\n", what, mcgb_last_error());
    exit(1);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Save point (MCnptColloid.c:958-960): the snapshot begun after the save sweep, appended to
 * Configuraciones.dat and written to fortSphere.1.BAK. */
static void save_point(mcgb_ctx *ctx, int N, double *x, double *y, double *z, double *ux,
                       double *uy, double *uz, double *box, double R)
{
    if (mcgb_snapshot_wait(ctx, 0, N, x, y, z, ux, uy, uz, box) != MCGB_OK)
        die("cannot read the snapshot");
    if (mcgb_write_trajectory("Configuraciones.dat", 1, N, x, y, z, ux, uy, uz, box) != MCGB_OK)
        die("cannot write Configuraciones.dat");
    if (mcgb_write_fort1("fortSphere.1.BAK", N, x, y, z, ux, uy, uz, box, R) != MCGB_OK)
        die("cannot write fortSphere.1.BAK");
}

/* 1 when the resume file's state is the configuration read from fortSphere.1 (printed with 8
 * decimals: positions, orientations and box agree within 1e-6, positions up to a period). */
static int resume_matches(int N, const uint32_t *qx, const uint32_t *qy, const uint32_t *qz,
                          const double *rux, const double *ruy, const double *ruz,
                          const double *rbox, const double *x, const double *y, const double *z,
                          const double *ux, const double *uy, const double *uz,
                          const double *box)
{
    const double tol = 1e-6;
    for (int d = 0; d < 3; ++d)
        if (fabs(rbox[d] - box[d]) > tol) return 0;
    const uint32_t *q[3] = {qx, qy, qz};
    const double *c[3] = {x, y, z};
    for (int i = 0; i < N; ++i) {
        if (fabs(rux[i] - ux[i]) > tol || fabs(ruy[i] - uy[i]) > tol || fabs(ruz[i] - uz[i]) > tol)
            return 0;
        for (int d = 0; d < 3; ++d) {
            double e = fabs(ldexp((double)q[d][i], -32) * rbox[d] - c[d][i]);
            if (fmin(e, fabs(e - rbox[d])) > tol) return 0;
        }
    }
    return 1;
}

int main(int argc, char **argv)
{
    const char *path = argc > 1 ? argv[1] : "parametersFile.txt";
    mcgb_params p;
    if (mcgb_params_load(path, &p) != MCGB_OK) die("cannot load parameters");
    const int N = p.num_particles;
    double *x = malloc(sizeof(double) * N), *y = malloc(sizeof(double) * N),
           *z = malloc(sizeof(double) * N), *ux = malloc(sizeof(double) * N),
           *uy = malloc(sizeof(double) * N), *uz = malloc(sizeof(double) * N);
    uint32_t *qx = malloc(sizeof(uint32_t) * N), *qy = malloc(sizeof(uint32_t) * N),
             *qz = malloc(sizeof(uint32_t) * N);
    double *rux = malloc(sizeof(double) * N), *ruy = malloc(sizeof(double) * N),
           *ruz = malloc(sizeof(double) * N);
    if (!x || !y || !z || !ux || !uy || !uz || !qx || !qy || !qz || !rux || !ruy || !ruz)
        die("out of memory");
    double rbox[3];
    mcgb_resume rec;
    int resume = 0;
    double box[3], R = p.colloid_radius;
    const double Rc = p.colloid_wall_option == 1 ? p.colloid_radius : 0.0;
    int ok = MCGB_OK;
    switch (p.initial_configuration_type) {
    case 0: ok = mcgb_read_fort1("fortSphere.1", N, x, y, z, ux, uy, uz, box, &R); break;
    /* types 1 and 2 size the box with colloid_radius whatever colloid_wall_option says, as the
     * reference's generators do (MCnptColloid.c:2843, :2932) */
    case 1: ok = mcgb_init_random(N, p.system_density, p.colloid_radius,
                                  p.gb_kappa * p.gb_sigma_sidetoside, p.seed, 1000000,
                                  x, y, z, ux, uy, uz, box); break;
    case 2: ok = mcgb_init_cubic(N, p.system_density, p.colloid_radius, p.seed,
                                 x, y, z, ux, uy, uz, box); break;
    default: ok = mcgb_init_lattice(N, p.system_density, Rc, p.colloid_exclusion_radius,
                                    p.lattice_jitter, p.seed, x, y, z, ux, uy, uz, box); break;
    }
    if (ok != MCGB_OK) die("cannot build the start configuration");
    if (p.initial_configuration_type == 0 &&
        mcgb_read_resume("fortSphere.1.mcgb", N, qx, qy, qz, rux, ruy, ruz, rbox, &rec) == MCGB_OK) {
        resume = rec.seed == p.seed &&
                 resume_matches(N, qx, qy, qz, rux, ruy, ruz, rbox, x, y, z, ux, uy, uz, box);
        if (!resume)
            printf("fortSphere.1.mcgb belongs to another configuration or seed: "
                   "starting from fortSphere.1\n");
    }
    /* the first frame is the start configuration; a resumed one is the exact state, known once
     * it is on the GPU */
    if (!resume &&
        mcgb_write_trajectory("Configuraciones.dat", 0, N, x, y, z, ux, uy, uz, box) != MCGB_OK)
        die("cannot write Configuraciones.dat");
    mcgb_ctx *ctx;
    if (mcgb_create(&p, 1, &ctx) != MCGB_OK) die("cannot create the GPU context");
    if (resume) {
        if (mcgb_resume_state(ctx, 0, N, qx, qy, qz, rux, ruy, ruz, rbox, &rec) != MCGB_OK ||
            mcgb_get_state(ctx, 0, N, x, y, z, ux, uy, uz, box) != MCGB_OK)
            die("cannot resume from fortSphere.1.mcgb");
        printf("resuming from fortSphere.1.mcgb after %llu sweeps\n",
               (unsigned long long)rec.sweep_index);
        if (mcgb_write_trajectory("Configuraciones.dat", 0, N, x, y, z, ux, uy, uz, box) != MCGB_OK)
            die("cannot write Configuraciones.dat");
    } else if (mcgb_set_state(ctx, 0, N, x, y, z, ux, uy, uz, box) != MCGB_OK) {
        die("cannot upload the state");
    }

    printf("================================================\n");
    printf("N = %d  box = %.5f x %.5f x %.5f  T* = %.4f  %s  precision FP%d\n", N, box[0], box[1],
           box[2], p.system_temperature, p.mc_ensemble_type == 2 ? "NPT" : "NVT", p.precision);
    printf("================================================\n");
    printf("%10s %16s %12s %10s %10s %10s\n", "sweep", "E/N", "P2", "density", "acc.ratio", "moves/s");

    const int total = p.total_mc_cycles;
    const int every = p.statistics_print_interval > 0 ? p.statistics_print_interval : total;
    const int save = p.configuration_save_interval;
    double t0 = now_s();
    uint64_t last_att = 0;
    double last_t = t0;
    int snap = 0;   /* a save-point snapshot is in flight: written while the next sweeps run */
    for (int done = 0; done < total;) {
        int chunk = every < total - done ? every : total - done;
        if (mcgb_sweep_async(ctx, chunk) != MCGB_OK) die("sweep failed");
        if (snap) {
            save_point(ctx, N, x, y, z, ux, uy, uz, box, R);
            snap = 0;
        }
        if (mcgb_sync(ctx) != MCGB_OK) die("sweep failed");
        done += chunk;
        mcgb_stats s;
        double p2, dir[3];
        if (mcgb_get_stats(ctx, 0, &s) != MCGB_OK || mcgb_order_parameter(ctx, 0, &p2, dir) != MCGB_OK)
            die("cannot read statistics");
        double t = now_s();
        double rate = (double)(s.attempted - last_att) / (t - last_t);
        last_att = s.attempted;
        last_t = t;
        printf("%10d %16.8f %12.6f %10.6f %10.4f %10.3e\n", done, s.energy / N, p2, s.density,
               s.attempted ? (double)s.accepted / s.attempted : 0.0, rate);
        if (save > 0 && (done % save == 0 || done == total)) {
            if (mcgb_snapshot_begin(ctx, 0) != MCGB_OK) die("cannot take a snapshot");
            snap = 1;
        }
    }
    if (snap) save_point(ctx, N, x, y, z, ux, uy, uz, box, R);
    if (mcgb_get_state(ctx, 0, N, x, y, z, ux, uy, uz, box) != MCGB_OK) die("cannot read state");
    if (mcgb_write_qmga("qmga_mcnvtGB.cnf", 0, N, x, y, z, ux, uy, uz, box) != MCGB_OK)
        die("cannot write qmga_mcnvtGB.cnf");
    if (mcgb_write_fort1("fortSphere.1", N, x, y, z, ux, uy, uz, box, R) != MCGB_OK)
        die("cannot write fortSphere.1");
    if (mcgb_get_state_fixed(ctx, 0, N, qx, qy, qz, rux, ruy, ruz, rbox) != MCGB_OK ||
        mcgb_get_resume(ctx, 0, &rec) != MCGB_OK ||
        mcgb_write_resume("fortSphere.1.mcgb", N, qx, qy, qz, rux, ruy, ruz, rbox, &rec) != MCGB_OK)
        die("cannot write fortSphere.1.mcgb");
    if (mcgb_write_xyz("DiscosEsfera.xyz", 0, N, x, y, z, box) != MCGB_OK ||
        mcgb_write_xyz("DiscosEsfera.xyz", 1, N, x, y, z, box) != MCGB_OK)
        die("cannot write DiscosEsfera.xyz");
    mcgb_stats s;
    mcgb_get_stats(ctx, 0, &s);
    printf("done: %llu trial moves in %.3f s, acceptance %.4f, max FP64 drift %.3e\n",
           (unsigned long long)s.attempted, now_s() - t0,
           s.attempted ? (double)s.accepted / s.attempted : 0.0, s.max_drift);
    mcgb_destroy(ctx);
    free(x); free(y); free(z); free(ux); free(uy); free(uz);
    free(qx); free(qy); free(qz); free(rux); free(ruy); free(ruz);
    return 0;
}
