/*
 * mcgb_host.c — host-side C part of libmcgb: parameter file loader, derived constants,
 * synthetic start configurations, integrated autocorrelation time, 3x3 symmetric eigen solver,
 * fortSphere.1 restart files, last-error text.  No GPU code.
 */
#include "mcgb_internal.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ error text */

static __thread char g_err[512];

void mcgb_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *mcgb_last_error(void) { return g_err; }

/* ------------------------------------------------------------------ parameters */

/* Key names in the reference's order (MCnptColloid.c:4104-4112). */
static const char *const k_required[MCGB_NUM_REQUIRED_KEYS] = {
    "mc_ensemble_type", "num_particles", "system_density", "system_temperature",
    "system_pressure", "gb_sigma_sidetoside", "gb_epsilon0", "gb_kappa",
    "gb_kappa_prime", "gb_mu_exponent", "gb_nu_exponent", "max_displacement",
    "max_rotation", "initial_configuration_type", "cutoff_radius",
    "verlet_radius", "total_mc_cycles", "configuration_save_interval",
    "statistics_print_interval", "colloid_wall_option", "colloid_interaction_epsilon",
    "colloid_interaction_amplitude", "colloid_radius"};

void mcgb_params_default(mcgb_params *p)
{
    memset(p, 0, sizeof(*p));
    p->seed = 20261203ull;
    p->precision = 32;
    p->npt_max_dlnv = 0.01;
    p->energy_check_interval = 100;
    p->lattice_jitter = 0.0;
    p->colloid_exclusion_radius = -1.0;      /* < 0: colloid_radius + 0.5 */
    p->waves_per_cell = 2;
    p->pass_schedule = 1;
    p->adapt_moves = 0;
    p->sublattice_colors = 2;
}

enum { T_INT, T_DOUBLE, T_FLOAT, T_U64 };

static int parse_value(const char *value, int type, void *dst)
{
    /* sscanf(value, "%d"/"%lf"/"%f") as at MCnptColloid.c:3955-4089: leading number accepted */
    switch (type) {
    case T_INT: return sscanf(value, "%d", (int *)dst);
    case T_DOUBLE: return sscanf(value, "%lf", (double *)dst);
    case T_FLOAT: return sscanf(value, "%f", (float *)dst);
    default: {
        unsigned long long v;
        int n = sscanf(value, "%llu", &v);
        if (n == 1) *(uint64_t *)dst = (uint64_t)v;
        return n;
    }
    }
}

int mcgb_params_load(const char *path, mcgb_params *out)
{
    mcgb_params p;
    mcgb_params_default(&p);
    FILE *f = fopen(path, "r");
    if (!f) {
        mcgb_set_error("Error opening the parameters file: %s (make sure the file exists)", path);
        return MCGB_ERR_IO;
    }
    struct { const char *name; int type; void *dst; } keys[] = {
        {"mc_ensemble_type", T_INT, &p.mc_ensemble_type},
        {"num_particles", T_INT, &p.num_particles},
        {"system_density", T_DOUBLE, &p.system_density},
        {"system_temperature", T_DOUBLE, &p.system_temperature},
        {"system_pressure", T_DOUBLE, &p.system_pressure},
        {"gb_sigma_sidetoside", T_DOUBLE, &p.gb_sigma_sidetoside},
        {"gb_epsilon0", T_DOUBLE, &p.gb_epsilon0},
        {"gb_kappa", T_DOUBLE, &p.gb_kappa},
        {"gb_kappa_prime", T_DOUBLE, &p.gb_kappa_prime},
        {"gb_mu_exponent", T_DOUBLE, &p.gb_mu_exponent},
        {"gb_nu_exponent", T_DOUBLE, &p.gb_nu_exponent},
        {"max_displacement", T_FLOAT, &p.max_displacement},
        {"max_rotation", T_FLOAT, &p.max_rotation},
        {"initial_configuration_type", T_INT, &p.initial_configuration_type},
        {"cutoff_radius", T_DOUBLE, &p.cutoff_radius},
        {"verlet_radius", T_DOUBLE, &p.verlet_radius},
        {"total_mc_cycles", T_INT, &p.total_mc_cycles},
        {"configuration_save_interval", T_INT, &p.configuration_save_interval},
        {"statistics_print_interval", T_INT, &p.statistics_print_interval},
        {"colloid_wall_option", T_INT, &p.colloid_wall_option},
        {"colloid_interaction_epsilon", T_DOUBLE, &p.colloid_interaction_epsilon},
        {"colloid_interaction_amplitude", T_DOUBLE, &p.colloid_interaction_amplitude},
        {"colloid_radius", T_DOUBLE, &p.colloid_radius},
        /* optional keys */
        {"seed", T_U64, &p.seed},
        {"precision", T_INT, &p.precision},
        {"npt_max_dlnv", T_DOUBLE, &p.npt_max_dlnv},
        {"energy_check_interval", T_INT, &p.energy_check_interval},
        {"lattice_jitter", T_DOUBLE, &p.lattice_jitter},
        {"colloid_exclusion_radius", T_DOUBLE, &p.colloid_exclusion_radius},
        {"stage_capacity", T_INT, &p.stage_capacity},
        {"waves_per_cell", T_INT, &p.waves_per_cell},
        {"pass_schedule", T_INT, &p.pass_schedule},
        {"adapt_moves", T_INT, &p.adapt_moves},
        {"sublattice_colors", T_INT, &p.sublattice_colors},
    };
    const int nkeys = (int)(sizeof(keys) / sizeof(keys[0]));
    char line[256], name[256], value[256];
    int total = 0;
    /* read_next_parameter, MCnptColloid.c:3892-3906 */
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '\n' || line[0] == '#' || line[0] == '\0') continue;
        value[0] = '\0';
        if (sscanf(line, "%255s %255[^\n]", name, value) < 1) continue;
        int k;
        for (k = 0; k < nkeys; ++k)
            if (strcmp(name, keys[k].name) == 0) break;
        if (k == nkeys) {
            printf("Warning: Unknown parameter '%s' in configuration file\n", name);
            continue;
        }
        if (parse_value(value, keys[k].type, keys[k].dst) != 1) {
            fclose(f);
            mcgb_set_error("Error reading parameter: %s", name);
            return MCGB_ERR_PARAM;
        }
        if (k < MCGB_NUM_REQUIRED_KEYS) {
            p.keys_read |= 1u << k;
            total++;
        }
    }
    fclose(f);
    fflush(stdout);
    if (total != MCGB_NUM_REQUIRED_KEYS || p.keys_read != (1u << MCGB_NUM_REQUIRED_KEYS) - 1u) {
        char buf[512];
        int len = snprintf(buf, sizeof(buf),
                           "Error: Not all parameters were read correctly. Expected 23, got %d\n"
                           "Missing parameters:\n", total);
        for (int i = 0; i < MCGB_NUM_REQUIRED_KEYS && len < (int)sizeof(buf); ++i)
            if (!(p.keys_read & (1u << i)))
                len += snprintf(buf + len, sizeof(buf) - (size_t)len, "  - %s%s\n", k_required[i],
                                (i == 5 || i == 7) ? " cannot generate gb_sigma_endtoend variable." : "");
        mcgb_set_error("%s", buf);
        return MCGB_ERR_PARAM;
    }
    *out = p;
    return MCGB_OK;
}

int mcgb_params_derived(const mcgb_params *p, double out[11])
{
    if (!p || !out || p->system_temperature <= 0.0 || p->gb_mu_exponent == 0.0) {
        mcgb_set_error("mcgb_params_derived: need a parameter block with T* > 0 and mu != 0");
        return MCGB_ERR_ARG;
    }
    /* get_system_const, MCnptColloid.c:2752-2783, in dependency order */
    double k2 = p->gb_kappa * p->gb_kappa;
    double chi = (k2 - 1.0) / (k2 + 1.0);
    double kimu = pow(p->gb_kappa_prime, 1.0 / p->gb_mu_exponent);
    double sigma_ee = p->gb_sigma_sidetoside * p->gb_kappa;   /* :4130 */
    out[0] = 1.0 / p->system_temperature;
    out[1] = 4.0 * p->gb_epsilon0;
    out[2] = p->cutoff_radius * p->cutoff_radius;
    out[3] = k2;
    out[4] = chi;
    out[5] = chi * chi;
    out[6] = kimu;
    out[7] = (kimu - 1.0) / (kimu + 1.0);
    out[8] = sigma_ee;
    out[9] = 4.0 * sigma_ee;
    out[10] = p->colloid_radius * p->colloid_radius;
    return MCGB_OK;
}

/* ------------------------------------------------------------------ host Philox and helpers */

void mcgb_philox_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1; c3 = (uint32_t)p0; c0 = n0; c2 = n2;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

int mcgb_cell_dim(double L, double rc)
{
    int n = 2 * (int)floor(L / (2.0 * rc * (1.0 + 1e-6)));
    return n >= 4 ? n : 0;
}

uint32_t mcgb_to_fixed(double x, double L)
{
    double f = x / L;
    f -= floor(f);
    double q = floor(f * 4294967296.0);
    if (q >= 4294967296.0) q = 0.0;
    if (q < 0.0) q = 0.0;
    return (uint32_t)(uint64_t)q;
}

/* ------------------------------------------------------------------ synthetic starts */

typedef struct { uint32_t key[2]; uint32_t ctr; uint32_t buf[4]; int used; } host_rng;

static uint32_t rng_next(host_rng *g)
{
    if (g->used == 4) {
        uint32_t c[4] = {g->ctr++, 0u, 0u, 0x5EED5EEDu};
        mcgb_philox_host(c, g->key, g->buf);
        g->used = 0;
    }
    return g->buf[g->used++];
}

static double rng_u01(host_rng *g) { return ((double)rng_next(g) + 0.5) * (1.0 / 4294967296.0); }

int mcgb_init_lattice(int n, double density, double colloid_radius, double exclusion_radius,
                      double jitter, uint64_t seed,
                      double *x, double *y, double *z, double *ux, double *uy, double *uz,
                      double box[3])
{
    if (n <= 0 || density <= 0.0 || !x || !y || !z || !ux || !uy || !uz || !box) {
        mcgb_set_error("mcgb_init_lattice: bad arguments");
        return MCGB_ERR_ARG;
    }
    const double pi = 3.14159265358979323846;
    double V = (double)n / density + (4.0 / 3.0) * pi * colloid_radius * colloid_radius * colloid_radius;
    double L = cbrt(V);
    double excl = exclusion_radius < 0.0 ? colloid_radius + 0.5 : exclusion_radius;
    double c = 0.5 * L;
    int ns = (int)ceil(cbrt((double)n) - 1e-9);
    if (ns < 1) ns = 1;
    int64_t nsites = 0;
    int64_t *site = NULL;
    for (;; ++ns) {
        double a = L / ns;
        int64_t tot = (int64_t)ns * ns * ns;
        free(site);
        site = (int64_t *)malloc(sizeof(int64_t) * (size_t)tot);
        if (!site) { mcgb_set_error("mcgb_init_lattice: out of memory"); return MCGB_ERR_NOMEM; }
        nsites = 0;
        for (int64_t s = 0; s < tot; ++s) {
            int i = (int)(s % ns), j = (int)((s / ns) % ns), k = (int)(s / ((int64_t)ns * ns));
            double px = (i + 0.5) * a - c, py = (j + 0.5) * a - c, pz = (k + 0.5) * a - c;
            if (colloid_radius > 0.0 && px * px + py * py + pz * pz < excl * excl) continue;
            site[nsites++] = s;
        }
        if (nsites >= n) break;
    }
    host_rng g = {{(uint32_t)seed, (uint32_t)(seed >> 32)}, 0u, {0, 0, 0, 0}, 4};
    /* partial Fisher-Yates: the first n entries become a uniform random subset */
    for (int64_t s = 0; s < n; ++s) {
        uint64_t r = ((uint64_t)rng_next(&g) << 32) | rng_next(&g);
        int64_t j = s + (int64_t)(r % (uint64_t)(nsites - s));
        int64_t t = site[s]; site[s] = site[j]; site[j] = t;
    }
    double a = L / ns;
    for (int p = 0; p < n; ++p) {
        int64_t s = site[p];
        int i = (int)(s % ns), j = (int)((s / ns) % ns), k = (int)(s / ((int64_t)ns * ns));
        double px = (i + 0.5) * a + jitter * (2.0 * rng_u01(&g) - 1.0);
        double py = (j + 0.5) * a + jitter * (2.0 * rng_u01(&g) - 1.0);
        double pz = (k + 0.5) * a + jitter * (2.0 * rng_u01(&g) - 1.0);
        x[p] = px - L * floor(px / L);
        y[p] = py - L * floor(py / L);
        z[p] = pz - L * floor(pz / L);
        /* uniform on the sphere: z = 1 - 2U, phi = 2 pi U' */
        double zz = 1.0 - 2.0 * rng_u01(&g), phi = 2.0 * pi * rng_u01(&g);
        double sxy = sqrt(fmax(0.0, 1.0 - zz * zz));
        ux[p] = sxy * cos(phi);
        uy[p] = sxy * sin(phi);
        uz[p] = zz;
    }
    free(site);
    box[0] = box[1] = box[2] = L;
    return MCGB_OK;
}

/* ------------------------------------------------------------------ statistics */

int mcgb_tau_int(const double *x, int n, double c, double *tau, int *window)
{
    if (!x || n < 4 || !tau) {
        mcgb_set_error("mcgb_tau_int: need at least 4 samples");
        return MCGB_ERR_ARG;
    }
    double mean = 0.0;
    for (int i = 0; i < n; ++i) mean += x[i];
    mean /= n;
    double c0 = 0.0;
    for (int i = 0; i < n; ++i) c0 += (x[i] - mean) * (x[i] - mean);
    c0 /= n;
    if (c0 <= 0.0) { *tau = 1.0; if (window) *window = 0; return MCGB_OK; }
    double t = 1.0;
    int M;
    for (M = 1; M < n / 2; ++M) {
        double ct = 0.0;
        for (int i = 0; i + M < n; ++i) ct += (x[i] - mean) * (x[i + M] - mean);
        ct /= n;
        t += 2.0 * ct / c0;
        if ((double)M >= c * t) break;       /* Sokal's automatic window */
    }
    if (t < 1.0 / n) t = 1.0 / n;
    *tau = t;
    if (window) *window = M;
    return MCGB_OK;
}

/* Symmetric 3x3 eigen-decomposition: closed-form eigenvalues (trigonometric solution of the
 * characteristic cubic); eigenvectors by the extreme-eigenvalue-first construction: the vector
 * of the better separated extreme eigenvalue from the largest cross product of two rows of
 * (A - lambda I), the middle one from the 2x2 projection of A on its orthogonal complement, the
 * last by a cross product (robust for repeated eigenvalues). */
static void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3h(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void unit_vector_for(const double A[3][3], double lam, double v[3])
{
    double r0[3] = {A[0][0] - lam, A[0][1], A[0][2]};
    double r1[3] = {A[1][0], A[1][1] - lam, A[1][2]};
    double r2[3] = {A[2][0], A[2][1], A[2][2] - lam};
    double c[3][3];
    cross3(r0, r1, c[0]);
    cross3(r0, r2, c[1]);
    cross3(r1, r2, c[2]);
    int best = 0;
    double nb = dot3h(c[0], c[0]);
    for (int k = 1; k < 3; ++k) {
        double nk = dot3h(c[k], c[k]);
        if (nk > nb) { nb = nk; best = k; }
    }
    if (nb <= 0.0) { v[0] = 1.0; v[1] = 0.0; v[2] = 0.0; return; }
    double sc = 1.0 / sqrt(nb);
    for (int i = 0; i < 3; ++i) v[i] = c[best][i] * sc;
}

int mcgb_eigen_sym3(const double q[6], double evals[3], double evec[9])
{
    const double a00 = q[0], a11 = q[1], a22 = q[2], a01 = q[3], a02 = q[4], a12 = q[5];
    double p1 = a01 * a01 + a02 * a02 + a12 * a12;
    double m = (a00 + a11 + a22) / 3.0;
    double l[3];
    double d0 = a00 - m, d1 = a11 - m, d2 = a22 - m;
    double p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * p1;
    if (p2 <= 0.0) {
        l[0] = l[1] = l[2] = m;
    } else {
        double p = sqrt(p2 / 6.0);
        double b00 = d0 / p, b11 = d1 / p, b22 = d2 / p, b01 = a01 / p, b02 = a02 / p, b12 = a12 / p;
        double detb = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
                      b02 * (b01 * b12 - b11 * b02);
        double r = 0.5 * detb;
        if (r > 1.0) r = 1.0;
        if (r < -1.0) r = -1.0;
        double phi = acos(r) / 3.0;
        const double two_pi_3 = 2.0943951023931954923;
        l[2] = m + 2.0 * p * cos(phi);
        l[0] = m + 2.0 * p * cos(phi + two_pi_3);
        l[1] = 3.0 * m - l[0] - l[2];
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2 - i; ++j)
            if (l[j] > l[j + 1]) { double t = l[j]; l[j] = l[j + 1]; l[j + 1] = t; }
    evals[0] = l[0]; evals[1] = l[1]; evals[2] = l[2];
    if (!evec) return MCGB_OK;
    double V[3][3];   /* columns */
    if (p2 <= 0.0) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) evec[3 * i + j] = i == j ? 1.0 : 0.0;
        return MCGB_OK;
    }
    const double A[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
    int ext = (l[2] - l[1] >= l[1] - l[0]) ? 2 : 0;
    double ve[3];
    unit_vector_for(A, l[ext], ve);
    /* orthonormal basis (U, W) of the complement of ve */
    double U[3], W[3];
    if (fabs(ve[0]) > fabs(ve[1])) {
        double inv = 1.0 / sqrt(ve[0] * ve[0] + ve[2] * ve[2]);
        U[0] = -ve[2] * inv; U[1] = 0.0; U[2] = ve[0] * inv;
    } else {
        double inv = 1.0 / sqrt(ve[1] * ve[1] + ve[2] * ve[2]);
        U[0] = 0.0; U[1] = ve[2] * inv; U[2] = -ve[1] * inv;
    }
    cross3(ve, U, W);
    double AU[3], AW[3];
    for (int i = 0; i < 3; ++i) {
        AU[i] = A[i][0] * U[0] + A[i][1] * U[1] + A[i][2] * U[2];
        AW[i] = A[i][0] * W[0] + A[i][1] * W[1] + A[i][2] * W[2];
    }
    double m00 = dot3h(U, AU) - l[1], m01 = dot3h(U, AW), m11 = dot3h(W, AW) - l[1];
    double w0, w1;
    if (fabs(m00) >= fabs(m11)) {
        double n = sqrt(m00 * m00 + m01 * m01);
        if (n > 0.0) { w0 = -m01 / n; w1 = m00 / n; } else { w0 = 1.0; w1 = 0.0; }
    } else {
        double n = sqrt(m11 * m11 + m01 * m01);
        if (n > 0.0) { w0 = m11 / n; w1 = -m01 / n; } else { w0 = 1.0; w1 = 0.0; }
    }
    double vm[3], vo[3];
    for (int i = 0; i < 3; ++i) vm[i] = w0 * U[i] + w1 * W[i];
    if (ext == 2) {
        cross3(vm, ve, vo);           /* eigenvector of l[0] */
        for (int i = 0; i < 3; ++i) { V[i][2] = ve[i]; V[i][1] = vm[i]; V[i][0] = vo[i]; }
    } else {
        cross3(ve, vm, vo);           /* eigenvector of l[2] */
        for (int i = 0; i < 3; ++i) { V[i][0] = ve[i]; V[i][1] = vm[i]; V[i][2] = vo[i]; }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) evec[3 * i + j] = V[i][j];
    return MCGB_OK;
}

/* ------------------------------------------------------------------ fortSphere.1 */

int mcgb_write_fort1(const char *path, int n, const double *x, const double *y, const double *z,
                     const double *ux, const double *uy, const double *uz,
                     const double box[3], double colloid_radius)
{
    FILE *f = fopen(path, "w");
    if (!f) { mcgb_set_error("cannot open %s for writing", path); return MCGB_ERR_IO; }
    /* esc_fort1, MCnptColloid.c:3193-3210 */
    fprintf(f, "%d %15.8lf %15.8lf %15.8lf %15.8lf\n", n, box[0], box[1], box[2], colloid_radius);
    for (int i = 0; i < n; ++i) {
        fprintf(f, "%15.8lf %15.8lf %15.8lf\n", x[i], y[i], z[i]);
        fprintf(f, "%15.8lf %15.8lf %15.8lf\n", 0.0, 0.0, 0.0);
        fprintf(f, "%15.8lf %15.8lf %15.8lf\n", ux[i], uy[i], uz[i]);
        fprintf(f, "%15.8lf %15.8lf %15.8lf\n", 0.0, 0.0, 0.0);
    }
    if (fclose(f) != 0) { mcgb_set_error("error writing %s", path); return MCGB_ERR_IO; }
    return MCGB_OK;
}

int mcgb_read_fort1(const char *path, int n, double *x, double *y, double *z,
                    double *ux, double *uy, double *uz, double box[3], double *colloid_radius)
{
    FILE *f = fopen(path, "r");
    if (!f) { mcgb_set_error("cannot open %s", path); return MCGB_ERR_IO; }
    /* read_initial_configuration, MCnptColloid.c:2786-2831 */
    int ntmp = 0, nval = 0;
    double b;
    if (fscanf(f, "%d %lf %lf %lf %lf", &ntmp, &box[0], &box[1], &box[2], colloid_radius) == 5) nval += 5;
    if (ntmp < n) {
        fclose(f);
        mcgb_set_error("%s holds %d particles, %d requested", path, ntmp, n);
        return MCGB_ERR_IO;
    }
    for (int i = 0; i < n; ++i) {
        if (fscanf(f, "%lf %lf %lf", &x[i], &y[i], &z[i]) == 3) nval += 3;
        if (fscanf(f, "%lf %lf %lf", &b, &b, &b) == 3) nval += 3;
        if (fscanf(f, "%lf %lf %lf", &ux[i], &uy[i], &uz[i]) == 3) nval += 3;
        if (fscanf(f, "%lf %lf %lf", &b, &b, &b) == 3) nval += 3;
    }
    fclose(f);
    if (nval != 12 * n + 5) { mcgb_set_error("error reading %s", path); return MCGB_ERR_IO; }
    return MCGB_OK;
}

/* ------------------------------------------------------------------ trajectory writers */

static FILE *open_for(const char *path, int append)
{
    FILE *f = fopen(path, append ? "a" : "w");
    if (!f) mcgb_set_error("cannot open %s for %s", path, append ? "appending" : "writing");
    return f;
}

static int close_checked(FILE *f, const char *path)
{
    if (fclose(f) != 0) { mcgb_set_error("error writing %s", path); return MCGB_ERR_IO; }
    return MCGB_OK;
}

/* ------------------------------------------------------------------ resume record */

/* Layout (text): a header line "mcgb-resume 1 N", a line with the seed, sweep index, the
 * three adaptation counters and the two capped-term counts, a line with d_max, d_rot, the
 * finite running energy and the box as hexadecimal floating-point literals (%a, exact), then
 * one line per particle: qx qy qz as unsigned integers and ux uy uz as hexadecimal literals. */
int mcgb_write_resume(const char *path, int n, const uint32_t *qx, const uint32_t *qy,
                      const uint32_t *qz, const double *ux, const double *uy, const double *uz,
                      const double box[3], const mcgb_resume *rec)
{
    if (!path || n < 0 || !qx || !qy || !qz || !ux || !uy || !uz || !box || !rec) {
        mcgb_set_error("mcgb_write_resume: null argument or negative count");
        return MCGB_ERR_ARG;
    }
    FILE *f = open_for(path, 0);
    if (!f) return MCGB_ERR_IO;
    fprintf(f, "mcgb-resume 1 %d\n", n);
    fprintf(f, "%llu %llu %llu %llu %llu %lld %lld\n", (unsigned long long)rec->seed,
            (unsigned long long)rec->sweep_index, (unsigned long long)rec->adapt_attempted,
            (unsigned long long)rec->adapt_accepted, (unsigned long long)rec->adaptations,
            (long long)rec->capped_pairs, (long long)rec->capped_colloid);
    fprintf(f, "%a %a %a %a %a %a\n", (double)rec->max_displacement, (double)rec->max_rotation,
            rec->energy_finite, box[0], box[1], box[2]);
    for (int i = 0; i < n; ++i)
        fprintf(f, "%u %u %u %a %a %a\n", qx[i], qy[i], qz[i], ux[i], uy[i], uz[i]);
    return close_checked(f, path);
}

int mcgb_read_resume(const char *path, int n, uint32_t *qx, uint32_t *qy, uint32_t *qz,
                     double *ux, double *uy, double *uz, double box[3], mcgb_resume *rec)
{
    if (!path || n < 0 || !qx || !qy || !qz || !ux || !uy || !uz || !box || !rec) {
        mcgb_set_error("mcgb_read_resume: null argument or negative count");
        return MCGB_ERR_ARG;
    }
    FILE *f = fopen(path, "r");
    if (!f) { mcgb_set_error("cannot open %s", path); return MCGB_ERR_IO; }
    int version = 0, nf = -1, ok = 1;
    unsigned long long v[5];
    long long cp, cc;
    double dm, dr, ef;
    ok = fscanf(f, "mcgb-resume %d %d", &version, &nf) == 2 && version == 1;
    if (ok && nf != n) {
        fclose(f);
        mcgb_set_error("%s holds %d particles, %d requested", path, nf, n);
        return MCGB_ERR_IO;
    }
    ok = ok && fscanf(f, "%llu %llu %llu %llu %llu %lld %lld", &v[0], &v[1], &v[2], &v[3], &v[4],
                      &cp, &cc) == 7;
    ok = ok && fscanf(f, "%la %la %la %la %la %la", &dm, &dr, &ef, &box[0], &box[1], &box[2]) == 6;
    for (int i = 0; ok && i < n; ++i) {
        unsigned a, b, c;
        ok = fscanf(f, "%u %u %u %la %la %la", &a, &b, &c, &ux[i], &uy[i], &uz[i]) == 6;
        qx[i] = a; qy[i] = b; qz[i] = c;
    }
    fclose(f);
    if (!ok) { mcgb_set_error("error reading %s", path); return MCGB_ERR_IO; }
    rec->seed = v[0];
    rec->sweep_index = v[1];
    rec->adapt_attempted = v[2];
    rec->adapt_accepted = v[3];
    rec->adaptations = v[4];
    rec->max_displacement = (float)dm;
    rec->max_rotation = (float)dr;
    rec->capped_pairs = cp;
    rec->capped_colloid = cc;
    rec->energy_finite = ef;
    return MCGB_OK;
}

/* print_Trayectoria, MCnptColloid.c:3081-3107 (same layout, caller's path). */
int mcgb_write_trajectory(const char *path, int append, int n, const double *x, const double *y,
                          const double *z, const double *ux, const double *uy, const double *uz,
                          const double box[3])
{
    if (!path || n < 0 || (n > 0 && (!x || !y || !z || !ux || !uy || !uz)) || !box) {
        mcgb_set_error("mcgb_write_trajectory: bad arguments");
        return MCGB_ERR_ARG;
    }
    FILE *f = open_for(path, append);
    if (!f) return MCGB_ERR_IO;
    fprintf(f, "%d \n", n);
    fprintf(f, "%3.5lf %3.5lf %3.5lf \n", box[0], box[1], box[2]);
    for (int i = 0; i < n; ++i)
        fprintf(f, "%5.7lf \t %5.7lf \t %5.7lf \t %5.7lf \t %5.7lf \t %5.7lf \n",
                x[i], y[i], z[i], ux[i], uy[i], uz[i]);
    return close_checked(f, path);
}

/* print_Qmga, MCnptColloid.c:3110-3157: the qmga snapshot layout, positions relative to the
 * box centre, 12 columns then "0 0"; the reference's format string carries a 36-space gap
 * between column groups (a line splice inside the literal), reproduced here. */
int mcgb_write_qmga(const char *path, int append, int n, const double *x, const double *y,
                    const double *z, const double *ux, const double *uy, const double *uz,
                    const double box[3])
{
    if (!path || n < 0 || (n > 0 && (!x || !y || !z || !ux || !uy || !uz)) || !box) {
        mcgb_set_error("mcgb_write_qmga: bad arguments");
        return MCGB_ERR_ARG;
    }
    FILE *f = open_for(path, append);
    if (!f) return MCGB_ERR_IO;
    fprintf(f, "%d\n", n);
    fprintf(f, "%lf\n", box[0]);
    fprintf(f, "%lf\n", box[1]);
    fprintf(f, "%lf\n", box[2]);
    fprintf(f, "%lf %lf\n", 0.0, 0.0);
    static const char gap[] = "                                   ";   /* 35 spaces */
    for (int i = 0; i < n; ++i) {
        fprintf(f, "%15.8lf %15.8lf %15.8lf %s", x[i] - 0.5 * box[0], y[i] - 0.5 * box[1],
                z[i] - 0.5 * box[2], gap);
        fprintf(f, "%15.8lf %15.8lf %15.8lf %s", 0.0, 0.0, 0.0, gap);
        fprintf(f, "%15.8lf %15.8lf %15.8lf %s", ux[i], uy[i], uz[i], gap);
        fprintf(f, "%15.8lf %15.8lf %15.8lf 0 0\n", 0.0, 0.0, 0.0);
    }
    return close_checked(f, path);
}

/* print_pos_xyz, MCnptColloid.c:3160-3180: append = 0 starts the file with the two header
 * lines only (the reference's option 0); append = 1 adds one "H x y z" line per particle. */
int mcgb_write_xyz(const char *path, int append, int n, const double *x, const double *y,
                   const double *z, const double box[3])
{
    if (!path || n < 0 || (append && n > 0 && (!x || !y || !z)) || !box) {
        mcgb_set_error("mcgb_write_xyz: bad arguments");
        return MCGB_ERR_ARG;
    }
    FILE *f = open_for(path, append);
    if (!f) return MCGB_ERR_IO;
    if (!append) {
        fprintf(f, "%10d %15lf %15lf %15lf \n", n, box[0], box[1], box[2]);
        fprintf(f, "Hola, cara de bola  \n");
    } else {
        for (int i = 0; i < n; ++i) fprintf(f, "H %15lf %15lf %15lf \n", x[i], y[i], z[i]);
    }
    return close_checked(f, path);
}

/* ------------------------------------------------------------------ initial configurations */

/* Box of initialize_random_positions / initialize_cubic_lattice_positions
 * (MCnptColloid.c:2843-2846, :2932-2935): V = N/rho + 4/3 pi R^3, L = V^(1/3), cubic.  pi is
 * the reference's 14-digit literal (MCnptColloid.c:10), so the box agrees to the last bit. */
static double init_box(int n, double density, double colloid_radius)
{
    const double pi = 3.1415926535898;
    double vcoll = (4. / 3.) * pi * pow(colloid_radius, 3);
    return pow((double)n / density + vcoll, 1.0 / 3.0);
}

static void sphere_point(host_rng *g, double *ux, double *uy, double *uz)
{
    const double pi = 3.14159265358979323846;
    double zz = 1.0 - 2.0 * rng_u01(g), phi = 2.0 * pi * rng_u01(g);
    double sxy = sqrt(fmax(0.0, 1.0 - zz * zz));
    *ux = sxy * cos(phi);
    *uy = sxy * sin(phi);
    *uz = zz;
}

/* Minimum-image squared distance (compute_distance_sq, MCnptColloid.c:2690-2721). */
static double min_image_sq(double ax, double ay, double az, double bx, double by, double bz,
                           double L)
{
    double d[3] = {ax - bx, ay - by, az - bz};
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        if (d[k] > 0.5 * L) d[k] -= L;
        else if (d[k] < -0.5 * L) d[k] += L;
        r2 += d[k] * d[k];
    }
    return r2;
}

/* initial_configuration_type 1: initialize_random_positions (MCnptColloid.c:2835-2920).
 * Random sequential addition: a uniform point of the box is kept when no earlier particle lies
 * within sigma_ee (minimum image) and it is not inside the colloid (distance to the centre
 * >= R); orientations uniform on the sphere.  Philox-seeded (the reference uses rand());
 * after max_tries rejected points for one particle it stops with MCGB_ERR_GEOMETRY (the
 * reference loops forever).  A cell grid of width >= sigma_ee replaces the O(N^2) scan. */
int mcgb_init_random(int n, double density, double colloid_radius, double sigma_ee, uint64_t seed,
                     int max_tries, double *x, double *y, double *z, double *ux, double *uy,
                     double *uz, double box[3])
{
    if (n <= 0 || density <= 0.0 || sigma_ee < 0.0 || max_tries <= 0 || !x || !y || !z || !ux ||
        !uy || !uz || !box) {
        mcgb_set_error("mcgb_init_random: bad arguments");
        return MCGB_ERR_ARG;
    }
    const double L = init_box(n, density, colloid_radius);
    const double c = 0.5 * L, R2 = colloid_radius * colloid_radius, s2 = sigma_ee * sigma_ee;
    int nc = sigma_ee > 0.0 ? (int)floor(L / sigma_ee) : 1;
    if (nc < 3) nc = 1;                          /* small box: every cell is a neighbour */
    if (nc > 1024) nc = 1024;
    const size_t ncell = (size_t)nc * nc * nc;
    int *head = (int *)malloc(sizeof(int) * ncell);
    int *next = (int *)malloc(sizeof(int) * (size_t)n);
    if (!head || !next) {
        free(head); free(next);
        mcgb_set_error("mcgb_init_random: out of memory");
        return MCGB_ERR_NOMEM;
    }
    for (size_t k = 0; k < ncell; ++k) head[k] = -1;
    host_rng g = {{(uint32_t)seed, (uint32_t)(seed >> 32)}, 0u, {0, 0, 0, 0}, 4};
    for (int i = 0; i < n; ++i) {
        int tries = 0, ok = 0, ci = 0;
        double px = 0.0, py = 0.0, pz = 0.0;
        while (!ok) {
            if (tries++ >= max_tries) {
                free(head); free(next);
                mcgb_set_error("mcgb_init_random: no room for particle %d after %d tries "
                               "(density %g, sigma_ee %g)", i, max_tries, density, sigma_ee);
                return MCGB_ERR_GEOMETRY;
            }
            px = rng_u01(&g) * L; py = rng_u01(&g) * L; pz = rng_u01(&g) * L;
            ok = min_image_sq(px, py, pz, c, c, c, L) >= R2;
            int cx = (int)(px / L * nc), cy = (int)(py / L * nc), cz = (int)(pz / L * nc);
            cx = cx >= nc ? nc - 1 : cx; cy = cy >= nc ? nc - 1 : cy; cz = cz >= nc ? nc - 1 : cz;
            ci = cx + nc * (cy + nc * cz);
            const int span = nc == 1 ? 0 : 1;
            for (int dz = -span; dz <= span && ok; ++dz)
                for (int dy = -span; dy <= span && ok; ++dy)
                    for (int dx = -span; dx <= span && ok; ++dx) {
                        int ex = (cx + dx + nc) % nc, ey = (cy + dy + nc) % nc, ez = (cz + dz + nc) % nc;
                        for (int j = head[ex + nc * (ey + nc * ez)]; j >= 0 && ok; j = next[j])
                            if (min_image_sq(px, py, pz, x[j], y[j], z[j], L) < s2) ok = 0;
                    }
        }
        x[i] = px; y[i] = py; z[i] = pz;
        sphere_point(&g, &ux[i], &uy[i], &uz[i]);
        next[i] = head[ci];
        head[ci] = i;
    }
    free(head); free(next);
    box[0] = box[1] = box[2] = L;
    return MCGB_OK;
}

/* initial_configuration_type 2: initialize_cubic_lattice_positions (MCnptColloid.c:2924-3033).
 * n_s = smallest integer >= 2 with n_s^3 >= N; sites (i, j, k) * L/n_s filled x-fastest,
 * skipping sites within the colloid (distance to the centre <= R when R > 0); orientations
 * uniform on the sphere (Philox-seeded).  The reference resets the spacing to sigma_ee when
 * L/n_s < 1.1 sigma_ee (D20), which places sites outside the box; here the spacing stays
 * L/n_s, so every site is inside.  Running out of sites is MCGB_ERR_GEOMETRY. */
int mcgb_init_cubic(int n, double density, double colloid_radius, uint64_t seed,
                    double *x, double *y, double *z, double *ux, double *uy, double *uz,
                    double box[3])
{
    if (n <= 0 || density <= 0.0 || !x || !y || !z || !ux || !uy || !uz || !box) {
        mcgb_set_error("mcgb_init_cubic: bad arguments");
        return MCGB_ERR_ARG;
    }
    const double L = init_box(n, density, colloid_radius);
    const double c = 0.5 * L, R2 = colloid_radius * colloid_radius;
    int ns = 2;
    while ((long long)ns * ns * ns < n) ns++;
    const double delta = L / (double)ns;
    int ix = 0, iy = 0, iz = 0;
    host_rng g = {{(uint32_t)seed, (uint32_t)(seed >> 32)}, 0u, {0, 0, 0, 0}, 4};
    for (int i = 0; i < n; ++i) {
        for (;;) {
            if (iz >= ns) {
                mcgb_set_error("mcgb_init_cubic: %d^3 lattice sites outside the colloid hold "
                               "only %d particles", ns, i);
                return MCGB_ERR_GEOMETRY;
            }
            const double px = (double)ix * delta, py = (double)iy * delta, pz = (double)iz * delta;
            if (!(R2 > 0.0 && min_image_sq(px, py, pz, c, c, c, L) <= R2)) break;
            if (++ix == ns) { ix = 0; if (++iy == ns) { iy = 0; ++iz; } }
        }
        x[i] = (double)ix * delta;
        y[i] = (double)iy * delta;
        z[i] = (double)iz * delta;
        sphere_point(&g, &ux[i], &uy[i], &uz[i]);
        if (++ix == ns) { ix = 0; if (++iy == ns) { iy = 0; ++iz; } }
    }
    box[0] = box[1] = box[2] = L;
    return MCGB_OK;
}
