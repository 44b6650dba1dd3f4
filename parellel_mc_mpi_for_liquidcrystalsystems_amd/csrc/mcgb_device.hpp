// mcgb_device.hpp — device-side data layout, counter-based RNG and pair physics for the
// Gay-Berne checkerboard sweep on CDNA4 (gfx950).
//
// Positions are 32-bit fixed-point fractions of the box (x = q * L / 2^32): the minimum image
// of a pair is the wrap-around int32 difference of the q's, cell assignment is an exact integer
// multiply-shift, and an isotropic volume change leaves every q unchanged.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcgb {

constexpr int kMaxReplicas = 1024;
constexpr double kTwo32 = 4294967296.0;
constexpr uint32_t kSetupTag = 0xFFFF0000u;   // 4th counter word of the per-sweep draws

// ------------------------------------------------------------------ per-replica tables

// Static per-replica constants (uploaded by the host; read with scalar loads).
struct RepConst {
    int N;                 // particles
    int part_off;          // first particle slot in the particle arrays
    int cell_off;          // first cell slot (counts/accumulators), capacity ncell_cap
    int cs_off;            // first cell_start slot (ncell_cap + 1 entries)
    int ncap[3];           // grid laid out (cubic box, NPT allowance): its product is the
                           // cell capacity; a box's grid takes any shape within it (fit_grid)
    int act_off;           // first block of this replica in the pass grid
    int nact_cap;          // blocks reserved for this replica (= ncap product / col^3)
    uint32_t key0, key1;   // Philox key
    int colloid_on;        // colloid_wall_option == 1
    int npt;               // NPT ln V moves
    int rec_off;           // first per-move draw record (capacity N + ncell_cap)
    // physics, FP64 master copies (get_system_const, MCnptColloid.c:2752-2783)
    double sigma0, four_eps0, chi, chi2, chip, chip2, mu, nu, rc, rc2;
    double temperature, beta, pressure;   // step sizes live in RepDyn (adapted on the device)
    double R, R2, colloid_cut, kappa_prime, one_minus_chi, one_minus_chip;
    double npt_max_dlnv;
    int adapt_moves;       // step-size adaptation threshold (attempted moves), 0 = off
    int col;               // sublattice colours per axis (2 or 3): col^3 passes per sweep
    float min_thick;       // min(sigma_ee, sigma_ss) as the reference's float (:833-834)
    double vcoll;          // 4/3 pi R^3 (density convention, MCnptColloid.c:2799-2804)
    // FP32 copies for the pair math
    float f_sigma0, f_four_eps0, f_chi, f_chi2, f_chip, f_chip2, f_mu, f_half_nu, f_rc2;
    // g(om) = (c0 + c1 (|p|^2+|q|^2)/r^2 + om^2 (r.w)^2/r^2) / (d0 + om^2 |w|^2)
    float f_g1c0, f_g1c1, f_g1d0, f_g2c0, f_g2c1, f_g2d0;
    float f_rc2_pre;       // candidate pre-filter, slightly above rc^2 (final test in FP64)
    // exponents of eps as small integers: 0 general; mu + 2 (nu - 1) for mu, nu in {1, 2}
    // (1: mu = nu = 1, 2: mu = 2, nu = 1, 3: mu = 1, nu = 2, 4: mu = nu = 2)
    int f_ipow;
};

// Per-replica state updated on the device.
struct RepDyn {
    double L[3];                 // box
    double s[3];                 // L / 2^32
    int n[3];                    // cells per axis this sweep
    int ncell, nact, k;          // cells, active cells per pass, moves per active cell
    long long K[3];              // fixed-point step amplitude round(d_max 2^32 / L)
    uint32_t t;                  // sweep index
    uint32_t g[3];               // grid shift
    uint32_t order[27];          // sublattice order (col^3 entries)
    uint32_t vol[2];             // NPT draws
    long long cap_pairs, cap_coll;
    double e_fin;                // running energy = cap_pairs*1e9 + cap_coll*1e8 + e_fin
    unsigned long long attempted, accepted, left, hard, cand, pairs, coll, vol_att, vol_acc;
    double moved_abs;            // sum |finite dE| of accepted moves this sweep
    unsigned long long moved_q[8];   // the same per k_end block, rounded and clamped (1e15),
                                     // summed as integers into 8 words (blocks b mod 8, so that
                                     // the atomics spread) for the FP64 re-summation test
    double max_drift;
    int resum;                   // re-sum requested for this sweep
    int overflow;                // a stencil exceeded the LDS staging capacity
    int series_count;
    int cells_g0;                // 1: current cell list built with shift 0 for box L
    // NPT trial box
    double Lt[3];
    int nt[3];
    int trial_ok;
    int pad1;
    double e_res[4];             // last full recompute {cap_pairs, cap_coll, fin, hard}
    unsigned long long resums;   // re-summations of the running energy
    unsigned long long da_s1, da_s2;   // delayed-acceptance rejections, surrogate / exact stage
    float dmax, drot;            // step sizes (float, as the reference's globals :108-109)
    unsigned long long ad_att, ad_acc;   // moves since the last step-size adaptation
    unsigned long long adapts;   // step-size adaptations made
    double max_drift_forced;     // largest |running - recomputed| at a forced re-summation
    unsigned long long resums_forced;   // re-summations forced by a sweep's moved energy
    double qsum[6];              // sum over particles of u_a u_b (xx, yy, zz, xy, xz, yz) after
                                 // the passes of the last sweep (order-parameter record)
    int pace;                    // sequential pass: the pace of its cells' moves, 10 ns ticks
                                 // per move (from the last sweep; 0: no pacing)
    int pad2;
};

// The moved-energy total of a sweep (k_end's integer sums).
__host__ __device__ inline unsigned long long moved_total(const RepDyn& D) {
    unsigned long long t = 0;
    for (int k = 0; k < 8; ++k) t += D.moved_q[k];
    return t;
}

// Draws of one trial move (draw_move): particle selector, fixed-point step, rotation vector,
// Metropolis threshold -log(U) T, and the threshold of the surrogate stage of delayed
// acceptance -log(U') T (U' an independent uniform).
struct MoveDraw {
    uint32_t j;
    int32_t dq[3];
    double thr;
    double v[3];
    double thr1;
};

// FP32 constants of the pair function, held in registers by the pass kernel.
struct GBf {
    float sigma0, four_eps0, chi2, chip2, mu, half_nu;
    float g1c0, g1c1, g1d0, g2c0, g2c1, g2d0;
    double sigma0d;
    int ipow;            // exponent case of eps (see f_ipow)
};

// FP64 constants of the pair function.
struct GBd {
    double sigma0, four_eps0, chi, chi2, chip, mu, nu;
};

// Per-active-cell accumulators of one pass (each cell is active exactly once per sweep).
struct CellAcc {
    int att, acc, left, hard;
    int dcap, dccap, pairs, coll;
    long long cand;
    double dfin, moved_abs;
    int s1rej, s2rej;          // delayed acceptance: rejections by the surrogate / exact stage
    int ticks, tmoves;         // sequential pass: the cell's 10 ns ticks and moves (pacing)
};

// ------------------------------------------------------------------ Philox4x32-10

__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
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

__host__ __device__ inline double u01(uint32_t r) { return ((double)r + 0.5) * (1.0 / kTwo32); }

// Symmetric fixed-point step: trunc((2r - 2^32 + 1) * K / 2^33), uniform in (-K/2, K/2).
__host__ __device__ inline int32_t step_q(uint32_t r, long long K) {
    long long m = 2LL * (long long)r - 4294967295LL;
    long long p = m * K;
    return (int32_t)(p >= 0 ? (p >> 33) : -((-p) >> 33));
}

__host__ __device__ inline uint32_t cell_coord(uint32_t q, uint32_t g, uint32_t n) {
    return (uint32_t)(((uint64_t)(uint32_t)(q - g) * n) >> 32);
}

__host__ __device__ inline int cell_dim(double L, double rc) {
    int n = 2 * (int)floor(L / (2.0 * rc * (1.0 + 1e-6)));
    return n >= 4 ? n : 0;
}

// Sweep grid for col sublattice colours per axis: col = 2 -> cell_dim; col = 3 -> the largest
// multiple of 3 with L/n >= r_c/2 (n >= 6).  Cells of one colour are col - 1 cells (>= r_c)
// apart, and a move's partners lie within +-(col - 1) cells.
__host__ __device__ inline int grid_dim(double L, double rc, int col) {
    if (col != 3) return cell_dim(L, rc);
    int n = 3 * (int)floor(L / (1.5 * rc * (1.0 + 1e-6)));
    return n >= 6 ? n : 0;
}

// The cell arrays hold `cap` cells in all (the grid of the cubic box the context was laid out
// for, or of its NPT growth allowance).  A box whose own grid needs more (an edge longer than
// the layout planned for) coarsens its longest axis first, col cells at a time, down to at most
// cap cells: wider cells keep the checkerboard exact, and the box keeps its own grid wherever
// that fits.  Every axis stays at >= 2 col cells (2 col per axis fit: each allocated axis
// holds at least that many).
__host__ __device__ inline void fit_grid(int n[3], int cap, int col) {
    while (n[0] * n[1] * n[2] > cap) {
        int a = n[1] > n[0] ? 1 : 0;
        if (n[2] > n[a]) a = 2;
        n[a] -= col;
    }
}

// ------------------------------------------------------------------ GB pair, FP32

// Gay-Berne energy (pij_GB/g_omega, MCnptColloid.c:2607-2681) for unit orientation vectors,
// rewritten so that no step cancels.  With p = e_i x r, q = e_j x r, w = e_i x e_j
// (Lagrange and Binet-Cauchy identities: 1 - a^2 = |p|^2/r^2, c - ab = p.q/r^2, 1 - c^2 = |w|^2,
// |p x q|^2 = (r.w)^2 r^2):
//   g(om) = [ (1-om)^2 + om(1-om)(|p|^2 + |q|^2)/r^2 + om^2 (r.w)^2/r^2 ] / [ (1-om^2) + om^2 |w|^2 ],
// a ratio of sums of non-negative terms for 0 <= om < 1 (the reference's form subtracts two
// nearly equal numbers near end-to-end contact).  sigma = sigma0 sqrt(D/N) = sigma0 D rsq(N D),
// D = 1 - chi^2 c^2.  eps1^nu eps2^mu = (N'/D')^mu D^(-nu/2): for integer exponents
// mu, nu in {1, 2} (GB(3,5,2,1) among them) N'^mu rsq(D'^(2 mu) D^nu), one reciprocal square
// root; otherwise exp2(mu (log2 N' - log2 D') - (nu/2) log2 D).  The distance r comes from
// the exact FP64 r^2 (r2d) with one FP64 Newton step on the hardware rsqrt, and
// r - sigma + sigma0 is formed in FP64: the remaining error is that of sigma (a few FP32 ulps).
// Transcendentals: v_rsq x3, v_rcp x1 (general exponents: v_rsq x2, v_rcp x1, v_log x3,
// v_exp x1).  cap = the 1e9 cap.
// Written for Q pairs that share e_i, each step done for every pair before the next step, so
// that Q independent dependency chains interleave; Q = 1 is the single-pair function.  Every
// pair goes through the same operations whatever Q is.
#define MCGB_EACH for (int q = 0; q < Q; ++q)
template <int Q>
__device__ inline void gb_pair_q_f32(const GBf& C, float uix, float uiy, float uiz,
                                     const float* ujx, const float* ujy, const float* ujz,
                                     const float* dx, const float* dy, const float* dz,
                                     const double* r2d, float* u, bool* cap) {
    float rinv[Q], ir2[Q], wx[Q], wy[Q], wz[Q], w2[Q], tau[Q], ab[Q], tt[Q];
    float n1[Q], d1[Q], n2[Q], d2[Q], eps[Q], sigma[Q];
    double dist[Q];
#pragma unroll
    MCGB_EACH {
        const float r2 = (float)r2d[q];
        rinv[q] = __builtin_amdgcn_rsqf(r2);
        ir2[q] = rinv[q] * rinv[q];
    }
    // w = e_i x e_j, p = e_i x dr, q = e_j x dr
#pragma unroll
    MCGB_EACH {
        wx[q] = uiy * ujz[q] - uiz * ujy[q];
        wy[q] = uiz * ujx[q] - uix * ujz[q];
        wz[q] = uix * ujy[q] - uiy * ujx[q];
    }
#pragma unroll
    MCGB_EACH {
        const float px = uiy * dz[q] - uiz * dy[q], py = uiz * dx[q] - uix * dz[q],
                    pz = uix * dy[q] - uiy * dx[q];
        const float qx = ujy[q] * dz[q] - ujz[q] * dy[q], qy = ujz[q] * dx[q] - ujx[q] * dz[q],
                    qz = ujx[q] * dy[q] - ujy[q] * dx[q];
        w2[q] = wx[q] * wx[q] + wy[q] * wy[q] + wz[q] * wz[q];
        tau[q] = dx[q] * wx[q] + dy[q] * wy[q] + dz[q] * wz[q];
        ab[q] = (px * px + py * py + pz * pz + qx * qx + qy * qy + qz * qz) * ir2[q];
        tt[q] = tau[q] * tau[q] * ir2[q];
    }
    // shape g(chi) and energy g(chi') as N / D
#pragma unroll
    MCGB_EACH {
        n1[q] = C.g1c0 + C.g1c1 * ab[q] + C.chi2 * tt[q];
        d1[q] = C.g1d0 + C.chi2 * w2[q];
        n2[q] = C.g2c0 + C.g2c1 * ab[q] + C.chip2 * tt[q];
        d2[q] = C.g2d0 + C.chip2 * w2[q];
    }
#pragma unroll
    MCGB_EACH sigma[q] = C.sigma0 * d1[q] * __builtin_amdgcn_rsqf(n1[q] * d1[q]);
    // the exponent case is uniform (a scalar branch)
    const int ip = __builtin_amdgcn_readfirstlane(C.ipow);
    if (ip == 2) {                                   // mu = 2, nu = 1
#pragma unroll
        MCGB_EACH {
            const float d2s = d2[q] * d2[q];
            eps[q] = n2[q] * n2[q] * __builtin_amdgcn_rsqf(d2s * d2s * d1[q]);
        }
    } else if (ip != 0) {                            // mu, nu in {1, 2}
        const bool mu2 = ip == 4, nu2 = ip >= 3;
#pragma unroll
        MCGB_EACH {
            const float d2s = d2[q] * d2[q];
            const float num = mu2 ? n2[q] * n2[q] : n2[q];
            const float den = (mu2 ? d2s * d2s : d2s) * (nu2 ? d1[q] * d1[q] : d1[q]);
            eps[q] = num * __builtin_amdgcn_rsqf(den);
        }
    } else {
#pragma unroll
        MCGB_EACH eps[q] = __builtin_amdgcn_exp2f(
            C.mu * (__builtin_amdgcn_logf(n2[q]) - __builtin_amdgcn_logf(d2[q])) -
            C.half_nu * __builtin_amdgcn_logf(d1[q]));
    }
    // r in FP64: one Newton step on the FP32 reciprocal square root
#pragma unroll
    MCGB_EACH {
        const double y = (double)rinv[q];
        const double y1 = y * (1.5 - 0.5 * r2d[q] * y * y);
        dist[q] = (r2d[q] * y1 - (double)sigma[q]) + C.sigma0d;
        cap[q] = dist[q] < 0.001;
    }
#pragma unroll
    MCGB_EACH {
        const float rho = C.sigma0 * __builtin_amdgcn_rcpf((float)dist[q]);
        const float rho2 = rho * rho;
        const float rho6 = rho2 * rho2 * rho2;
        u[q] = C.four_eps0 * eps[q] * rho6 * (rho6 - 1.0f);
    }
}
#undef MCGB_EACH
__device__ inline float gb_pair_f32(const GBf& C, float uix, float uiy, float uiz,
                                    float ujx, float ujy, float ujz, float dx, float dy, float dz,
                                    double r2d, bool* cap) {
    float u;
    gb_pair_q_f32<1>(C, uix, uiy, uiz, &ujx, &ujy, &ujz, &dx, &dy, &dz, &r2d, &u, cap);
    return u;
}

// ------------------------------------------------------------------ GB pair, FP64

// pow with the exponents 1 and 2 taken as x and x*x (the correctly rounded value, which pow
// returns to within its 1-ulp bound); any other exponent goes through pow.
__device__ inline double pow_int_fast(double x, double e) {
    if (e == 1.0) return x;
    if (e == 2.0) return x * x;
    return pow(x, e);
}

__device__ inline double g_omega64(double c, double ap2, double am2, double w) {
    double oe = w * c;
    return 1.0 - 0.5 * w * (ap2 / (1.0 + oe) + am2 / (1.0 - oe));
}

// Term for term as pij_GB (MCnptColloid.c:2618-2681) with the inverse norms and cosine of GBij
// (:1524-1535); returns the finite energy, *cap = the 1e9 cap.
__device__ inline double gb_pair_f64(const GBd& C, double uix, double uiy, double uiz,
                                     double ujx, double ujy, double ujz, double dx, double dy,
                                     double dz, double r2, bool* cap) {
    double rij = sqrt(r2);
    double irij = 1.0 / rij;
    double iei = 1.0 / sqrt(uix * uix + uiy * uiy + uiz * uiz);
    double iej = 1.0 / sqrt(ujx * ujx + ujy * ujy + ujz * ujz);
    double c = iei * iej * (uix * ujx + uiy * ujy + uiz * ujz);
    double a = irij * iei * (dx * uix + dy * uiy + dz * uiz);
    double b = irij * iej * (dx * ujx + dy * ujy + dz * ujz);
    double ap = a + b, am = a - b;
    double ap2 = ap * ap, am2 = am * am;
    double eps1 = 1.0 / sqrt(1.0 - C.chi2 * c * c);
    double ctep1 = pow_int_fast(eps1, C.nu);
    double ctep2 = pow_int_fast(g_omega64(c, ap2, am2, C.chip), C.mu);
    double schi = g_omega64(c, ap2, am2, C.chi);
    double sigma = C.sigma0 / sqrt(schi);
    double dist = rij - sigma + C.sigma0;
    *cap = dist < 0.001;
    if (*cap) return 0.0;
    double e = C.sigma0 / dist;
    double e2 = e * e;
    double e6 = e2 * e2 * e2;
    return C.four_eps0 * (ctep1 * ctep2) * e6 * (e6 - 1.0);
}

// ------------------------------------------------------------------ colloid-rod, FP64

// GBCollRod (MCnptColloid.c:1758-1854); rel = rod centre - colloid centre.  Evaluated in FP64
// in both precisions, only near the colloid, and kept out of line so that the sweep kernel does
// not carry its registers.  cap = the 1e8 cap.  The result comes back by value, in registers
// (an out-parameter of a call that is not inlined would live in scratch memory); colloid_f64
// is the inline form with the cap as an out-parameter.
struct ColloidTerm {
    double e;
    int cap;
};
__device__ __attribute__((noinline)) ColloidTerm colloid_term(const RepConst& C, double rx, double ry,
                                                              double rz, double ex, double ey, double ez) {
    // the reference's expression with four divisions instead of nine: 1/d, 1/x and the two
    // ratios; the rest as products (rounding within a few FP64 ulps of the literal form)
    const double d = sqrt(rx * rx + ry * ry + rz * rz);
    const double dij = d - C.R;
    if (d <= 1e-12) return {0.0, 1};
    if (dij > C.colloid_cut) return {0.0, 0};
    const double id = 1.0 / d;
    double cth = (rx * ex + ry * ey + rz * ez) * id;
    cth = cth > 1.0 ? 1.0 : (cth < -1.0 ? -1.0 : cth);
    const double sin2 = 1.0 - cth * cth;
    const double s0 = C.sigma0;
    double den = C.one_minus_chi, num = 1.0 - C.chi * sin2;
    if (den <= 0.0) den = 1e-12;
    if (num <= 1e-12) num = 1e-12;
    const double scr = s0 * sqrt(num / den);
    double den_e = 1.0 - C.chip * sin2;
    if (den_e <= 1e-12) den_e = 1e-12;
    const double eps = C.kappa_prime * pow_int_fast(C.one_minus_chip / den_e, C.mu);
    const double x = dij - scr + s0;
    if (x <= 1e-10) return {0.0, 1};
    const double s3 = s0 * s0 * s0, s9 = s3 * s3 * s3;
    const double ix = 1.0 / x;
    const double ix2 = ix * ix, ix3 = ix2 * ix, ix8 = ix3 * ix3 * ix2, ix9 = ix8 * ix;
    return {eps * (s9 * ix9 * (1.0 / 45.0) - s3 * ix3 * (1.0 / 6.0) - s9 * ix8 * id * (1.0 / 40.0) +
                   s3 * ix2 * id * 0.25), 0};
}
__device__ inline double colloid_f64(const RepConst& C, double rx, double ry, double rz, double ex,
                                     double ey, double ez, bool* cap) {
    const ColloidTerm t = colloid_term(C, rx, ry, rz, ex, ey, ez);
    *cap = t.cap != 0;
    return t.e;
}

__host__ __device__ inline GBf make_gbf(const RepConst& C) {
    GBf g;
    g.sigma0 = C.f_sigma0; g.four_eps0 = C.f_four_eps0; g.chi2 = C.f_chi2; g.chip2 = C.f_chip2;
    g.mu = C.f_mu; g.half_nu = C.f_half_nu;
    g.g1c0 = C.f_g1c0; g.g1c1 = C.f_g1c1; g.g1d0 = C.f_g1d0;
    g.g2c0 = C.f_g2c0; g.g2c1 = C.f_g2c1; g.g2d0 = C.f_g2d0;
    g.sigma0d = C.sigma0;
    g.ipow = C.f_ipow;
    return g;
}

__host__ __device__ inline GBd make_gbd(const RepConst& C) {
    GBd g;
    g.sigma0 = C.sigma0; g.four_eps0 = C.four_eps0; g.chi = C.chi; g.chi2 = C.chi2;
    g.chip = C.chip; g.mu = C.mu; g.nu = C.nu;
    return g;
}

}  // namespace mcgb
