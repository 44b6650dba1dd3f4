// mcgb_pass.hpp — the sublattice-pass kernel: W wavefronts per active cell make the cell's k
// trial moves in sequence (the TODO body of MCnptColloid.c:683-700 filled with the move of
// attempt_monte_carlo_move, :2445-2604).
//
// Per activation: the stencil ((2R+1)^3 cell segments of the device cell list, R = 1 for 8
// sublattices, 2 for 27) is staged in LDS.  The moves run in chunks of 16: the chunk's draws
// (written by k_count_draws) arrive from HBM one chunk ahead, and 16 lanes prepare one record per
// move (draws, the particle's state, trial orientation, trial colloid term, the leave-cell and
// colloid-core tests, the pre-filter radius).  Per move: each
// wavefront tests its share of the stencil against the old and trial positions (FP32
// midpoint pre-filter), compacts the hits into a wavefront-private LDS list, evaluates both GB
// energies of every listed pair (cutoff decided on the FP64 r^2) two items per lane at a time,
// and reduces (DPP) the energy difference; the wavefront partials meet in LDS, every thread
// takes the same Metropolis decision, and an accepted move is committed to LDS (in the FP32
// sequential build by every wavefront, which then reads the next move's record and first scan
// group ahead of the closing barrier) and by thread 0 to HBM.  Two workgroup barriers per
// move.  Cells behind the pace set from the last sweep raise their issue priority (pacing).
// Stencils beyond the staging capacity read HBM directly.
#pragma once

#include <type_traits>

#include "mcgb_device.hpp"
#include "mcgb_surrogate.hpp"
#include "mcgb.h"

namespace mcgb {

// ------------------------------------------------------------------ DPP wave reductions

// Sum over the 64 lanes of a wavefront; the total lands in lane 63 (gfx9 DPP row/bank modes).
template <int CTRL, int ROWMASK>
__device__ inline int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWMASK, 0xF, false);
}
template <int CTRL, int ROWMASK>
__device__ inline float dpp_f(float v) {
    return __builtin_bit_cast(float, dpp_i<CTRL, ROWMASK>(__builtin_bit_cast(int, v)));
}
// FP64 halves moved with an undefined "old" operand: a lane that a step's row mask leaves out
// gets an arbitrary value, but such lanes never feed lane 63's chain (rows 0 and 2 in the two
// broadcast steps), and no move of a zero or a copy precedes each DPP move.
template <int CTRL, int ROWMASK>
__device__ inline double dpp_d(double v) {
    unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xFFFFFFFFull), CTRL, ROWMASK, 0xF, false);
    int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, ROWMASK, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

#define MCGB_DPP_STEPS(X) X(0xB1, 0xF) X(0x4E, 0xF) X(0x141, 0xF) X(0x140, 0xF) X(0x142, 0xA) X(0x143, 0xC)

__device__ inline int wave_sum_i63(int v) {
#define S_(c, m) v += dpp_i<c, m>(v);
    MCGB_DPP_STEPS(S_)
#undef S_
    return v;
}
__device__ inline float wave_sum_f63(float v) {
#define S_(c, m) v += dpp_f<c, m>(v);
    MCGB_DPP_STEPS(S_)
#undef S_
    return v;
}
__device__ inline double wave_sum_d63(double v) {
#define S_(c, m) v += dpp_d<c, m>(v);
    MCGB_DPP_STEPS(S_)
#undef S_
    return v;
}
// Wavefront-uniform values read from LDS or computed per lane: held in SGPRs from here on
// (frees VGPRs for occupancy).
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ inline float uni(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ inline double uni(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(u & 0xFFFFFFFFull));
    const int hi = __builtin_amdgcn_readfirstlane((int)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// ------------------------------------------------------------------ arguments and LDS layout

struct PassArgs {
    const RepConst* RC;
    RepDyn* RD;
    const int* blk_rep;
    uint32_t *qx, *qy, *qz;
    void *ux, *uy, *uz;
    const int* pid;
    const int* cell_start;
    CellAcc* acc;
    mcgb_move_trace* trace;
    const MoveDraw* draws;
    int* ovf_count;   // [8] cells of each pass left to the HBM kernel
    int* ovf_list;    // [8][nblk] their block indices
    int nblk;
    int pass;
    int cap;          // LDS staging capacity (particles)
    // delayed acceptance
    const MlpPacked* mlp;             // packed BF16 operands of the surrogate
    unsigned long long* da_hist;      // [replica][kDaBins + 2] surrogate error histogram
    int da_judge;                     // 1: exact stage evaluated for every move (diagnostics)
};

struct WavePartial {
    double d;         // finite dE (FP64 lane sums, both precisions)
    int capd;         // change of capped pairs
    int pad;
    double scale;     // sum |terms| (trace only)
};

// One trial move as prepared at the start of its chunk: its draws, and the particle's state at
// that point with the trial orientation and trial colloid term derived from it (valid unless an
// earlier accepted move of the chunk moved the same particle).  A move reads its record in one
// batch of LDS loads.
template <typename Real>
struct alignas(16) MovePrep {
    double thr, thr1;           // Metropolis thresholds: log U and (delayed acceptance) stage 1
    double en;                  // trial-state colloid term (cells near the colloid)
    uint32_t qx, qy, qz;        // position at the start of the chunk
    int32_t dqx, dqy, dqz;      // displacement, fixed point
    int32_t jsel;               // cell-local index of the particle
    uint32_t flags;             // bit q: move q of the chunk draws the same particle; bit 16: capped
                                // trial term; bit 17: the trial position leaves the cell; bit 18:
                                // it lies in the colloid core (tested when bit 17 is clear)
    Real ux, uy, uz;            // orientation at the start of the chunk
    Real nx, ny, nz;            // trial orientation from it
    Real vx, vy, vz;            // rotation draw
    float fr2;                  // squared pre-filter radius (fixed-point units for a cubic box)
};
constexpr uint32_t kPrepLeft = 1u << 17, kPrepHard = 1u << 18;

template <typename Real, int W>
struct PassLds {
    static constexpr int kDraws = 16;                // moves whose draws are staged at a time (16 x 7 8-byte words: two per lane of one wavefront)
    __host__ __device__ static inline int list_cap(int cap) {
        return ((cap + 64 * W - 1) / (64 * W)) * 64;
    }
    // staged positions: the capacity rounded up to whole wavefronts, the tail filled with a
    // far-away sentinel so that the scan needs no bounds test
    __host__ __device__ static inline int stage_cap(int cap) { return ((cap + 63) / 64) * 64; }
    // per-wavefront candidate list: list_cap entries and a slot that unselected lanes write
    __host__ __device__ static inline int list_stride(int cap) { return list_cap(cap) + 8; }
    // byte offsets: stencil tables 0, positions kHeader, orientations, lists, draws
    // (16-aligned); 12 * stage_cap is a multiple of 768, so every region keeps its alignment
    // s_start[126], s_off[126]: the positions begin below 1024 bytes, so a record's address
    // offsets fit the 8-bit dword fields of the paired LDS loads (no address add per load)
    static constexpr int kHeader = 1008;
    __host__ __device__ static inline int off_su(int cap) { return kHeader + 12 * stage_cap(cap); }
    __host__ __device__ static inline int off_list(int cap) {
        return off_su(cap) + cap * 3 * (int)sizeof(Real);
    }
    __host__ __device__ static inline int off_draws(int cap) {
        return (off_list(cap) + W * list_stride(cap) * 2 + 15) & ~15;
    }
    // inside the draws region: the prepared records of kDraws moves, the wavefront partials,
    // the surrogate partials, then the colloid terms of up to kCache cell particles at their
    // current states and their cap flags
    static constexpr int kCache = 64;
    static constexpr int kDrawBytes = kDraws * (int)sizeof(MovePrep<Real>);   // a multiple of 16
    static constexpr int kSurOff = (kDrawBytes + W * (int)sizeof(WavePartial) + 15) & ~15;
    static constexpr int kPrepOff = (kSurOff + W * (int)sizeof(float) + 15) & ~15;
    static constexpr int kPrepBytes = kCache * 8 + kCache;
    static constexpr int kRawOff = (kPrepOff + kPrepBytes + 15) & ~15;   // raw draws of a chunk
    static constexpr int kRawBytes = kDraws * (int)sizeof(MoveDraw);
    __host__ __device__ static inline size_t bytes(int cap) {
        return (size_t)off_draws(cap) + kRawOff + kRawBytes + 16;
    }
};

// Box scales and cutoff held in registers for the whole pass.
struct PairGeom {
    double s0, s1, s2, rc2;
};

// The pair constants moved into vector registers (one copy per lane): as scalars, together
// with the move's uniform values, they overflow the scalar register file, and the compiler
// then parks scalars in vector-register lanes, one v_writelane/v_readlane per spill and use.
template <typename T> __device__ inline void to_vgpr(T& x) { asm volatile("" : "+v"(x)); }
__device__ inline void pin_vgpr(GBf& G, PairGeom& X) {
    to_vgpr(G.sigma0); to_vgpr(G.four_eps0); to_vgpr(G.chi2); to_vgpr(G.chip2);
    to_vgpr(G.mu); to_vgpr(G.half_nu); to_vgpr(G.g1c0); to_vgpr(G.g1c1); to_vgpr(G.g1d0);
    to_vgpr(G.g2c0); to_vgpr(G.g2c1); to_vgpr(G.g2d0); to_vgpr(G.sigma0d);
    to_vgpr(X.s0); to_vgpr(X.s1); to_vgpr(X.s2); to_vgpr(X.rc2);
}
__device__ inline void pin_vgpr(GBd&, PairGeom&) {}

template <typename Real> struct GBConst;
template <> struct GBConst<float> { using type = GBf; };
template <> struct GBConst<double> { using type = GBd; };

__device__ inline GBf gb_const(const RepConst& C, float*) { return make_gbf(C); }
__device__ inline GBd gb_const(const RepConst& C, double*) { return make_gbd(C); }

// Pair energy of the moved particle (orientation ui) and stencil particle j at fixed-point
// separation (ox, oy, oz) = q_j - q_i.  Returns false when the pair is beyond the cutoff (the
// decision is taken on the FP64 r^2 in both precisions).
__device__ inline bool pair_energy(const GBf& G, const PairGeom& X, float uix, float uiy, float uiz,
                                   float ujx, float ujy, float ujz, int ox, int oy, int oz, float* u,
                                   bool* cap) {
    const double dx = (double)ox * X.s0, dy = (double)oy * X.s1, dz = (double)oz * X.s2;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < X.rc2)) return false;
    *u = gb_pair_f32(G, uix, uiy, uiz, ujx, ujy, ujz, (float)dx, (float)dy, (float)dz, r2, cap);
    return true;
}
__device__ inline bool pair_energy(const GBd& G, const PairGeom& X, double uix, double uiy,
                                   double uiz, double ujx, double ujy, double ujz, int ox, int oy,
                                   int oz, double* u, bool* cap) {
    const double dx = (double)ox * X.s0, dy = (double)oy * X.s1, dz = (double)oz * X.s2;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < X.rc2)) return false;
    *u = gb_pair_f64(G, uix, uiy, uiz, ujx, ujy, ujz, dx, dy, dz, r2, cap);
    return true;
}

// The same without the branch: the term is computed for every pair, the result says whether it
// counts (within the cutoff).  Lets two pairs' dependency chains interleave.
__device__ inline bool pair_energy_all(const GBf& G, const PairGeom& X, float uix, float uiy,
                                       float uiz, float ujx, float ujy, float ujz, int ox, int oy,
                                       int oz, float* u, bool* cap) {
    const double dx = (double)ox * X.s0, dy = (double)oy * X.s1, dz = (double)oz * X.s2;
    const double r2 = dx * dx + dy * dy + dz * dz;
    *u = gb_pair_f32(G, uix, uiy, uiz, ujx, ujy, ujz, (float)dx, (float)dy, (float)dz, r2, cap);
    return r2 < X.rc2;
}
__device__ inline bool pair_energy_all(const GBd& G, const PairGeom& X, double uix, double uiy,
                                       double uiz, double ujx, double ujy, double ujz, int ox,
                                       int oy, int oz, double* u, bool* cap) {
    const double dx = (double)ox * X.s0, dy = (double)oy * X.s1, dz = (double)oz * X.s2;
    const double r2 = dx * dx + dy * dy + dz * dz;
    *u = gb_pair_f64(G, uix, uiy, uiz, ujx, ujy, ujz, dx, dy, dz, r2, cap);
    return r2 < X.rc2;
}

// Two pairs of the same moved particle (orientation ui) at once: in FP32 their dependency
// chains interleave (gb_pair_q_f32<2>); in FP64 they are evaluated one after the other.
__device__ inline void pair_energy_two(const GBf& G, const PairGeom& X, float uix, float uiy,
                                       float uiz, const float* ujx, const float* ujy,
                                       const float* ujz, const int* ox, const int* oy,
                                       const int* oz, float* u, bool* cap, bool* in) {
    float fx[2], fy[2], fz[2];
    double r2[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const double dx = (double)ox[q] * X.s0, dy = (double)oy[q] * X.s1, dz = (double)oz[q] * X.s2;
        r2[q] = dx * dx + dy * dy + dz * dz;
        in[q] = r2[q] < X.rc2;
        fx[q] = (float)dx; fy[q] = (float)dy; fz[q] = (float)dz;
    }
    gb_pair_q_f32<2>(G, uix, uiy, uiz, ujx, ujy, ujz, fx, fy, fz, r2, u, cap);
}
__device__ inline void pair_energy_two(const GBd& G, const PairGeom& X, double uix, double uiy,
                                       double uiz, const double* ujx, const double* ujy,
                                       const double* ujz, const int* ox, const int* oy,
                                       const int* oz, double* u, bool* cap, bool* in) {
    for (int q = 0; q < 2; ++q)
        in[q] = pair_energy_all(G, X, uix, uiy, uiz, ujx[q], ujy[q], ujz[q], ox[q], oy[q], oz[q],
                                &u[q], &cap[q]);
}

// Per-lane accumulator of a finite energy difference, reduced across lanes in a fixed DPP order
// and across wavefronts in a fixed order: deterministic.
template <typename Real> struct LaneAcc;
// FP32 pair terms: each lane sums its few terms in FP64 and the wavefront reduces them in FP64,
// so a move's dE carries no summation error beyond the pair terms' own FP32 rounding (and
// near-cap terms up to ~1e36 cannot overflow).  FP64 adds run at the rate of unpacked FP32.
template <> struct LaneAcc<float> {
    double s = 0.0;
    __device__ void add(float v) { s += (double)v; }
    __device__ void to_partial(WavePartial& p) { p.d = wave_sum_d63(s); }
    __device__ static double total(const WavePartial* p, int w) {
        double s = 0.0;
        for (int k = 0; k < w; ++k) s += p[k].d;
        return s;
    }
};
template <> struct LaneAcc<double> {
    double s = 0.0;
    __device__ void add(double v) { s += v; }
    __device__ void to_partial(WavePartial& p) {
        p.d = wave_sum_d63(s);
    }
    __device__ static double total(const WavePartial* p, int w) {
        double s = 0.0;
        for (int k = 0; k < w; ++k) s += p[k].d;
        return s;
    }
};

// Trial orientation normalize(u + d_rot v) (MCnptColloid.c:2526-2550; kept when |.|^2 <= 1e-12).
// Products and sums are not fused into FMAs, so that every kernel, wherever it computes the
// trial orientation, rounds it the same way (in FP64, as the oracle's plain double arithmetic).
__device__ inline void rotate(double drot, float ux, float uy, float uz, float vx, float vy,
                              float vz, float& ox, float& oy, float& oz) {
#pragma clang fp contract(off)
    const float d = (float)drot;
    const float tx = ux + d * vx, ty = uy + d * vy, tz = uz + d * vz;
    const float n2 = tx * tx + ty * ty + tz * tz;
    if (n2 <= 1e-12f) { ox = ux; oy = uy; oz = uz; return; }
    float rn = __builtin_amdgcn_rsqf(n2);
    rn = rn * (1.5f - 0.5f * n2 * rn * rn);          // one Newton step
    ox = tx * rn; oy = ty * rn; oz = tz * rn;
}
__device__ inline void rotate(double drot, double ux, double uy, double uz, double vx, double vy,
                              double vz, double& ox, double& oy, double& oz) {
#pragma clang fp contract(off)
    const double tx = ux + drot * vx, ty = uy + drot * vy, tz = uz + drot * vz;
    const double n2 = tx * tx + ty * ty + tz * tz;
    if (n2 <= 1e-12) { ox = ux; oy = uy; oz = uz; return; }
    const double nr = sqrt(n2);
    ox = tx / nr; oy = ty / nr; oz = tz / nr;
}

// Candidate pre-filter of a move: sphere around the midpoint of the old and trial positions
// with radius sqrt(rc2p) + |dq|/2 (slightly enlarged for FP32 rounding), FP32 distances.
// For a cubic box (one scale on every axis) the test runs in fixed-point units, r2u = r2/s^2,
// one multiply per axis fewer.
struct MidFilter {
    uint32_t mx, my, mz;
    float sx, sy, sz, r2, r2u;
    bool iso;
    __device__ inline bool near(uint32_t qx, uint32_t qy, uint32_t qz) const {
        const float fx = (float)(int)(qx - mx) * sx, fy = (float)(int)(qy - my) * sy,
                    fz = (float)(int)(qz - mz) * sz;
        return fx * fx + fy * fy + fz * fz < r2;
    }
    template <typename Q>
    __device__ inline bool near(const Q& q, std::false_type) const { return near(q.x, q.y, q.z); }
    template <typename Q>
    __device__ inline bool near(const Q& q, std::true_type) const {
        const float fx = (float)(int)(q.x - mx), fy = (float)(int)(q.y - my), fz = (float)(int)(q.z - mz);
        return fx * fx + fy * fy + fz * fz < r2u;
    }
};

// Rank of this lane among the set lanes of a ballot below it.
__device__ inline int lane_rank(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Staged records: one address per particle, the components at fixed offsets.
struct Pos3 { uint32_t x, y, z; };
template <typename Real> struct Vec3 { Real x, y, z; };
// Per-pass constants of the filter: sqrt(rc2p) and 1/s^2 for the cubic-box test.
struct FilterConst {
    float sx, sy, sz, rc_root, inv_s2;
    bool iso;
};
__device__ inline FilterConst filter_const(float sx, float sy, float sz, float rc2p) {
    FilterConst F;
    F.sx = sx; F.sy = sy; F.sz = sz;
    F.rc_root = sqrtf(rc2p);
    F.inv_s2 = 1.0f / (sx * sx);
    F.iso = sx == sy && sy == sz;
    return F;
}
// The hardware square root (<= 1 ulp) is covered by the (1 + 1e-6) factors.  Per lane: the
// squared radius r2 and (cubic box) the same in fixed-point units r2u, rounded up by one part
// in 1e6 (conservative).
__device__ inline void filter_radius2(int dqx, int dqy, int dqz, const FilterConst& F, float* r2,
                                      float* r2u) {
    const float hx = (float)dqx * F.sx, hy = (float)dqy * F.sy, hz = (float)dqz * F.sz;
    const float rm = (F.rc_root + 0.5f * __builtin_amdgcn_sqrtf(hx * hx + hy * hy + hz * hz)) *
                         (1.0f + 1e-6f) + 2.0f * (F.sx + F.sy + F.sz);
    *r2 = rm * rm;
    *r2u = rm * rm * F.inv_s2 * (1.0f + 1e-6f);
}
// The filter of a move (wavefront-uniform inputs; the radius held as a scalar).
__device__ inline MidFilter mid_filter(uint32_t qix, uint32_t qiy, uint32_t qiz, int dqx, int dqy,
                                       int dqz, const FilterConst& F) {
    MidFilter M;
    M.mx = qix + (uint32_t)(dqx >> 1);
    M.my = qiy + (uint32_t)(dqy >> 1);
    M.mz = qiz + (uint32_t)(dqz >> 1);
    M.sx = F.sx; M.sy = F.sy; M.sz = F.sz;
    float r2, r2u;
    filter_radius2(dqx, dqy, dqz, F, &r2, &r2u);
    M.r2 = uni(r2);
    M.iso = F.iso;
    M.r2u = uni(r2u);
    return M;
}

// The same filter with its squared radius already known (r2u for a cubic box, else r2).
__device__ inline MidFilter mid_filter_r(uint32_t qix, uint32_t qiy, uint32_t qiz, int dqx,
                                         int dqy, int dqz, const FilterConst& F, float fr2) {
    MidFilter M;
    M.mx = qix + (uint32_t)(dqx >> 1);
    M.my = qiy + (uint32_t)(dqy >> 1);
    M.mz = qiz + (uint32_t)(dqz >> 1);
    M.sx = F.sx; M.sy = F.sy; M.sz = F.sz;
    M.r2 = fr2;
    M.r2u = fr2;
    M.iso = F.iso;
    return M;
}

// A move's outcome tests from the particle's state (q) and displacement: kPrepLeft when the
// trial position leaves the active cell (cx, cy, cz) of the grid (n, shift g), else kPrepHard
// when it lies in the colloid core (cells within reach of the colloid: MCnptColloid.c:2516-2520);
// *fr2 = the squared pre-filter radius of mid_filter (r2u for a cubic box, else r2).  Per lane
// (the chunk's records are prepared one move per lane).
__device__ inline uint32_t prep_tests(uint32_t qx, uint32_t qy, uint32_t qz, int dqx, int dqy,
                                      int dqz, uint32_t gx, uint32_t gy, uint32_t gz, int nx,
                                      int ny, int nz, int cx, int cy, int cz, bool cnear,
                                      double R2, const PairGeom& X, const FilterConst& F,
                                      float* fr2) {
    const uint32_t qnx = qx + (uint32_t)dqx, qny = qy + (uint32_t)dqy, qnz = qz + (uint32_t)dqz;
    float r2, r2u;
    filter_radius2(dqx, dqy, dqz, F, &r2, &r2u);
    *fr2 = F.iso ? r2u : r2;
    if (cell_coord(qnx, gx, (uint32_t)nx) != (uint32_t)cx ||
        cell_coord(qny, gy, (uint32_t)ny) != (uint32_t)cy ||
        cell_coord(qnz, gz, (uint32_t)nz) != (uint32_t)cz)
        return kPrepLeft;
    if (cnear) {
        const double rx = (double)(int32_t)(qnx - 0x80000000u) * X.s0;
        const double ry = (double)(int32_t)(qny - 0x80000000u) * X.s1;
        const double rz = (double)(int32_t)(qnz - 0x80000000u) * X.s2;
        if (rx * rx + ry * ry + rz * rz <= R2) return kPrepHard;
    }
    return 0u;
}

// Flat stencil index e -> global particle slot (segments of the ns neighbour cells).
__device__ inline int stencil_slot(const int* s_off, const int* s_start, int part_off, int e,
                                   int ns = 27) {
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s_off[mid] <= e) lo = mid; else hi = mid;
    }
    return part_off + s_start[lo] + (e - s_off[lo]);
}

// GLOBAL = false: one workgroup per active cell, stencil staged in LDS; a cell whose stencil
// exceeds the staging capacity is appended to the pass's overflow list and left untouched.
// GLOBAL = true: the workgroups walk that overflow list and read the stencil from HBM.
// TR = true: every trial move is recorded in P.trace (with the sum of |pair terms|).
// DA = true: delayed acceptance (Christen & Fox 2005).  Stage 1 accepts with
// min(1, exp(-beta dE~)), dE~ = the surrogate's GB pair terms; only then is the exact GB energy
// evaluated, and stage 2 accepts with min(1, exp(-beta (dE - dE~))), dE the exact energy change
// (GB pairs, caps and colloid term).  With a zero network stage 1 always passes and stage 2 is
// the plain Metropolis test with the plain sweep's draw: the plain chain, colloid or not.
// The surrogate orders each pair by stencil slot (lower slot first), a fixed function of the
// configuration for the whole move, so the composite kernel keeps detailed balance exactly.
// pass_cell makes the moves of active-cell block b (every early return is uniform over the
// workgroup); k_pass runs one block per workgroup, or (GLOBAL) walks the overflow list.
template <typename Real, int W, bool GLOBAL, bool TR, bool DA, bool O4 = false>
__device__ __forceinline__ void pass_cell(const PassArgs& P, int b, unsigned char* smem) {
    static_assert(!(DA && (GLOBAL || TR)), "delayed acceptance runs in the staged, untraced pass");
    const int r = P.blk_rep[b];
    const RepConst& C = P.RC[r];
    RepDyn& D = P.RD[r];
    const int a = b - C.act_off;
    const int nact = D.nact;
    if (a >= nact) return;                    // uniform: whole block
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = uni(tid >> 6);             // wavefront index (uniform)
    const int cap = P.cap;
    Real* __restrict__ Ux = (Real*)P.ux;
    Real* __restrict__ Uy = (Real*)P.uy;
    Real* __restrict__ Uz = (Real*)P.uz;
    using L = PassLds<Real, W>;

    int* s_start = (int*)smem;                                 // 126 (<= 125 used)
    int* s_off = s_start + 126;                                // 126 (<= 126 used)
    using R3 = Vec3<Real>;
    Pos3* sq = (Pos3*)(smem + L::kHeader);                     // stage_cap positions
    R3* su = (R3*)(smem + L::off_su(cap));                     // cap orientations
    unsigned short* slist = (unsigned short*)(smem + L::off_list(cap)) + w * L::list_stride(cap);
    unsigned char* dp = smem + L::off_draws(cap);
    constexpr int KD = L::kDraws;                              // moves per draw chunk
    MovePrep<Real>* sprep = (MovePrep<Real>*)dp;               // KD prepared moves
    WavePartial* sred = (WavePartial*)(dp + L::kDrawBytes);
    float* ssur = (float*)(dp + L::kSurOff);
    double* scache = (double*)(dp + L::kPrepOff);              // kCache: current colloid terms
    unsigned char* scc = (unsigned char*)(scache + L::kCache); // kCache: cap flags of scache
    uint2* sdraw2 = (uint2*)(dp + L::kRawOff);                 // KD draws as read from HBM
    MlpRegs MR;
    if (DA) mlp_load(P.mlp, lane, MR);

    typename GBConst<Real>::type G = gb_const(C, (Real*)nullptr);
    PairGeom X = {D.s[0], D.s[1], D.s[2], C.rc2};
    // not in the 128- and 256-register builds; with eight wavefronts per cell the registers
    // cost the second resident workgroup per CU (measured)
    if constexpr (!O4 && !DA && W <= 4) pin_vgpr(G, X);
    // the delayed-acceptance build has room for the box scales only (250 of 256 registers)
    if constexpr (DA && std::is_same<Real, float>::value) {
        to_vgpr(X.s0); to_vgpr(X.s1); to_vgpr(X.s2); to_vgpr(X.rc2);
    }
    // the state arrays as the accepted moves write them (addresses in vector registers in the
    // sequential builds, for the same reason; they fit the four-wavefront budget too)
    uint32_t* wq[3] = {P.qx, P.qy, P.qz};
    Real* wu[3] = {Ux, Uy, Uz};
    if constexpr (!DA && W <= 4 && std::is_same<Real, float>::value)
        for (int d = 0; d < 3; ++d) { to_vgpr(wq[d]); to_vgpr(wu[d]); }
    const int part_off = C.part_off;
    const int colloid_on = C.colloid_on;
    const double R2 = C.R2;
    const double drot = (double)D.drot;
    const int nx = D.n[0], ny = D.n[1], nz = D.n[2];
    const int col = C.col;
    const int hx = nx / col, hy = ny / col;
    const uint32_t sig = D.order[P.pass];
    const int ax = a % hx, ay = (a / hx) % hy, az = a / (hx * hy);
    const int cx = col * ax + (int)(sig % (uint32_t)col),
              cy = col * ay + (int)((sig / (uint32_t)col) % (uint32_t)col),
              cz = col * az + (int)(sig / (uint32_t)(col * col));
    const int cell = cx + nx * (cy + ny * cz);
    const int* cs = P.cell_start + C.cs_off;
    const uint32_t gx = D.g[0], gy = D.g[1], gz = D.g[2];

    // the draws of the cell's moves (k_count_draws), KD moves (KD * 56 contiguous bytes) at a time,
    // spread over wavefront 0 as 8-byte words: the first chunk's loads are issued here, ahead of
    // the staging, and every later chunk's while the chunk before it runs
    const int k = D.k;
    const size_t tbase = ((size_t)P.pass * (size_t)nact + (size_t)a) * (size_t)k;
    const MoveDraw* __restrict__ drw = P.draws + C.rec_off + tbase;
    const uint2* __restrict__ drw2 = (const uint2*)drw;
    constexpr int kW2 = (int)(sizeof(MoveDraw) / 8);          // 8-byte words per draw
    static_assert(sizeof(MoveDraw) % 8 == 0 && KD * kW2 <= 128, "a chunk's draws fit two words per lane");
    uint2 dpre = make_uint2(0u, 0u), dpre_b = make_uint2(0u, 0u);   // words lane, lane + 64
    if (w == 0 && lane < min(k, KD) * kW2) dpre = drw2[lane];
    if (w == 0 && lane + 64 < min(k, KD) * kW2) dpre_b = drw2[lane + 64];

    // pacing of the sequential pass: the cell timed on the 100 MHz clock from the start of its
    // workgroup (a slow staging counts against it), the pace set by the last sweep's end fold
    const unsigned long long pace0 = (!GLOBAL && !TR && !DA && !O4) ? __builtin_amdgcn_s_memrealtime() : 0ull;
    // stencil table: NS = (2R+1)^3 neighbour cells in (dz, dy, dx) order, the active cell at
    // NS/2; two cells per lane, one inclusive scan per half
    const int RS = col - 1, S1 = 2 * RS + 1, NS = S1 * S1 * S1;
    if (w == 0) {
        int cnt[2] = {0, 0};
        for (int h = 0; h < 2; ++h) {
            const int c = lane + 64 * h;
            if (c < NS) {
                int ex = cx + (c % S1) - RS, ey = cy + ((c / S1) % S1) - RS, ez = cz + c / (S1 * S1) - RS;
                ex += ex < 0 ? nx : 0; ex -= ex >= nx ? nx : 0;
                ey += ey < 0 ? ny : 0; ey -= ey >= ny ? ny : 0;
                ez += ez < 0 ? nz : 0; ez -= ez >= nz ? nz : 0;
                const int nb = ex + nx * (ey + ny * ez);
                const int st = cs[nb];
                cnt[h] = cs[nb + 1] - st;
                s_start[c] = st;
            }
        }
        int i0 = cnt[0], i1 = cnt[1];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v0 = __shfl_up(i0, off), v1 = __shfl_up(i1, off);
            if (lane >= off) { i0 += v0; i1 += v1; }
        }
        const int tot0 = __shfl(i0, 63);
        if (lane < NS) s_off[lane + 1] = i0;
        if (lane + 64 < NS) s_off[lane + 65] = tot0 + i1;
        if (lane == 0) s_off[0] = 0;
    }
    __syncthreads();
    const int S = uni(s_off[NS]);
    const int coff = s_off[NS / 2];
    const int nc = s_off[NS / 2 + 1] - coff;
    const int cstart = part_off + s_start[NS / 2];
    constexpr bool staged = !GLOBAL;
    if (!GLOBAL && S > cap) {                  // uniform: leave this cell to the HBM kernel
        if (tid == 0) {
            atomicOr(&D.overflow, 1);
            const int slot = atomicAdd(&P.ovf_count[P.pass], 1);
            P.ovf_list[(size_t)P.pass * P.nblk + slot] = b;
        }
        return;
    }
    // the stencil padded to whole wavefronts with a sentinel: the point opposite the centre of
    // the active cell (half a box away on every axis).  With >= 4 cells of >= r_c per axis (8
    // sublattices) it lies >= 1.5 cells (>= 1.5 r_c) from any cell member per axis, with >= 6
    // cells of >= r_c/2 (27 sublattices) >= 2.5 cells (>= 1.25 r_c): in both cases beyond every
    // move's filter radius (<= r_c + half a cell diagonal, trial positions being confined to
    // the cell).
    const int S_pad = (S + 63) & ~63;
    if (staged) {
        // kB records per thread per step: the slot searches, then every HBM load, then the LDS
        // stores (6 kB loads in flight per thread; with W = 2, 768 records in one step)
        constexpr int kB = W == 2 ? 6 : 4;
        for (int e0 = tid; e0 < S; e0 += kB * 64 * W) {
            int gi[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const int e = e0 + j * 64 * W;
                gi[j] = e < S ? stencil_slot(s_off, s_start, part_off, e, NS) : -1;
            }
            uint32_t x[kB], y[kB], z[kB];
            Real a0[kB], a1[kB], a2[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j)
                if (gi[j] >= 0) {
                    x[j] = P.qx[gi[j]]; y[j] = P.qy[gi[j]]; z[j] = P.qz[gi[j]];
                    a0[j] = Ux[gi[j]]; a1[j] = Uy[gi[j]]; a2[j] = Uz[gi[j]];
                }
#pragma unroll
            for (int j = 0; j < kB; ++j)
                if (gi[j] >= 0) {
                    const int e = e0 + j * 64 * W;
                    sq[e] = Pos3{x[j], y[j], z[j]};
                    su[e] = R3{a0[j], a1[j], a2[j]};
                }
        }
        const uint32_t fx = gx + (uint32_t)(((uint64_t)(2 * cx + 1) << 31) / (uint64_t)nx) + 0x80000000u;
        const uint32_t fy = gy + (uint32_t)(((uint64_t)(2 * cy + 1) << 31) / (uint64_t)ny) + 0x80000000u;
        const uint32_t fz = gz + (uint32_t)(((uint64_t)(2 * cz + 1) << 31) / (uint64_t)nz) + 0x80000000u;
        for (int e = S + tid; e < S_pad; e += 64 * W) sq[e] = Pos3{fx, fy, fz};
    }

    constexpr bool tracing = TR;
    int n_att = 0, n_acc = 0, n_left = 0, n_hard = 0, n_coll = 0, n_s1 = 0, n_s2 = 0;
    int dcap_sum = 0, dccap_sum = 0;
    int my_pairs = 0;
    long long n_cand = 0;
    double dfin_sum = 0.0, moved = 0.0;
    const float sx = (float)X.s0, sy = (float)X.s1, sz = (float)X.s2;
    const float rc2p = C.f_rc2_pre;     // pre-filter; the cutoff itself is decided in FP64
    // colloid-rod terms only matter within R + 4 sigma_ee of the colloid centre
    const double creach = C.R + C.colloid_cut + 0.5;
    const double creach2 = creach * creach;
    const FilterConst FC = filter_const(sx, sy, sz, rc2p);
    // colloid tests only in cells that reach within creach of the box centre: the distance of
    // the cell centre minus half the cell diagonal (with a margin) bounds every position of
    // the cell, old and trial (trial positions are confined to the cell)
    bool cnear = false;
    if (colloid_on) {
        const double wx = D.L[0] / nx, wy = D.L[1] / ny, wz = D.L[2] / nz;
        const double ccx = (double)(int32_t)(gx + (uint32_t)(((uint64_t)(2 * cx + 1) << 31) / (uint64_t)nx) - 0x80000000u) * X.s0;
        const double ccy = (double)(int32_t)(gy + (uint32_t)(((uint64_t)(2 * cy + 1) << 31) / (uint64_t)ny) - 0x80000000u) * X.s1;
        const double ccz = (double)(int32_t)(gz + (uint32_t)(((uint64_t)(2 * cz + 1) << 31) / (uint64_t)nz) - 0x80000000u) * X.s2;
        const double dc = sqrt(ccx * ccx + ccy * ccy + ccz * ccz);
        const double hd = 0.5 * sqrt(wx * wx + wy * wy + wz * wz);
        cnear = dc - hd <= creach * (1.0 + 1e-9) + 1e-6;
    }
    auto rel = [](uint32_t q, double s) { return (double)(int32_t)(q - 0x80000000u) * s; };
    // Cells near the colloid keep the colloid term of each of their particles at its current
    // state in LDS (set here, replaced on acceptance), so a move evaluates only its trial
    // state's term, and that one is prepared with the chunk's draws, one move per lane.  The
    // term is exactly 0 beyond the colloid cutoff, so evaluating it everywhere in such cells
    // gives the same sums as testing the reach first.
    const bool ccache = staged && cnear && nc <= L::kCache;
    // the cell's staged records come from every wavefront's staging loads: a barrier before
    // the cache reads them (uniform: ccache is the workgroup's)
    if (ccache) __syncthreads();
    if (ccache)
        for (int j = tid; j < nc; j += 64 * W) {
            const Pos3 qj = sq[coff + j];
            const R3 uj = su[coff + j];
            bool cj;
            scache[j] = colloid_f64(C, rel(qj.x, X.s0), rel(qj.y, X.s1), rel(qj.z, X.s2),
                                    (double)uj.x, (double)uj.y, (double)uj.z, &cj);
            scc[j] = cj ? 1 : 0;
        }

    const int pace = (!GLOBAL && !TR && !DA && !O4) ? uni(D.pace) : 0;
    if (nc == 0) {
        if (tracing)
            for (int m = tid; m < k; m += 64 * W) {
                mcgb_move_trace rec = {-1, MCGB_MOVE_EMPTY, 0, 0, 0.0, 0.0};
                P.trace[tbase + m] = rec;
            }
    } else {
        for (int m0 = 0; m0 < k; m0 += KD) {
            __syncthreads();
            int myj = -1;
            MovePrep<Real> mp = {};
            if (w == 0) {
                // this chunk's draws into LDS (read back below by the same wavefront), the
                // next chunk's loads issued
                if (lane < min(k - m0, KD) * kW2) sdraw2[lane] = dpre;
                if (lane + 64 < min(k - m0, KD) * kW2) sdraw2[lane + 64] = dpre_b;
                if (lane < min(k - m0 - KD, KD) * kW2) dpre = drw2[(size_t)(m0 + KD) * kW2 + lane];
                if (lane + 64 < min(k - m0 - KD, KD) * kW2)
                    dpre_b = drw2[(size_t)(m0 + KD) * kW2 + lane + 64];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (tid < KD && m0 + tid < k) {   // the draws of up to KD moves
                const MoveDraw d = ((const MoveDraw*)sdraw2)[tid];
                const int js = (int)(((uint64_t)d.j * (uint32_t)nc) >> 32);
                myj = js;
                mp.thr = d.thr;
                mp.thr1 = d.thr1;
                mp.dqx = d.dq[0]; mp.dqy = d.dq[1]; mp.dqz = d.dq[2];
                mp.jsel = js;
                mp.vx = (Real)d.v[0]; mp.vy = (Real)d.v[1]; mp.vz = (Real)d.v[2];
                if (staged) {
                    // the particle's state at the start of the chunk, the move's trial
                    // orientation (and trial colloid term) from it
                    const Pos3 qj = sq[coff + js];
                    const R3 uj = su[coff + js];
                    mp.qx = qj.x; mp.qy = qj.y; mp.qz = qj.z;
                    mp.ux = uj.x; mp.uy = uj.y; mp.uz = uj.z;
                    rotate(drot, uj.x, uj.y, uj.z, mp.vx, mp.vy, mp.vz, mp.nx, mp.ny, mp.nz);
                    if (ccache) {
                        bool cn;
                        mp.en = colloid_f64(C, rel(qj.x + (uint32_t)d.dq[0], X.s0),
                                            rel(qj.y + (uint32_t)d.dq[1], X.s1),
                                            rel(qj.z + (uint32_t)d.dq[2], X.s2), (double)mp.nx,
                                            (double)mp.ny, (double)mp.nz, &cn);
                        mp.flags = cn ? 1u << 16 : 0u;
                    }
                    // the move's outcome tests and pre-filter radius, from the same state
                    mp.flags |= prep_tests(qj.x, qj.y, qj.z, d.dq[0], d.dq[1], d.dq[2], gx, gy, gz,
                                           nx, ny, nz, cx, cy, cz, cnear, R2, X, FC, &mp.fr2);
                }
            }
            if (w == 0) {
                // bit q: move q of the chunk (q earlier than this one) draws the same particle
                uint32_t dup = 0;
#pragma unroll
                for (int q = 0; q < KD - 1; ++q) {
                    const int jq = __shfl(myj, q);
                    if (q < tid && jq == myj) dup |= 1u << q;
                }
                mp.flags |= dup;
                if (tid < KD) sprep[tid] = mp;
            }
            __syncthreads();
            uint32_t acc_mask = 0;            // moves of this chunk accepted so far
            const int mend = min(k, m0 + KD);
            // FP32 sequential build: the next move's record read before the barrier that ends
            // a move (the records do not change during the chunk), its LDS latency overlapping
            // the barrier (with the group below, 168 VGPRs: still three wavefronts per SIMD; the
            // other builds, and eight wavefronts per cell, have no room for them)
            constexpr bool kPre = std::is_same<Real, float>::value && !O4 && !DA && W <= 4;
            MovePrep<Real> pm_next;
            if (kPre) pm_next = sprep[0];
            // and, read at the same point, the first position group of the next move's scan
            // (group w: it does not depend on the move); each wavefront writes an accepted
            // move's new position itself before reading it
            const int ngs = S_pad >> 6;
            Pos3 pre0;
            if (kPre && staged) pre0 = sq[min(w, ngs - 1) * 64 + lane];
            for (int m = m0; m < mend; ++m) {
                const int ml = m - m0;
                // the whole record in one batch of LDS loads (reading fields where they are
                // used puts LDS latency on the move's critical path)
                const MovePrep<Real> pm = kPre ? pm_next : sprep[ml];
                const int jsel = uni(pm.jsel);
                // the chunk's preparation of this move holds unless an earlier accepted move
                // of the chunk moved the same particle
                const uint32_t fl = uni(pm.flags);
                const bool fresh = staged && (fl & acc_mask & 0xffffu) == 0u;
                const int li = coff + jsel;          // flat stencil index of the moved particle
                const int gsel = cstart + jsel;      // its global slot
                uint32_t qix, qiy, qiz;
                Real uix, uiy, uiz;
                if (fresh) {
                    // the particle's state and the step stay in vector registers (every lane
                    // read the same record): they feed vector arithmetic only, and as scalars
                    // they cost a readfirstlane and, for use by the lanes, a move back each
                    qix = pm.qx; qiy = pm.qy; qiz = pm.qz;
                    uix = pm.ux; uiy = pm.uy; uiz = pm.uz;
                } else if (staged) {
                    const Pos3 qi = sq[li];
                    const R3 ui = su[li];
                    qix = qi.x; qiy = qi.y; qiz = qi.z;
                    uix = ui.x; uiy = ui.y; uiz = ui.z;
                } else {
                    qix = P.qx[gsel]; qiy = P.qy[gsel]; qiz = P.qz[gsel];
                    uix = Ux[gsel]; uiy = Uy[gsel]; uiz = Uz[gsel];
                }
                const int dqx = pm.dqx, dqy = pm.dqy, dqz = pm.dqz;
                const uint32_t qnx = qix + (uint32_t)dqx, qny = qiy + (uint32_t)dqy,
                               qnz = qiz + (uint32_t)dqz;
                int outcome;
                int capd = 0, ccd = 0;
                double dfin = 0.0, scale = 0.0;
                ++n_att;
                // the outcome tests and the pre-filter radius: from the chunk's record, or
                // again from the current state
                float fr2;
                const uint32_t tests =
                    fresh ? fl & (kPrepLeft | kPrepHard)
                          : uni(prep_tests(qix, qiy, qiz, dqx, dqy, dqz, gx, gy, gz, nx, ny, nz, cx,
                                           cy, cz, cnear, R2, X, FC, &fr2));
                if (fresh) fr2 = pm.fr2;
                if (tests & kPrepLeft) {
                    outcome = MCGB_MOVE_LEFT;       // moves are confined to the active cell
                    ++n_left;
                } else if (tests & kPrepHard) {
                    outcome = MCGB_MOVE_HARD;       // MCnptColloid.c:2516-2520
                    ++n_hard;
                } else {
                    const double rnx = (double)(int32_t)(qnx - 0x80000000u) * X.s0;
                    const double rny = (double)(int32_t)(qny - 0x80000000u) * X.s1;
                    const double rnz = (double)(int32_t)(qnz - 0x80000000u) * X.s2;
                    Real unx, uny, unz;
                    if (fresh) {
                        unx = pm.nx; uny = pm.ny; unz = pm.nz;
                    } else {
                        rotate(drot, uix, uiy, uiz, pm.vx, pm.vy, pm.vz, unx, uny, unz);
                    }
                    LaneAcc<Real> acc;
                    Real scl = 0;
                    int cpl = 0;
                    int base = 0;                        // this wavefront's listed candidates
                    if (staged) {
                        // candidates of both states, wavefront-private compaction: a partner
                        // within the cutoff of either state lies within rc + |dq|/2 of their
                        // midpoint (triangle inequality), so one FP32 distance per candidate
                        const MidFilter M = mid_filter_r(qix, qiy, qiz, dqx, dqy, dqz, FC, fr2);
                        // up to four wavefront-wide groups of 64 records per step (this
                        // wavefront's groups w, w + W, ...), their LDS loads issued before the
                        // tests; the padded tail holds sentinels the filter rejects
                        auto scan = [&](auto iso) {
                            constexpr int kG = 4;
                            const int ng = S_pad >> 6;
                            for (int g0 = w; g0 < ng; g0 += kG * W) {
                                // uniform; taken through an opaque scalar move so that the
                                // compiler does not hoist the group tests out of the move loop
                                // as lane masks (they overflow the scalar registers and come
                                // back by v_readlane, one per group and move)
                                int nv;
                                asm volatile("s_mov_b32 %0, %1" : "=s"(nv) : "s"(min(kG, (ng - g0 + W - 1) / W)));
                                // every load issued unconditionally (groups past the end read
                                // the last one again): no branch between them, so they are in
                                // flight together
                                Pos3 p[kG];
                                if (kPre && g0 == w) {          // uniform
                                    p[0] = pre0;
#pragma unroll
                                    for (int j = 1; j < kG; ++j)
                                        p[j] = sq[min(g0 + j * W, ng - 1) * 64 + lane];
                                } else {
#pragma unroll
                                    for (int j = 0; j < kG; ++j)
                                        p[j] = sq[min(g0 + j * W, ng - 1) * 64 + lane];
                                }
#pragma unroll
                                for (int j = 0; j < kG; ++j)
                                    if (j < nv) {
                                        const int e = (g0 + j * W) * 64 + lane;
                                        // the moved particle itself passes the filter (it lies
                                        // within |dq|/2 of the midpoint): listed, and dropped
                                        // where its items are evaluated
                                        const bool kk = M.near(p[j], iso);
                                        const unsigned long long mk = __builtin_amdgcn_ballot_w64(kk);
                                        // the selected lanes write (under the compare's mask)
                                        if (kk) slist[base + lane_rank(mk)] = (unsigned short)e;
                                        base += __popcll(mk);
                                    }
                            }
                        };
                        if (M.iso) scan(std::true_type{});
                        else scan(std::false_type{});
                        // the list is private to this wavefront: a wavefront-scope fence suffices
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    }
                    // colloid-rod term of both states in FP64, only near the colloid
                    double cacc = 0.0, cscl = 0.0;
                    double cen = 0.0;                 // the trial state's term (for the cache)
                    bool ccn = false;
                    if (colloid_on) n_coll += 2;
                    if (ccache) {
                        const double eo = scache[jsel];
                        const bool co = scc[jsel] != 0;
                        if (fresh) {
                            cen = pm.en;
                            ccn = (fl >> 16) & 1u;
                        } else {
                            cen = colloid_f64(C, rnx, rny, rnz, (double)unx, (double)uny,
                                              (double)unz, &ccn);
                        }
                        ccd = (ccn ? 1 : 0) - (co ? 1 : 0);
                        cacc = (ccn ? 0.0 : cen) - (co ? 0.0 : eo);
                        cscl = fabs(ccn ? 0.0 : cen) + fabs(co ? 0.0 : eo);
                    } else if (cnear) {
                        const double rox = (double)(int32_t)(qix - 0x80000000u) * X.s0;
                        const double roy = (double)(int32_t)(qiy - 0x80000000u) * X.s1;
                        const double roz = (double)(int32_t)(qiz - 0x80000000u) * X.s2;
                        const bool nearc = rox * rox + roy * roy + roz * roz <= creach2 ||
                                           rnx * rnx + rny * rny + rnz * rnz <= creach2;
                        if (nearc) {
                            bool co, cn;
                            const double eo = colloid_f64(C, rox, roy, roz, (double)uix, (double)uiy,
                                                          (double)uiz, &co);
                            const double en = colloid_f64(C, rnx, rny, rnz, (double)unx, (double)uny,
                                                          (double)unz, &cn);
                            ccd = (cn ? 1 : 0) - (co ? 1 : 0);
                            cacc = (cn ? 0.0 : en) - (co ? 0.0 : eo);
                            cscl = fabs(cn ? 0.0 : en) + fabs(co ? 0.0 : eo);
                        }
                    }
                    // delayed acceptance, stage 1: the surrogate over this wavefront's list,
                    // 32 (slot, state) items per matrix-core tile
                    bool exact = true, stage1 = true;
                    double sur = 0.0;
                    if (DA) {
                        float sacc = 0.0f;
                        const int items = 2 * base;
                        for (int t0 = 0; t0 < items; t0 += 32) {
                            const int it = t0 + (lane & 31);
                            bool on = false;
                            float fr = 0.f, fa = 0.f, fb = 0.f, fc = 0.f;
                            if (lane < 32 && it < items) {
                                const int e = slist[it >> 1];
                                const bool nw = it & 1;
                                const Pos3 pe = sq[e];
                                const double dx = (double)(int)(pe.x - (nw ? qnx : qix)) * X.s0;
                                const double dy = (double)(int)(pe.y - (nw ? qny : qiy)) * X.s1;
                                const double dz = (double)(int)(pe.z - (nw ? qnz : qiz)) * X.s2;
                                const double r2 = dx * dx + dy * dy + dz * dz;
                                if (r2 < X.rc2 && e != li) {
                                    on = true;
                                    const float ex = (float)(nw ? unx : uix), ey = (float)(nw ? uny : uiy),
                                                ez = (float)(nw ? unz : uiz);
                                    const R3 ue = su[e];
                                    const float jx = (float)ue.x, jy = (float)ue.y, jz = (float)ue.z;
                                    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
                                    fr = sqrtf((float)r2);
                                    const float ir = 1.0f / fr;
                                    float a = (fx * ex + fy * ey + fz * ez) * ir;
                                    float bb = (fx * jx + fy * jy + fz * jz) * ir;
                                    if (li > e) {             // pair seen from the lower slot
                                        const float ta = a;
                                        a = -bb;
                                        bb = -ta;
                                    }
                                    fa = a; fb = bb;
                                    fc = ex * jx + ey * jy + ez * jz;
                                }
                            }
                            const float y = mlp_tile32(MR, mlp_input(fr, fa, fb, fc, on, lane));
                            if (on) sacc += (it & 1) ? y : -y;
                        }
                        sacc = wave_sum_f63(sacc);
                        if (lane == 63) ssur[w] = sacc;
                        __syncthreads();
                        for (int q = 0; q < W; ++q) sur += (double)ssur[q];
                        stage1 = sur <= pm.thr1;
                        exact = stage1 || P.da_judge;
                    }
                    if (staged && exact) {
                        // items it and it + 64 per lane and step (the same state: 64 is even),
                        // both evaluated without a branch so that their dependency chains
                        // interleave; each lane adds its terms in item order, as one item per
                        // step would.  A missing second item repeats the first and is dropped.
                        const int items = 2 * base;
                        // a lane's items all hold one state (item parity = lane parity): its
                        // capped items are counted, and signed once after the loop
                        int ncap = 0;
                        for (int it = lane; it < items; it += 128) {
                            const bool hb = it + 64 < items;
                            const int ea = slist[it >> 1], eb = slist[(hb ? it + 64 : it) >> 1];
                            const bool nw = it & 1;
                            const uint32_t px = nw ? qnx : qix, py = nw ? qny : qiy, pz = nw ? qnz : qiz;
                            const Real ox = nw ? unx : uix, oy = nw ? uny : uiy, oz = nw ? unz : uiz;
                            const Pos3 pa = sq[ea], pb = sq[eb];
                            const R3 ua = su[ea], ub = su[eb];
                            const Real jx[2] = {ua.x, ub.x}, jy[2] = {ua.y, ub.y}, jz[2] = {ua.z, ub.z};
                            const int dxs[2] = {(int)(pa.x - px), (int)(pb.x - px)};
                            const int dys[2] = {(int)(pa.y - py), (int)(pb.y - py)};
                            const int dzs[2] = {(int)(pa.z - pz), (int)(pb.z - pz)};
                            Real v2[2];
                            bool c2[2], in2[2];
                            pair_energy_two(G, X, ox, oy, oz, jx, jy, jz, dxs, dys, dzs, v2, c2, in2);
                            // selects, not branches: a branch would let the compiler sink
                            // each pair's chain into its own block (adding 0 leaves a sum as is)
                            const bool ia = in2[0] & (ea != li), ib = in2[1] & hb & (eb != li);
                            const bool fa = ia & !c2[0], fb = ib & !c2[1];
                            const Real ta = fa ? (nw ? v2[0] : -v2[0]) : (Real)0;
                            const Real tb = fb ? (nw ? v2[1] : -v2[1]) : (Real)0;
                            my_pairs += (int)ia + (int)ib;
                            ncap += (int)(ia & c2[0]) + (int)(ib & c2[1]);
                            acc.add(ta);
                            acc.add(tb);
                            if (TR) scl += (ta < 0 ? -ta : ta) + (tb < 0 ? -tb : tb);
                        }
                        cpl += (lane & 1) ? ncap : -ncap;
                    } else if (!staged) {
                        // HBM path (stencil beyond the staging capacity): both states per
                        // candidate, no compaction
                        for (int e = tid; e < S; e += 64 * W) {
                            if (e == li) continue;
                            const int gj = stencil_slot(s_off, s_start, part_off, e, NS);
                            const uint32_t jx = P.qx[gj], jy = P.qy[gj], jz = P.qz[gj];
                            const Real ujx = Ux[gj], ujy = Uy[gj], ujz = Uz[gj];
                            for (int st = 0; st < 2; ++st) {
                                const bool nw = st == 1;
                                bool cp;
                                Real u;
                                if (pair_energy(G, X, nw ? unx : uix, nw ? uny : uiy, nw ? unz : uiz,
                                                ujx, ujy, ujz, (int)(jx - (nw ? qnx : qix)),
                                                (int)(jy - (nw ? qny : qiy)), (int)(jz - (nw ? qnz : qiz)),
                                                &u, &cp)) {
                                    ++my_pairs;
                                    if (cp) cpl += nw ? 1 : -1;
                                    else {
                                        acc.add(nw ? u : -u);
                                        if (TR) scl += u < 0 ? -u : u;
                                    }
                                }
                            }
                        }
                    }
                    double dpair = 0.0;
                    n_cand += 2LL * (S - 1);
                    if (exact) {
                        // wavefront partials -> LDS
                        WavePartial part;
                        acc.to_partial(part);
                        // capped pairs are rare once a start's overlaps are gone: their count is
                        // reduced only when some lane holds one
                        part.capd = __builtin_amdgcn_ballot_w64(cpl != 0) ? wave_sum_i63(cpl) : 0;
                        part.scale = tracing ? wave_sum_d63((double)scl) : 0.0;
                        if (lane == 63) sred[w] = part;
                        __syncthreads();
                        dpair = LaneAcc<Real>::total(sred, W);
                        capd = 0;
                        scale = cscl;
                        for (int q = 0; q < W; ++q) {
                            capd += sred[q].capd;
                            scale += sred[q].scale;
                        }
                    }
                    dfin = dpair + cacc;
                    bool accept;
                    if (DA) {
                        // stage 2: the exact energy change against the surrogate's
                        accept = stage1 &&
                                 (double)capd * 1.0e9 + (double)ccd * 1.0e8 + (dfin - sur) <= pm.thr;
                        if (exact && capd == 0 && tid == 0) {
                            const float x = (float)((sur - dpair) / C.temperature) / kDaRange;
                            int bin = (int)floorf((x + 1.0f) * (0.5f * kDaBins));
                            bin = bin < 0 ? kDaBins : (bin >= kDaBins ? kDaBins + 1 : bin);
                            atomicAdd(&P.da_hist[(size_t)r * (kDaBins + 2) + bin], 1ull);
                        }
                        if (!stage1) ++n_s1;
                        else if (!accept) ++n_s2;
                    } else {
                        // log U <= -beta dE (MCnptColloid.c:2574)
                        accept = (double)capd * 1.0e9 + (double)ccd * 1.0e8 + dfin <= pm.thr;
                    }
                    // every lane computed the same decision: held as a scalar, the branches
                    // on it (and on what it feeds, the chunk's accepted-move mask) are uniform
                    accept = uni((int)accept) != 0;
                    if (accept) {
                        outcome = MCGB_MOVE_ACCEPT;
                        ++n_acc;
                        dcap_sum += capd;
                        dccap_sum += ccd;
                        dfin_sum += dfin;
                        moved += fabs(dfin);
                        acc_mask = uni(acc_mask | (1u << ml));
                        // the staged state: by every wavefront when it reads the next move's
                        // first scan groups ahead of the barrier (the same bytes)
                        if (staged && (kPre ? lane == 0 : tid == 0)) {
                            sq[li] = Pos3{qnx, qny, qnz};
                            su[li] = R3{unx, uny, unz};
                        }
                        if (tid == 0) {
                            if (ccache) {
                                scache[jsel] = cen;
                                scc[jsel] = ccn ? 1 : 0;
                            }
                            wq[0][gsel] = qnx; wq[1][gsel] = qny; wq[2][gsel] = qnz;
                            wu[0][gsel] = unx; wu[1][gsel] = uny; wu[2][gsel] = unz;
                        }
                    } else {
                        outcome = MCGB_MOVE_REJECT;
                    }
                }
                if (tracing && tid == 0) {
                    mcgb_move_trace rec;
                    rec.pid = P.pid[gsel];
                    rec.outcome = outcome;
                    rec.dcap = capd;
                    rec.dccap = ccd;
                    rec.dfin = dfin;
                    rec.scale = scale;
                    P.trace[tbase + m] = rec;
                }
                if (kPre && m + 1 < mend) pm_next = sprep[ml + 1];
                if (kPre && staged) pre0 = sq[min(w, ngs - 1) * 64 + lane];
                if (pace > 0) {
                    // pacing: a cell behind the pace (m + 1 moves done by (m + 1) * pace ticks
                    // after the first) takes issue priority for its next move over the cells
                    // sharing its SIMDs, in steps of half a move: the slowest cells set the pass
                    // time, the others have slack.  The clock is read before the barrier, so
                    // its latency overlaps the barrier.  Scheduling only: the chain is the same
                    // at every priority.
                    const long long el = (long long)(__builtin_amdgcn_s_memrealtime() - pace0);
                    const long long bh = el - (long long)(m + 1) * pace;
                    if (bh > pace / 2) __builtin_amdgcn_s_setprio(3);
                    else if (bh > 0) __builtin_amdgcn_s_setprio(2);
                    else if (bh > -pace / 2) __builtin_amdgcn_s_setprio(1);
                    else __builtin_amdgcn_s_setprio(0);
                }
                __syncthreads();
            }
        }
    }
    // per-cell accumulators (pairs counted per lane)
    const int wp = wave_sum_i63(my_pairs);
    __syncthreads();
    if (lane == 63) sred[w].capd = wp;
    __syncthreads();
    if (tid == 0) {
        int np = 0;
        for (int q = 0; q < W; ++q) np += sred[q].capd;
        CellAcc ca;
        ca.att = n_att; ca.acc = n_acc; ca.left = n_left; ca.hard = n_hard;
        ca.dcap = dcap_sum; ca.dccap = dccap_sum; ca.pairs = np; ca.coll = n_coll;
        ca.cand = n_cand; ca.dfin = dfin_sum; ca.moved_abs = moved;
        ca.s1rej = n_s1; ca.s2rej = n_s2;
        ca.ticks = (!GLOBAL && !TR && !DA && !O4 && nc > 0)
                       ? (int)(__builtin_amdgcn_s_memrealtime() - pace0) : 0;
        ca.tmoves = (!GLOBAL && !TR && !DA && !O4 && nc > 0) ? k : 0;
        P.acc[C.cell_off + cell] = ca;
    }
}

// Three wavefronts per SIMD, no more and no fewer (at most 168 VGPRs): with more allowed, the
// 1331 workgroups of a pass at N = 256k (W = 2) run 10% slower (measured), and the FP32 build
// with its prefetched record and scan group would otherwise take 170 VGPRs (two per SIMD);
// schedule 2 (k_pass_o4) is the four-wavefront build for workloads that gain from it.
template <typename Real, int W, bool GLOBAL, bool TR, bool DA = false>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_pass(PassArgs P) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (GLOBAL) {
        const int nov = P.ovf_count[P.pass];
        for (int j = blockIdx.x; j < nov; j += gridDim.x) {
            pass_cell<Real, W, GLOBAL, TR, DA>(P, P.ovf_list[(size_t)P.pass * P.nblk + j], smem);
            __syncthreads();                  // the next entry reuses the LDS tables
        }
    } else {
        pass_cell<Real, W, GLOBAL, TR, DA>(P, blockIdx.x, smem);
    }
}

// The delayed-acceptance pass compiled for at least two wavefronts per SIMD (at most 256
// registers, VGPRs and matrix-core accumulators together): with W = 1 every active cell of
// the 256k grid is then resident at once.
template <typename Real, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2))) void k_pass_da(PassArgs P) {
    extern __shared__ __align__(16) unsigned char smem[];
    pass_cell<Real, W, false, false, true>(P, blockIdx.x, smem);
}

// The FP32 sequential pass compiled for four wavefronts per SIMD (at most 128 VGPRs): more
// workgroups resident per CU; which one is faster depends on the number of active cells per
// pass (pass schedule 2, chosen by the autotuner).
template <int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) void k_pass_o4(PassArgs P) {
    extern __shared__ __align__(16) unsigned char smem[];
    pass_cell<float, W, false, false, false, true>(P, blockIdx.x, smem);
}

}  // namespace mcgb
