// mcgb_duo.hpp — the sublattice pass with two trial moves of a cell in flight at once.
//
// Same Markov chain as k_pass: the moves of an active cell are decided one after another, as in
// attempt_monte_carlo_move (MCnptColloid.c:2445-2604).  One workgroup of two wavefronts per
// active cell; each wavefront evaluates a whole move over the whole stencil by itself
// (wavefront-private candidate list, pair sums reduced within the wavefront), move m on
// wavefront 0 and move m + 1 on wavefront 1 at the same time.  When the two moves draw different
// particles a and b, the energy change of move m + 1 depends on the outcome of move m through
// one pair only, (b, a): wavefront 1 sums its pair terms against a at a's old state and against
// a at a's trial state into two separate sums, beside the sum over every other partner, and
// after the barrier that publishes move m's decision adds the one that holds.  Both moves
// commit after that barrier, and a second barrier closes the step.  Two moves per two barriers;
// moves m and m + 1 that draw the same particle run one per step.
//
// The stencil staging, the per-chunk move records, the midpoint pre-filter, the pair function
// and the commit are those of k_pass (mcgb_pass.hpp); the decisions equal the sequential
// schedule's up to the summation order.
#pragma once

#include "mcgb_spec.hpp"

namespace mcgb {

// LDS layout: stencil tables; positions (stage_cap records of 12 bytes, then one spare record
// that holds the trial state of move m for move m + 1); orientations (the same count); two
// wavefront-private candidate lists of stage_cap + 1 entries (+ a slot that unselected lanes
// write); the draws region of PassLds<Real, 2>; the two wavefronts' decisions.
template <typename Real>
struct DuoLds {
    using PL = PassLds<Real, 2>;
    static constexpr int kHeader = 1024;
    __host__ __device__ static inline int stage_cap(int cap) { return PL::stage_cap(cap); }
    __host__ __device__ static inline int spare(int cap) { return stage_cap(cap); }
    __host__ __device__ static inline int off_su(int cap) {
        return kHeader + ((12 * (stage_cap(cap) + 1) + 15) & ~15);
    }
    __host__ __device__ static inline int lcap(int cap) { return stage_cap(cap) + 1; }
    __host__ __device__ static inline int list_stride(int cap) { return lcap(cap) + 7; }
    __host__ __device__ static inline int off_list(int cap) {
        return off_su(cap) + (((stage_cap(cap) + 1) * 3 * (int)sizeof(Real) + 15) & ~15);
    }
    __host__ __device__ static inline int off_draws(int cap) {
        return (off_list(cap) + 2 * list_stride(cap) * 2 + 15) & ~15;
    }
    __host__ __device__ static inline int off_dec(int cap) {
        return off_draws(cap) + PL::kRawOff + PL::kRawBytes;
    }
    __host__ __device__ static inline size_t bytes(int cap) { return (size_t)off_dec(cap) + 32; }
};

// Pair sums of one move within its wavefront.  For the second move of a step, items against the
// first move's particle are routed: its current (old) state to `ro`, its trial state (the spare
// record) to `rt`; everything else to `m`.
template <typename Real>
struct DuoAcc {
    LaneAcc<Real> m, ro, rt;
    int cm = 0, co = 0, ct = 0;          // capped pairs (+1 trial state, -1 old state)
    int po = 0, pt = 0;                  // in-cutoff pairs against the first move's particle
    Real sm = 0, so = 0, st = 0;         // sum |terms| (trace only)
};

template <typename Real, bool TR>
__device__ __forceinline__ void pass_cell_duo(const PassArgs& P, int b, unsigned char* smem) {
    constexpr int W = 2;
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
    using L = DuoLds<Real>;
    using PL = PassLds<Real, 2>;
    const int lcap = L::lcap(cap);
    const int SP = L::spare(cap);

    int* s_start = (int*)smem;                                 // 128 (<= 125 used)
    int* s_off = s_start + 128;                                // 128 (<= 126 used)
    using R3 = Vec3<Real>;
    Pos3* sq = (Pos3*)(smem + L::kHeader);
    R3* su = (R3*)(smem + L::off_su(cap));
    unsigned short* slist = (unsigned short*)(smem + L::off_list(cap)) + w * L::list_stride(cap);
    unsigned char* dp = smem + L::off_draws(cap);
    constexpr int KD = PL::kDraws;
    MovePrep<Real>* sprep = (MovePrep<Real>*)dp;
    double* scache = (double*)(dp + PL::kPrepOff);
    unsigned char* scc = (unsigned char*)(scache + PL::kCache);
    uint2* sdraw2 = (uint2*)(dp + PL::kRawOff);
    int* sdec = (int*)(smem + L::off_dec(cap));                // [0], [1]: decisions; [2..]: sums
    WavePartial* sred = (WavePartial*)(dp + PL::kDrawBytes);

    typename GBConst<Real>::type G = gb_const(C, (Real*)nullptr);
    PairGeom X = {D.s[0], D.s[1], D.s[2], C.rc2};
    pin_vgpr(G, X);
    uint32_t* wq[3] = {P.qx, P.qy, P.qz};
    Real* wu[3] = {Ux, Uy, Uz};
    if constexpr (std::is_same<Real, float>::value)
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

    // the draws of the cell's moves, KD at a time, over wavefront 0 (as k_pass)
    const int k = D.k;
    const size_t tbase = ((size_t)P.pass * (size_t)nact + (size_t)a) * (size_t)k;
    const MoveDraw* __restrict__ drw = P.draws + C.rec_off + tbase;
    const uint2* __restrict__ drw2 = (const uint2*)drw;
    constexpr int kW2 = (int)(sizeof(MoveDraw) / 8);
    static_assert(sizeof(MoveDraw) % 8 == 0 && KD * kW2 <= 128, "a chunk's draws fit two words per lane");
    uint2 dpre = make_uint2(0u, 0u), dpre_b = make_uint2(0u, 0u);   // words lane, lane + 64
    if (w == 0 && lane < min(k, KD) * kW2) dpre = drw2[lane];
    if (w == 0 && lane + 64 < min(k, KD) * kW2) dpre_b = drw2[lane + 64];

    // stencil table (as k_pass)
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
    if (S > cap) {                             // uniform: leave this cell to the HBM kernel
        if (tid == 0) {
            atomicOr(&D.overflow, 1);
            const int slot = atomicAdd(&P.ovf_count[P.pass], 1);
            P.ovf_list[(size_t)P.pass * P.nblk + slot] = b;
        }
        return;
    }
    // staging with the sentinel padding (as k_pass: 6 records per thread per step)
    const int S_pad = (S + 63) & ~63;
    {
        constexpr int kB = 6;
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

    int n_att = 0, n_acc = 0, n_left = 0, n_hard = 0;
    int dcap_sum = 0, dccap_sum = 0;
    int my_pairs = 0;                   // per lane (main items), plus per-wave extras on lane 0
    double dfin_sum = 0.0, moved = 0.0;
    const float sx = (float)X.s0, sy = (float)X.s1, sz = (float)X.s2;
    const float rc2p = C.f_rc2_pre;
    const double creach = C.R + C.colloid_cut + 0.5;
    const double creach2 = creach * creach;
    const FilterConst FC = filter_const(sx, sy, sz, rc2p);
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
    const bool ccache = cnear && nc <= PL::kCache;
    // the cell's staged records come from both wavefronts' staging loads: a barrier before the
    // cache reads them (uniform: ccache is the workgroup's)
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

    if (nc == 0) {
        if (TR)
            for (int m = tid; m < k; m += 64 * W) {
                mcgb_move_trace rec = {-1, MCGB_MOVE_EMPTY, 0, 0, 0.0, 0.0};
                P.trace[tbase + m] = rec;
            }
    } else {
        for (int m0 = 0; m0 < k; m0 += KD) {
            __syncthreads();
            // the chunk's move records (as k_pass)
            int myj = -1;
            MovePrep<Real> mp = {};
            if (w == 0) {
                if (lane < min(k - m0, KD) * kW2) sdraw2[lane] = dpre;
                if (lane + 64 < min(k - m0, KD) * kW2) sdraw2[lane + 64] = dpre_b;
                if (lane < min(k - m0 - KD, KD) * kW2) dpre = drw2[(size_t)(m0 + KD) * kW2 + lane];
                if (lane + 64 < min(k - m0 - KD, KD) * kW2)
                    dpre_b = drw2[(size_t)(m0 + KD) * kW2 + lane + 64];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (tid < KD && m0 + tid < k) {
                const MoveDraw d = ((const MoveDraw*)sdraw2)[tid];
                const int js = (int)(((uint64_t)d.j * (uint32_t)nc) >> 32);
                myj = js;
                mp.thr = d.thr;
                mp.thr1 = d.thr1;
                mp.dqx = d.dq[0]; mp.dqy = d.dq[1]; mp.dqz = d.dq[2];
                mp.jsel = js;
                mp.vx = (Real)d.v[0]; mp.vy = (Real)d.v[1]; mp.vz = (Real)d.v[2];
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
                mp.flags |= prep_tests(qj.x, qj.y, qj.z, d.dq[0], d.dq[1], d.dq[2], gx, gy, gz,
                                       nx, ny, nz, cx, cy, cz, cnear, R2, X, FC, &mp.fr2);
            }
            if (w == 0) {
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
            for (int m = m0; m < mend;) {     // uniform: every value below comes from LDS
                const int ma = m;
                int mb = -1;
                if (ma + 1 < mend) {
                    const uint32_t flb = uni(sprep[ma + 1 - m0].flags);
                    if (!((flb >> (ma - m0)) & 1u)) mb = ma + 1;
                }
                m = mb >= 0 ? ma + 2 : ma + 1;
                const int mine = w == 0 ? ma : mb;           // uniform per wavefront
                const bool second = w == 1;                   // this wavefront speculates on ma
                int outcome = MCGB_MOVE_REJECT;
                bool evaluated = false;                       // energy computed, decision pending
                int capd = 0, ccd = 0;
                double dfin = 0.0, scale = 0.0;
                double cacc_m = 0.0, cscl_m = 0.0, cen_m = 0.0;
                bool ccn_m = false;
                int li = 0, gsel = 0, jsel = 0;
                uint32_t qnx = 0, qny = 0, qnz = 0;
                Real unx = 0, uny = 0, unz = 0;
                double thr = 0.0;
                // second wavefront: the pair sums against ma's particle, old and trial
                double d_main = 0.0, d_old = 0.0, d_trial = 0.0;
                int c_main = 0, c_old = 0, c_trial = 0, p_old = 0, p_trial = 0;
                double s_main = 0.0, s_old = 0.0, s_trial = 0.0;
                if (mine >= 0) {
                    const int ml = mine - m0;
                    const MovePrep<Real> pm = sprep[ml];
                    jsel = uni(pm.jsel);
                    const uint32_t fl = uni(pm.flags);
                    const bool fresh = (fl & acc_mask & 0xffffu) == 0u;
                    li = coff + jsel;
                    gsel = cstart + jsel;
                    uint32_t qix, qiy, qiz;
                    Real uix, uiy, uiz;
                    if (fresh) {
                        qix = uni(pm.qx); qiy = uni(pm.qy); qiz = uni(pm.qz);
                        uix = uni(pm.ux); uiy = uni(pm.uy); uiz = uni(pm.uz);
                    } else {
                        const Pos3 qi = sq[li];
                        const R3 ui = su[li];
                        qix = uni(qi.x); qiy = uni(qi.y); qiz = uni(qi.z);
                        uix = uni(ui.x); uiy = uni(ui.y); uiz = uni(ui.z);
                    }
                    const int dqx = uni(pm.dqx), dqy = uni(pm.dqy), dqz = uni(pm.dqz);
                    qnx = qix + (uint32_t)dqx; qny = qiy + (uint32_t)dqy; qnz = qiz + (uint32_t)dqz;
                    thr = pm.thr;
                    ++n_att;
                    float fr2;
                    const uint32_t tests =
                        fresh ? fl & (kPrepLeft | kPrepHard)
                              : prep_tests(qix, qiy, qiz, dqx, dqy, dqz, gx, gy, gz, nx, ny, nz,
                                           cx, cy, cz, cnear, R2, X, FC, &fr2);
                    if (fresh) fr2 = uni(pm.fr2);
                    if (tests & kPrepLeft) {
                        outcome = MCGB_MOVE_LEFT;
                        ++n_left;
                    } else if (tests & kPrepHard) {
                        outcome = MCGB_MOVE_HARD;
                        ++n_hard;
                    } else {
                        evaluated = true;
                        const double rnx = (double)(int32_t)(qnx - 0x80000000u) * X.s0;
                        const double rny = (double)(int32_t)(qny - 0x80000000u) * X.s1;
                        const double rnz = (double)(int32_t)(qnz - 0x80000000u) * X.s2;
                        if (fresh) {
                            unx = uni(pm.nx); uny = uni(pm.ny); unz = uni(pm.nz);
                        } else {
                            rotate(drot, uix, uiy, uiz, pm.vx, pm.vy, pm.vz, unx, uny, unz);
                            unx = uni(unx); uny = uni(uny); unz = uni(unz);
                        }
                        // the first move's particle as the second wavefront sees it: slot, and
                        // its trial state in the spare record (the first wavefront derives the
                        // same bits from the same record)
                        int lia = -1;
                        if (second) {
                            const MovePrep<Real> pa = sprep[ma - m0];
                            const int ja = uni(pa.jsel);
                            const bool fa = (uni(pa.flags) & acc_mask & 0xffffu) == 0u;
                            lia = coff + ja;
                            uint32_t ax0, ay0, az0;
                            Real aux, auy, auz, anx, a_ny, anz;
                            if (fa) {
                                ax0 = uni(pa.qx); ay0 = uni(pa.qy); az0 = uni(pa.qz);
                                anx = uni(pa.nx); a_ny = uni(pa.ny); anz = uni(pa.nz);
                            } else {
                                const Pos3 qa = sq[lia];
                                const R3 ua = su[lia];
                                ax0 = uni(qa.x); ay0 = uni(qa.y); az0 = uni(qa.z);
                                aux = uni(ua.x); auy = uni(ua.y); auz = uni(ua.z);
                                rotate(drot, aux, auy, auz, pa.vx, pa.vy, pa.vz, anx, a_ny, anz);
                                anx = uni(anx); a_ny = uni(a_ny); anz = uni(anz);
                            }
                            if (lane == 0) {
                                sq[SP] = Pos3{ax0 + (uint32_t)uni(pa.dqx), ay0 + (uint32_t)uni(pa.dqy),
                                              az0 + (uint32_t)uni(pa.dqz)};
                                su[SP] = R3{anx, a_ny, anz};
                            }
                        }
                        // candidates of both states over the whole stencil (wavefront-private)
                        const MidFilter M = mid_filter_r(qix, qiy, qiz, dqx, dqy, dqz, FC, fr2);
                        int base = 0;
                        auto scan = [&](auto iso) {
                            constexpr int kG = 4;
                            const int ng = S_pad >> 6;
                            for (int g0 = 0; g0 < ng; g0 += kG) {
                                const int nv = min(kG, ng - g0);   // uniform
                                Pos3 p[kG];
#pragma unroll
                                for (int j = 0; j < kG; ++j)
                                    if (j < nv) p[j] = sq[(g0 + j) * 64 + lane];
#pragma unroll
                                for (int j = 0; j < kG; ++j)
                                    if (j < nv) {
                                        const int e = (g0 + j) * 64 + lane;
                                        const bool kk = M.near(p[j], iso) & (e != li);
                                        const unsigned long long mk = __builtin_amdgcn_ballot_w64(kk);
                                        const int rank = lane_rank(mk);
                                        slist[kk ? base + rank : lcap] = (unsigned short)e;
                                        base += __popcll(mk);
                                    }
                            }
                        };
                        if (M.iso) scan(std::true_type{});
                        else scan(std::false_type{});
                        if (second) {                 // the first move's trial state, unfiltered
                            if (lane == 0) slist[base] = (unsigned short)SP;
                            base += 1;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        // colloid-rod term of both states in FP64 (as k_pass)
                        if (ccache) {
                            const double eo = scache[jsel];
                            const bool co = scc[jsel] != 0;
                            if (fresh) {
                                cen_m = pm.en;
                                ccn_m = (fl >> 16) & 1u;
                            } else {
                                cen_m = colloid_f64(C, rnx, rny, rnz, (double)unx, (double)uny,
                                                    (double)unz, &ccn_m);
                            }
                            ccd = (ccn_m ? 1 : 0) - (co ? 1 : 0);
                            cacc_m = (ccn_m ? 0.0 : cen_m) - (co ? 0.0 : eo);
                            cscl_m = fabs(ccn_m ? 0.0 : cen_m) + fabs(co ? 0.0 : eo);
                        } else if (cnear) {
                            const double rox = (double)(int32_t)(qix - 0x80000000u) * X.s0;
                            const double roy = (double)(int32_t)(qiy - 0x80000000u) * X.s1;
                            const double roz = (double)(int32_t)(qiz - 0x80000000u) * X.s2;
                            const bool nearc = rox * rox + roy * roy + roz * roz <= creach2 ||
                                               rnx * rnx + rny * rny + rnz * rnz <= creach2;
                            if (nearc) {
                                bool co, cn;
                                const double eo = colloid_f64(C, rox, roy, roz, (double)uix,
                                                              (double)uiy, (double)uiz, &co);
                                const double en = colloid_f64(C, rnx, rny, rnz, (double)unx,
                                                              (double)uny, (double)unz, &cn);
                                ccd = (cn ? 1 : 0) - (co ? 1 : 0);
                                cacc_m = (cn ? 0.0 : en) - (co ? 0.0 : eo);
                                cscl_m = fabs(cn ? 0.0 : en) + fabs(co ? 0.0 : eo);
                            }
                        }
                        // pair terms, items it and it + 64 per lane and step (as k_pass); the
                        // second wavefront routes the items against the first move's particle
                        DuoAcc<Real> A;
                        const int items = 2 * base;
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
                            const bool ia = in2[0], ib = in2[1] & hb;
                            const bool fa = ia & !c2[0], fb = ib & !c2[1];
                            const Real ta = fa ? (nw ? v2[0] : -v2[0]) : (Real)0;
                            const Real tb = fb ? (nw ? v2[1] : -v2[1]) : (Real)0;
                            const int ka = (int)(ia & c2[0]) * (nw ? 1 : -1);
                            const int kb = (int)(ib & c2[1]) * (nw ? 1 : -1);
                            if (second) {
                                const bool oa = ea == lia, ob = eb == lia;
                                const bool sa = ea == SP, sb = eb == SP;
                                const bool xa = oa | sa, xb = ob | sb;
                                A.m.add(xa ? (Real)0 : ta);
                                A.m.add(xb ? (Real)0 : tb);
                                A.ro.add(oa ? ta : (Real)0);
                                A.ro.add(ob ? tb : (Real)0);
                                A.rt.add(sa ? ta : (Real)0);
                                A.rt.add(sb ? tb : (Real)0);
                                A.cm += (xa ? 0 : ka) + (xb ? 0 : kb);
                                A.co += (oa ? ka : 0) + (ob ? kb : 0);
                                A.ct += (sa ? ka : 0) + (sb ? kb : 0);
                                my_pairs += (int)(ia & !xa) + (int)(ib & !xb);
                                A.po += (int)(ia & oa) + (int)(ib & ob);
                                A.pt += (int)(ia & sa) + (int)(ib & sb);
                                if (TR) {
                                    const Real aa = ta < 0 ? -ta : ta, ab = tb < 0 ? -tb : tb;
                                    A.sm += (xa ? (Real)0 : aa) + (xb ? (Real)0 : ab);
                                    A.so += (oa ? aa : (Real)0) + (ob ? ab : (Real)0);
                                    A.st += (sa ? aa : (Real)0) + (sb ? ab : (Real)0);
                                }
                            } else {
                                A.m.add(ta);
                                A.m.add(tb);
                                A.cm += ka + kb;
                                my_pairs += (int)ia + (int)ib;
                                if (TR) A.sm += (ta < 0 ? -ta : ta) + (tb < 0 ? -tb : tb);
                            }
                        }
                        // reductions within the wavefront (results in lane 63, broadcast)
                        WavePartial part;
                        A.m.to_partial(part);
                        d_main = from63(part.d);
                        c_main = __builtin_amdgcn_ballot_w64(A.cm != 0) ? from63(wave_sum_i63(A.cm)) : 0;
                        if (TR) s_main = from63(wave_sum_d63((double)A.sm));
                        if (second) {
                            A.ro.to_partial(part);
                            d_old = from63(part.d);
                            A.rt.to_partial(part);
                            d_trial = from63(part.d);
                            if (__builtin_amdgcn_ballot_w64((A.co | A.ct) != 0)) {
                                c_old = from63(wave_sum_i63(A.co));
                                c_trial = from63(wave_sum_i63(A.ct));
                            }
                            p_old = from63(wave_sum_i63(A.po));
                            p_trial = from63(wave_sum_i63(A.pt));
                            if (TR) {
                                s_old = from63(wave_sum_d63((double)A.so));
                                s_trial = from63(wave_sum_d63((double)A.st));
                            }
                        }
                    }
                }
                // the first move decides at once
                bool accept = false;
                if (!second && evaluated) {
                    capd = c_main;
                    dfin = d_main + cacc_m;
                    scale = s_main + cscl_m;
                    // log U <= -beta dE (MCnptColloid.c:2574)
                    accept = (double)capd * 1.0e9 + (double)ccd * 1.0e8 + dfin <= thr;
                    accept = uni((int)accept) != 0;
                }
                if (!second && lane == 0) sdec[0] = accept ? 1 : 0;
                __syncthreads();                      // the first move's decision is published
                const int acc_a = uni(sdec[0]);
                if (second && evaluated) {
                    // the pair with the first move's particle at the state that holds
                    capd = c_main + (acc_a ? c_trial : c_old);
                    dfin = d_main + (acc_a ? d_trial : d_old) + cacc_m;
                    scale = s_main + (acc_a ? s_trial : s_old) + cscl_m;
                    if (lane == 0) my_pairs += acc_a ? p_trial : p_old;
                    accept = (double)capd * 1.0e9 + (double)ccd * 1.0e8 + dfin <= thr;
                    accept = uni((int)accept) != 0;
                }
                if (mine >= 0) {
                    if (accept) {
                        outcome = MCGB_MOVE_ACCEPT;
                        ++n_acc;
                        dcap_sum += capd;
                        dccap_sum += ccd;
                        dfin_sum += dfin;
                        moved += fabs(dfin);
                        if (lane == 0) {
                            sq[li] = Pos3{qnx, qny, qnz};
                            su[li] = R3{unx, uny, unz};
                            if (ccache) {
                                scache[jsel] = cen_m;
                                scc[jsel] = ccn_m ? 1 : 0;
                            }
                            wq[0][gsel] = qnx; wq[1][gsel] = qny; wq[2][gsel] = qnz;
                            wu[0][gsel] = unx; wu[1][gsel] = uny; wu[2][gsel] = unz;
                        }
                    }
                    if (TR && lane == 0) {
                        mcgb_move_trace rec;
                        rec.pid = P.pid[gsel];
                        rec.outcome = outcome;
                        rec.dcap = evaluated ? capd : 0;
                        rec.dccap = evaluated ? ccd : 0;
                        rec.dfin = evaluated ? dfin : 0.0;
                        rec.scale = evaluated ? scale : 0.0;
                        P.trace[tbase + mine] = rec;
                    }
                }
                if (second && lane == 0) sdec[1] = accept ? 1 : 0;
                __syncthreads();                      // both moves committed
                if (second) {
                    acc_mask |= (acc_a ? 1u : 0u) << (ma - m0);
                    if (accept) acc_mask |= 1u << (mb - m0);
                } else {
                    if (accept) acc_mask |= 1u << (ma - m0);
                    if (mb >= 0 && uni(sdec[1])) acc_mask |= 1u << (mb - m0);
                }
                acc_mask = uni(acc_mask);
            }
        }
    }
    // per-cell accumulators: each wavefront counted its own moves
    const int wp = wave_sum_i63(my_pairs);
    __syncthreads();
    if (lane == 63) {
        sred[w].capd = wp;
        sred[w].d = dfin_sum;
        sred[w].scale = moved;
    }
    if (w == 1 && lane == 0) {
        int* x = sdec + 2;
        x[0] = n_att; x[1] = n_acc; x[2] = n_left; x[3] = n_hard;
        x[4] = dcap_sum; x[5] = dccap_sum;
    }
    __syncthreads();
    if (tid == 0) {
        const int* x = sdec + 2;
        CellAcc ca;
        ca.att = n_att + x[0]; ca.acc = n_acc + x[1]; ca.left = n_left + x[2];
        ca.hard = n_hard + x[3];
        ca.dcap = dcap_sum + x[4]; ca.dccap = dccap_sum + x[5];
        ca.pairs = sred[0].capd + sred[1].capd;
        // both counted per evaluated move (neither left the cell nor hit the colloid core)
        const int nev = n_att + x[0] - n_left - x[2] - n_hard - x[3];
        ca.coll = colloid_on ? 2 * nev : 0;
        ca.cand = 2LL * (S - 1) * (long long)nev;
        ca.dfin = sred[0].d + sred[1].d;
        ca.moved_abs = sred[0].scale + sred[1].scale;
        ca.s1rej = 0; ca.s2rej = 0;
        ca.ticks = 0; ca.tmoves = 0;
        P.acc[C.cell_off + cell] = ca;
    }
}

// Two wavefronts per cell, at most three wavefronts per SIMD (as k_pass).
template <typename Real, bool TR>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 3))) void k_pass_duo(PassArgs P) {
    extern __shared__ __align__(16) unsigned char smem[];
    pass_cell_duo<Real, TR>(P, blockIdx.x, smem);
}

}  // namespace mcgb
