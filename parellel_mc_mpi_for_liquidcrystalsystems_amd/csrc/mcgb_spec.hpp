// mcgb_spec.hpp — the sublattice pass with speculative evaluation of a cell's moves.
//
// Same Markov chain as k_pass (the moves of an active cell are decided one after another, as in
// attempt_monte_carlo_move, MCnptColloid.c:2445-2604), different schedule.  During a pass only
// the active cell's own particles can change, so the energy change of a move splits into
//   env  = sum over stencil particles outside the active cell (fixed for the whole pass) and
//   cell = sum over the other particles of the active cell (current states).
// Phase A: the W wavefronts of the workgroup evaluate env and the colloid term of up to
// kChunk moves in parallel (one move per wavefront at a time, no workgroup barrier), from the
// state at the start of the chunk.  Phase B: wavefront 0 takes the moves in order; a move whose
// particle was not moved earlier in the chunk adds the cell part at the current states and
// decides; a move whose particle was moved earlier (its phase-A numbers are stale) is evaluated
// again in full.  Decisions equal those of the sequential schedule up to the summation order.
#pragma once

#include "mcgb_pass.hpp"

namespace mcgb {

// Phase-A result of one move.
struct SpecRes {
    double env;       // finite env dE (FP64 lane sums)
    double cacc;      // colloid finite dE
    int capd;         // change of capped env pairs
    int ccd;          // change of capped colloid terms
    int outcome;      // 0 evaluated, MCGB_MOVE_LEFT, MCGB_MOVE_HARD
    int pad;
};

template <typename Real, int W>
struct SpecLds {
    static constexpr int kChunk = 32;   // moves per phase-A / phase-B round
    static constexpr int kRing = 128;   // pending candidates per wavefront (>= 63 + 64)
    __host__ __device__ static inline size_t bytes(int cap) {
        size_t b = 256;                                          // stencil tables
        b += (size_t)cap * (12 + 3 * sizeof(Real));              // staged SoA
        b += (size_t)W * kRing * 2;                              // candidate rings
        b = (b + 15) & ~(size_t)15;
        b += kChunk * (8 + 3 * sizeof(Real) + 12 + 4);           // draws
        b = (b + 15) & ~(size_t)15;
        b += kChunk * sizeof(SpecRes);                           // phase-A results
        b += kChunk * 2 + 16;                                    // moved list
        return b + 32;
    }
};

// Accumulates the pair terms of the moved particle (old state, trial state) against listed
// stencil slots: one (slot, state) item per lane.
template <typename Real>
struct SpecAcc {
    LaneAcc<Real> acc;
    int cpl = 0;
    int pairs = 0;
};

template <typename Real, typename G>
__device__ inline void spec_item(const G& Gc, const PairGeom& X, const uint32_t* sqx,
                                 const uint32_t* sqy, const uint32_t* sqz, const Real* sux,
                                 const Real* suy, const Real* suz, int e, bool nw, uint32_t qix,
                                 uint32_t qiy, uint32_t qiz, uint32_t qnx, uint32_t qny,
                                 uint32_t qnz, Real uix, Real uiy, Real uiz, Real unx, Real uny,
                                 Real unz, SpecAcc<Real>& A) {
    const uint32_t px = nw ? qnx : qix, py = nw ? qny : qiy, pz = nw ? qnz : qiz;
    bool cp;
    Real u;
    if (pair_energy(Gc, X, nw ? unx : uix, nw ? uny : uiy, nw ? unz : uiz, sux[e], suy[e], suz[e],
                    (int)(sqx[e] - px), (int)(sqy[e] - py), (int)(sqz[e] - pz), &u, &cp)) {
        ++A.pairs;
        if (cp) A.cpl += nw ? 1 : -1;
        else A.acc.add(nw ? u : -u);
    }
}

// Values held by lane 63 (the DPP reductions' result lane), broadcast to the wavefront.
__device__ inline int from63(int v) { return __builtin_amdgcn_readlane(v, 63); }
__device__ inline double from63(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xFFFFFFFFull), 63);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), 63);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ inline void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wavefront: the pair terms of a move against every staged slot in [0, S) except `li` and,
// when skip_cell, the active cell's slots [c0, c1).  FP32 pre-filter of both states, candidates
// through a wavefront-private ring, evaluated 32 slots (64 items) at a time.
template <typename Real, typename G>
__device__ inline void spec_scan(const G& Gc, const PairGeom& X, float sx, float sy, float sz,
                                 float rc2p, const uint32_t* sqx, const uint32_t* sqy,
                                 const uint32_t* sqz, const Real* sux, const Real* suy,
                                 const Real* suz, unsigned short* ring, int S, int li, bool skip_cell,
                                 int c0, int c1, uint32_t qix, uint32_t qiy, uint32_t qiz,
                                 int dqx, int dqy, int dqz, Real uix, Real uiy, Real uiz, Real unx,
                                 Real uny, Real unz, int lane, SpecAcc<Real>& A) {
    constexpr int R = SpecLds<Real, 1>::kRing;
    const uint32_t qnx = qix + (uint32_t)dqx, qny = qiy + (uint32_t)dqy, qnz = qiz + (uint32_t)dqz;
    const MidFilter M = mid_filter(qix, qiy, qiz, dqx, dqy, dqz, filter_const(sx, sy, sz, rc2p));
    int head = 0, tail = 0;
    auto flush = [&](int n) {          // n <= 32 slots, wavefront-uniform
        wave_sync_lds();
        if (lane < 2 * n) {
            const int e = ring[(head + (lane >> 1)) & (R - 1)];
            spec_item<Real>(Gc, X, sqx, sqy, sqz, sux, suy, suz, e, lane & 1, qix, qiy, qiz, qnx,
                            qny, qnz, uix, uiy, uiz, unx, uny, unz, A);
        }
        head += n;
        wave_sync_lds();
    };
    auto test = [&](int e) -> bool {      // branch-free: clamped in-bounds reads
        const int ce = min(e, S - 1);
        const bool valid = (e < S) & (e != li) & !(skip_cell & (e >= c0) & (e < c1));
        return M.near(sqx[ce], sqy[ce], sqz[ce]) & valid;
    };
    auto push = [&](int e, bool keep) {
        const unsigned long long mask = __ballot(keep);
        if (keep) ring[(tail + __popcll(mask & __lanemask_lt())) & (R - 1)] = (unsigned short)e;
        tail += __popcll(mask);
    };
    for (int e0 = 0; e0 < S; e0 += 128) {
        const int ea = e0 + lane, eb = e0 + 64 + lane;
        const bool ka = test(ea);
        const bool kb = test(eb);
        push(ea, ka);                  // pending <= 31 before a push, <= 95 after
        while (tail - head >= 32) flush(32);
        push(eb, kb);
        while (tail - head >= 32) flush(32);
    }
    while (tail - head >= 32) flush(32);
    if (tail > head) flush(tail - head);
}

template <typename Real, int W>
__global__ __launch_bounds__(64 * W) void k_pass_spec(PassArgs P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    const int r = P.blk_rep[b];
    const RepConst& C = P.RC[r];
    RepDyn& D = P.RD[r];
    const int a = b - C.act_off;
    const int nact = D.nact;
    if (a >= nact) return;                    // uniform: whole block
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int cap = P.cap;
    Real* __restrict__ Ux = (Real*)P.ux;
    Real* __restrict__ Uy = (Real*)P.uy;
    Real* __restrict__ Uz = (Real*)P.uz;
    using L = SpecLds<Real, W>;
    constexpr int KD = L::kChunk;

    int* s_start = (int*)smem;                                 // 32
    int* s_off = s_start + 32;                                 // 32 (28 used)
    uint32_t* sqx = (uint32_t*)(smem + 256);
    uint32_t* sqy = sqx + cap;
    uint32_t* sqz = sqy + cap;
    Real* sux = (Real*)(sqz + cap);
    Real* suy = sux + cap;
    Real* suz = suy + cap;
    unsigned short* rings = (unsigned short*)(suz + cap);
    unsigned short* ring = rings + (size_t)w * L::kRing;
    unsigned char* dp = (unsigned char*)(rings + (size_t)W * L::kRing);
    dp = (unsigned char*)(((uintptr_t)dp + 15) & ~(uintptr_t)15);
    double* sthr = (double*)dp;                                // KD
    Real* svx = (Real*)(sthr + KD);                            // 3 x KD
    Real* svy = svx + KD;
    Real* svz = svy + KD;
    int* sdqx = (int*)(svz + KD);                              // 3 x KD
    int* sdqy = sdqx + KD;
    int* sdqz = sdqy + KD;
    uint32_t* sj = (uint32_t*)(sdqz + KD);                     // KD
    SpecRes* sres = (SpecRes*)(((uintptr_t)(sj + KD) + 15) & ~(uintptr_t)15);
    unsigned short* smv = (unsigned short*)(sres + KD);        // KD moved cell members
    int* sred = (int*)(smv + KD + 8);                          // W pair counts

    typename GBConst<Real>::type G = gb_const(C, (Real*)nullptr);
    PairGeom X = {D.s[0], D.s[1], D.s[2], C.rc2};
    // FP32: the pair constants in vector registers (as in the sequential pass); with more
    // than two wavefronts per cell the extra registers cost a workgroup per CU (measured)
    if constexpr (W <= 2) pin_vgpr(G, X);
    const int part_off = C.part_off;
    const int colloid_on = C.colloid_on;
    const double R2 = C.R2;
    const double drot = (double)D.drot;
    const int nx = D.n[0], ny = D.n[1], nz = D.n[2];
    const int hx = nx >> 1, hy = ny >> 1;
    const uint32_t sig = D.order[P.pass];
    const int ax = a % hx, ay = (a / hx) % hy, az = a / (hx * hy);
    const int cx = 2 * ax + (int)(sig & 1u), cy = 2 * ay + (int)((sig >> 1) & 1u),
              cz = 2 * az + (int)((sig >> 2) & 1u);
    const int cell = cx + nx * (cy + ny * cz);
    const int* cs = P.cell_start + C.cs_off;
    const uint32_t gx = D.g[0], gy = D.g[1], gz = D.g[2];

    // stencil table: 27 neighbour cells in (dz, dy, dx) order, centre = 13
    if (w == 0) {
        int cnt = 0;
        if (lane < 27) {
            int ex = cx + (lane % 3) - 1, ey = cy + ((lane / 3) % 3) - 1, ez = cz + lane / 9 - 1;
            ex += ex < 0 ? nx : 0; ex -= ex >= nx ? nx : 0;
            ey += ey < 0 ? ny : 0; ey -= ey >= ny ? ny : 0;
            ez += ez < 0 ? nz : 0; ez -= ez >= nz ? nz : 0;
            int nb = ex + nx * (ey + ny * ez);
            int st = cs[nb];
            cnt = cs[nb + 1] - st;
            s_start[lane] = st;
        }
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) {
            int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (lane < 27) s_off[lane + 1] = incl;
        if (lane == 0) s_off[0] = 0;
    }
    __syncthreads();
    const int S = s_off[27];
    const int coff = s_off[13];
    const int nc = s_off[14] - coff;
    const int cstart = part_off + s_start[13];
    if (S > cap) {                            // uniform: leave this cell to the HBM kernel
        if (tid == 0) {
            atomicOr(&D.overflow, 1);
            const int slot = atomicAdd(&P.ovf_count[P.pass], 1);
            P.ovf_list[(size_t)P.pass * P.nblk + slot] = b;
        }
        return;
    }
    for (int e = tid; e < S; e += 64 * W) {
        int gi = stencil_slot(s_off, s_start, part_off, e);
        sqx[e] = P.qx[gi]; sqy[e] = P.qy[gi]; sqz[e] = P.qz[gi];
        sux[e] = Ux[gi]; suy[e] = Uy[gi]; suz[e] = Uz[gi];
    }

    const int k = D.k;
    const size_t tbase = ((size_t)P.pass * (size_t)nact + (size_t)a) * (size_t)k;
    const MoveDraw* __restrict__ drw = P.draws + C.rec_off + tbase;
    int n_att = 0, n_acc = 0, n_left = 0, n_hard = 0, n_coll = 0;
    int dcap_sum = 0, dccap_sum = 0;
    int my_pairs = 0;
    long long n_cand = 0;
    double dfin_sum = 0.0, moved = 0.0;
    const float sx = (float)X.s0, sy = (float)X.s1, sz = (float)X.s2;
    const float rc2p = C.f_rc2_pre;     // pre-filter; the cutoff itself is decided in FP64
    const double creach = C.R + C.colloid_cut + 0.5;
    const double creach2 = creach * creach;

    // colloid-rod terms of both states (FP64), only near the colloid
    auto colloid = [&](uint32_t qix, uint32_t qiy, uint32_t qiz, double rnx, double rny, double rnz,
                       Real uix, Real uiy, Real uiz, Real unx, Real uny, Real unz, double& cacc,
                       int& ccd) {
        cacc = 0.0;
        ccd = 0;
        if (!colloid_on) return;
        const double rox = (double)(int32_t)(qix - 0x80000000u) * X.s0;
        const double roy = (double)(int32_t)(qiy - 0x80000000u) * X.s1;
        const double roz = (double)(int32_t)(qiz - 0x80000000u) * X.s2;
        if (rox * rox + roy * roy + roz * roz <= creach2 || rnx * rnx + rny * rny + rnz * rnz <= creach2) {
            bool co, cn;
            const double eo = colloid_f64(C, rox, roy, roz, (double)uix, (double)uiy, (double)uiz, &co);
            const double en = colloid_f64(C, rnx, rny, rnz, (double)unx, (double)uny, (double)unz, &cn);
            ccd = (cn ? 1 : 0) - (co ? 1 : 0);
            cacc = (cn ? 0.0 : en) - (co ? 0.0 : eo);
        }
    };
    // early outcome of a move from state qi: 0, MCGB_MOVE_LEFT or MCGB_MOVE_HARD; trial position
    auto early = [&](uint32_t qnx, uint32_t qny, uint32_t qnz, double& rnx, double& rny, double& rnz) {
        rnx = (double)(int32_t)(qnx - 0x80000000u) * X.s0;
        rny = (double)(int32_t)(qny - 0x80000000u) * X.s1;
        rnz = (double)(int32_t)(qnz - 0x80000000u) * X.s2;
        if (cell_coord(qnx, gx, (uint32_t)nx) != (uint32_t)cx ||
            cell_coord(qny, gy, (uint32_t)ny) != (uint32_t)cy ||
            cell_coord(qnz, gz, (uint32_t)nz) != (uint32_t)cz)
            return (int)MCGB_MOVE_LEFT;                 // moves are confined to the active cell
        if (colloid_on && rnx * rnx + rny * rny + rnz * rnz <= R2)
            return (int)MCGB_MOVE_HARD;                 // MCnptColloid.c:2516-2520
        return 0;
    };

    if (nc > 0) {
        for (int m0 = 0; m0 < k; m0 += KD) {
            __syncthreads();
            if (tid < KD && m0 + tid < k) {   // the draws of up to KD moves (k_count_draws)
                const MoveDraw& d = drw[m0 + tid];
                sj[tid] = d.j;
                sdqx[tid] = d.dq[0];
                sdqy[tid] = d.dq[1];
                sdqz[tid] = d.dq[2];
                svx[tid] = (Real)d.v[0];
                svy[tid] = (Real)d.v[1];
                svz[tid] = (Real)d.v[2];
                sthr[tid] = d.thr;
            }
            __syncthreads();
            const int mn = min(k - m0, KD);
            // ---------------- phase A: env and colloid terms from the chunk's start state
            for (int ml = w; ml < mn; ml += W) {
                const int jsel = uni((int)(((uint64_t)sj[ml] * (uint32_t)nc) >> 32));
                const int li = coff + jsel;
                const uint32_t qix = uni(sqx[li]), qiy = uni(sqy[li]), qiz = uni(sqz[li]);
                const Real uix = uni(sux[li]), uiy = uni(suy[li]), uiz = uni(suz[li]);
                const int dqx = uni(sdqx[ml]), dqy = uni(sdqy[ml]), dqz = uni(sdqz[ml]);
                double rnx, rny, rnz;
                const int out = early(qix + (uint32_t)dqx, qiy + (uint32_t)dqy, qiz + (uint32_t)dqz,
                                      rnx, rny, rnz);
                SpecRes res = {0.0, 0.0, 0, 0, out, 0};
                if (out == 0) {
                    Real unx, uny, unz;
                    rotate(drot, uix, uiy, uiz, svx[ml], svy[ml], svz[ml], unx, uny, unz);
                    unx = uni(unx); uny = uni(uny); unz = uni(unz);
                    SpecAcc<Real> A;
                    spec_scan<Real>(G, X, sx, sy, sz, rc2p, sqx, sqy, sqz, sux, suy, suz, ring, S, li,
                                    true, coff, coff + nc, qix, qiy, qiz, dqx, dqy, dqz, uix, uiy,
                                    uiz, unx, uny, unz, lane, A);
                    my_pairs += A.pairs;
                    WavePartial part;
                    A.acc.to_partial(part);
                    res.env = LaneAcc<Real>::total(&part, 1);
                    res.capd = wave_sum_i63(A.cpl);
                    colloid(qix, qiy, qiz, rnx, rny, rnz, uix, uiy, uiz, unx, uny, unz, res.cacc, res.ccd);
                }
                if (lane == 63) sres[ml] = res;
            }
            __syncthreads();
            // ---------------- phase B: wavefront 0 decides the chunk's moves in order
            if (w == 0) {
                int nmv = 0;
                for (int ml = 0; ml < mn; ++ml) {
                    const int jsel = uni((int)(((uint64_t)sj[ml] * (uint32_t)nc) >> 32));
                    const int li = coff + jsel;
                    const int gsel = cstart + jsel;
                    const bool dirty = __ballot(lane < nmv && (int)smv[lane] == jsel) != 0ull;
                    const uint32_t qix = uni(sqx[li]), qiy = uni(sqy[li]), qiz = uni(sqz[li]);
                    const Real uix = uni(sux[li]), uiy = uni(suy[li]), uiz = uni(suz[li]);
                    const int dqx = uni(sdqx[ml]), dqy = uni(sdqy[ml]), dqz = uni(sdqz[ml]);
                    const uint32_t qnx = qix + (uint32_t)dqx, qny = qiy + (uint32_t)dqy,
                                   qnz = qiz + (uint32_t)dqz;
                    ++n_att;
                    double rnx, rny, rnz;
                    int out;
                    if (dirty) {
                        out = early(qnx, qny, qnz, rnx, rny, rnz);
                    } else {
                        out = sres[ml].outcome;
                    }
                    if (out == MCGB_MOVE_LEFT) { ++n_left; continue; }
                    if (out == MCGB_MOVE_HARD) { ++n_hard; continue; }
                    Real unx, uny, unz;
                    rotate(drot, uix, uiy, uiz, svx[ml], svy[ml], svz[ml], unx, uny, unz);
                    unx = uni(unx); uny = uni(uny); unz = uni(unz);
                    double dpair, cacc;
                    int capd, ccd;
                    SpecAcc<Real> A;
                    if (dirty) {
                        // full evaluation at the current states
                        spec_scan<Real>(G, X, sx, sy, sz, rc2p, sqx, sqy, sqz, sux, suy, suz, ring,
                                        S, li, false, 0, 0, qix, qiy, qiz, dqx, dqy, dqz, uix, uiy,
                                        uiz, unx, uny, unz, lane, A);
                        early(qnx, qny, qnz, rnx, rny, rnz);
                        colloid(qix, qiy, qiz, rnx, rny, rnz, uix, uiy, uiz, unx, uny, unz, cacc, ccd);
                        WavePartial part;
                        A.acc.to_partial(part);
                        dpair = from63(LaneAcc<Real>::total(&part, 1));
                        capd = from63(wave_sum_i63(A.cpl));
                    } else {
                        // the other particles of the active cell at their current states
                        for (int it = lane; it < 2 * nc; it += 64) {
                            const int cm = it >> 1;
                            if (cm == jsel) continue;
                            spec_item<Real>(G, X, sqx, sqy, sqz, sux, suy, suz, coff + cm, it & 1,
                                            qix, qiy, qiz, qnx, qny, qnz, uix, uiy, uiz, unx, uny,
                                            unz, A);
                        }
                        WavePartial part;
                        A.acc.to_partial(part);
                        const SpecRes& R0 = sres[ml];
                        dpair = R0.env + from63(LaneAcc<Real>::total(&part, 1));
                        capd = R0.capd + from63(wave_sum_i63(A.cpl));
                        cacc = R0.cacc;
                        ccd = R0.ccd;
                    }
                    my_pairs += A.pairs;
                    if (colloid_on) n_coll += 2;
                    n_cand += 2LL * (S - 1);
                    const double dfin = dpair + cacc;
                    // log U <= -beta dE (MCnptColloid.c:2574)
                    if ((double)capd * 1.0e9 + (double)ccd * 1.0e8 + dfin <= sthr[ml]) {
                        ++n_acc;
                        dcap_sum += capd;
                        dccap_sum += ccd;
                        dfin_sum += dfin;
                        moved += fabs(dfin);
                        if (lane == 0) {
                            sqx[li] = qnx; sqy[li] = qny; sqz[li] = qnz;
                            sux[li] = unx; suy[li] = uny; suz[li] = unz;
                            smv[nmv] = (unsigned short)jsel;
                            P.qx[gsel] = qnx; P.qy[gsel] = qny; P.qz[gsel] = qnz;
                            Ux[gsel] = unx; Uy[gsel] = uny; Uz[gsel] = unz;
                        }
                        ++nmv;
                        wave_sync_lds();
                    }
                }
            }
        }
    }
    // per-cell accumulators (pairs counted per lane, all wavefronts)
    const int wp = wave_sum_i63(my_pairs);
    __syncthreads();
    if (lane == 63) sred[w] = wp;
    __syncthreads();
    if (tid == 0) {
        int np = 0;
        for (int q = 0; q < W; ++q) np += sred[q];
        CellAcc ca;
        ca.att = n_att; ca.acc = n_acc; ca.left = n_left; ca.hard = n_hard;
        ca.dcap = dcap_sum; ca.dccap = dccap_sum; ca.pairs = np; ca.coll = n_coll;
        ca.cand = n_cand; ca.dfin = dfin_sum; ca.moved_abs = moved;
        ca.s1rej = 0; ca.s2rej = 0;
        ca.ticks = 0; ca.tmoves = 0;
        P.acc[C.cell_off + cell] = ca;
    }
}

}  // namespace mcgb
