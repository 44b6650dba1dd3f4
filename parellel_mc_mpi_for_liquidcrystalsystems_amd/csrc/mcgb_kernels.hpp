// mcgb_kernels.hpp — HIP kernels of the checkerboard Gay-Berne sweep (gfx950, wave64).
//
// One sweep of every replica is the launch sequence
//   k_begin -> k_zero -> k_count -> k_scan -> k_scatter -> k_sort_gather   (cell-list rebuild)
//   -> k_pass x 8 (sublattice passes) -> k_end_moves -> k_energy/k_energy_reduce (re-sum, only
//   when flagged) -> [NPT: trial cell list, k_energy, k_volume] -> k_observe -> k_observe_reduce.
// Nothing in it synchronises with the host.
#pragma once

#include "mcgb_device.hpp"
#include "mcgb.h"

namespace mcgb {

// ------------------------------------------------------------------ wave helpers

__device__ inline int bcast(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float bcast(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ inline double bcast(double v) {
    unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    int lo = __builtin_amdgcn_readfirstlane((int)(u & 0xFFFFFFFFull));
    int hi = __builtin_amdgcn_readfirstlane((int)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return bcast(v);
}

// ------------------------------------------------------------------ per-sweep set-up

// Draws of sweep t (grid shift, sublattice order, NPT draws), the cell grid and step sizes.
__global__ void k_begin(const RepConst* __restrict__ RC, RepDyn* __restrict__ RD, int nrep,
                        int resum_interval) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    const uint32_t t = D.t;
    uint32_t s0[4], s1[4], s2[4];
    philox4x32_10(0u, 0u, t, kSetupTag, C.key0, C.key1, s0);
    philox4x32_10(1u, 0u, t, kSetupTag, C.key0, C.key1, s1);
    philox4x32_10(2u, 0u, t, kSetupTag, C.key0, C.key1, s2);
    D.g[0] = s0[0]; D.g[1] = s0[1]; D.g[2] = s0[2];
    D.vol[0] = s0[3]; D.vol[1] = s2[3];
    // shuffle_int_array (MCnptColloid.c:3721-3737) with Philox draws
    uint32_t ord[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    uint32_t dr[7] = {s1[0], s1[1], s1[2], s1[3], s2[0], s2[1], s2[2]};
    for (int i = 7, d = 0; i >= 1; --i, ++d) {
        uint32_t j = (uint32_t)(((uint64_t)dr[d] * (uint64_t)(i + 1)) >> 32);
        uint32_t tmp = ord[i]; ord[i] = ord[j]; ord[j] = tmp;
    }
    for (int i = 0; i < 8; ++i) D.order[i] = ord[i];
    int nc = 1;
    for (int d = 0; d < 3; ++d) {
        int n = cell_dim(D.L[d], C.rc);
        if (n > C.ncap[d]) n = C.ncap[d];
        D.n[d] = n;
        nc *= n;
        D.s[d] = D.L[d] / kTwo32;
        D.K[d] = (long long)floor(C.dmax * kTwo32 / D.L[d] + 0.5);
    }
    D.ncell = nc;
    D.nact = nc / 8;
    D.k = (C.N + nc - 1) / nc;
    D.moved_abs = 0.0;
    D.resum = (resum_interval > 0 && ((t + 1) % (uint32_t)resum_interval) == 0) ? 1 : 0;
}

// Host-requested cell list with zero shift (set_state, explicit recompute): grid from D.L.
__global__ void k_grid_g0(const RepConst* __restrict__ RC, RepDyn* __restrict__ RD, int nrep) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    int nc = 1;
    for (int d = 0; d < 3; ++d) {
        int n = cell_dim(D.L[d], C.rc);
        if (n > C.ncap[d]) n = C.ncap[d];
        D.n[d] = n;
        nc *= n;
        D.s[d] = D.L[d] / kTwo32;
        D.K[d] = (long long)floor(C.dmax * kTwo32 / D.L[d] + 0.5);
        D.g[d] = 0u;
    }
    D.ncell = nc;
    D.nact = nc / 8;
    D.k = (C.N + nc - 1) / nc;
}

__global__ void k_zero(int* __restrict__ p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ------------------------------------------------------------------ cell-list rebuild

// Cell of every particle (get_owner_block_index/get_owner_cell_index, MCnptColloid.c:4371-4399,
// as one exact integer multiply-shift per axis) and cell occupancy counts.
// trial != 0: the NPT trial grid (D.nt, zero shift).
__global__ void k_count(const RepConst* __restrict__ RC, const RepDyn* __restrict__ RD,
                        const uint32_t* __restrict__ qx, const uint32_t* __restrict__ qy,
                        const uint32_t* __restrict__ qz, int* __restrict__ cell_id,
                        int* __restrict__ count, int trial) {
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    if (trial && !D.trial_ok) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.N) return;
    const int gi = C.part_off + i;
    const uint32_t nx = trial ? D.nt[0] : D.n[0], ny = trial ? D.nt[1] : D.n[1],
                   nz = trial ? D.nt[2] : D.n[2];
    const uint32_t gx = trial ? 0u : D.g[0], gy = trial ? 0u : D.g[1], gz = trial ? 0u : D.g[2];
    const uint32_t cx = cell_coord(qx[gi], gx, nx), cy = cell_coord(qy[gi], gy, ny),
                   cz = cell_coord(qz[gi], gz, nz);
    const int c = (int)(cx + nx * (cy + ny * cz));
    cell_id[gi] = c;
    atomicAdd(&count[C.cell_off + c], 1);
}

// Exclusive scan of the counts into cell_start (one workgroup per replica); the counts are
// zeroed for reuse as fill pointers.
__global__ __launch_bounds__(1024) void k_scan(const RepConst* __restrict__ RC,
                                               RepDyn* __restrict__ RD, int* __restrict__ count,
                                               int* __restrict__ cell_start, int trial) {
    __shared__ int part[1024];
    const int r = blockIdx.x;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    if (trial && !D.trial_ok) return;
    const int ncell = trial ? D.nt[0] * D.nt[1] * D.nt[2] : D.ncell;
    const int tid = threadIdx.x;
    const int chunk = (ncell + 1023) / 1024;
    const int c0 = tid * chunk;
    const int c1 = min(ncell, c0 + chunk);
    int* cnt = count + C.cell_off;
    int sum = 0;
    for (int c = c0; c < c1; ++c) sum += cnt[c];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {       // Hillis-Steele inclusive scan
        int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - sum;
    int* cs = cell_start + C.cs_off;
    for (int c = c0; c < c1; ++c) {
        int v = cnt[c];
        cs[c] = run;
        run += v;
        cnt[c] = 0;
    }
    if (tid == 0) cs[ncell] = C.N;
}

// Unordered scatter of particle slots into their cell segments.
__global__ void k_scatter(const RepConst* __restrict__ RC, const RepDyn* __restrict__ RD,
                          const int* __restrict__ cell_id, int* __restrict__ fill,
                          const int* __restrict__ cell_start, int* __restrict__ tmp, int trial) {
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    if (trial && !RD[r].trial_ok) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.N) return;
    const int c = cell_id[C.part_off + i];
    const int slot = atomicAdd(&fill[C.cell_off + c], 1);
    tmp[C.part_off + cell_start[C.cs_off + c] + slot] = i;
}

// Orders every cell segment by particle id (a deterministic cell list: dist_batch_to_cells,
// MCnptColloid.c:4942-4964, in a fixed order).  Gather mode copies the particle SoA from
// buffer `src` into cell order in buffer `dst`; index mode (NPT trial list) writes the ordered
// slot indices to list_out.  One wavefront per cell, four cells per workgroup.
template <typename Real>
__global__ __launch_bounds__(256) void k_sort_gather(
    const RepConst* __restrict__ RC, const RepDyn* __restrict__ RD, const int* __restrict__ cell_start,
    const int* __restrict__ tmp, const int* __restrict__ pid_src, int index_only, int trial,
    int* __restrict__ list_out,
    const uint32_t* __restrict__ sqx, const uint32_t* __restrict__ sqy, const uint32_t* __restrict__ sqz,
    const Real* __restrict__ sux, const Real* __restrict__ suy, const Real* __restrict__ suz,
    uint32_t* __restrict__ dqx, uint32_t* __restrict__ dqy, uint32_t* __restrict__ dqz,
    Real* __restrict__ dux, Real* __restrict__ duy, Real* __restrict__ duz, int* __restrict__ dpid) {
    __shared__ int spid[4][256];
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    if (trial && !D.trial_ok) return;                  // uniform over the block
    const int ncell = trial ? D.nt[0] * D.nt[1] * D.nt[2] : D.ncell;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + w;
    const bool valid = c < ncell;
    const int s = valid ? cell_start[C.cs_off + c] : 0;
    const int nc = valid ? cell_start[C.cs_off + c + 1] - s : 0;
    const int* seg = tmp + C.part_off + s;
    const int* pids = pid_src + C.part_off;
    const bool in_lds = nc <= 256;
    if (in_lds)
        for (int k = lane; k < nc; k += 64) spid[w][k] = pids[seg[k]];
    __syncthreads();
    for (int k = lane; k < nc; k += 64) {
        const int idx = seg[k];
        const int p = in_lds ? spid[w][k] : pids[idx];
        int rank = 0;
        for (int j = 0; j < nc; ++j) rank += (in_lds ? spid[w][j] : pids[seg[j]]) < p;
        const int dst = C.part_off + s + rank;
        if (index_only) {
            list_out[dst] = idx;
        } else {
            const int src = C.part_off + idx;
            dqx[dst] = sqx[src]; dqy[dst] = sqy[src]; dqz[dst] = sqz[src];
            dux[dst] = sux[src]; duy[dst] = suy[src]; duz[dst] = suz[src];
            dpid[dst] = p;
        }
    }
}

// ------------------------------------------------------------------ sweep pass

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
    int pass;
    int cap;          // LDS staging capacity (particles)
};

template <typename Real>
struct PassLds {
    // byte offsets inside the dynamic LDS block
    static __host__ __device__ inline size_t bytes(int cap) {
        return 256 + (size_t)cap * (12 + 3 * sizeof(Real) + 2) + 64 * (4 + 12 + 3 * sizeof(Real) + 8) + 64;
    }
};

// Pair energy of the moved particle (orientation ui) and stencil particle j at fixed-point
// separation (ox, oy, oz) = q_j - q_i.  Returns false when the pair is beyond the cutoff (the
// decision is taken on the FP64 r^2 in both precisions).
template <typename Real>
__device__ inline bool pair_energy(const RepConst& C, const RepDyn& D, Real uix, Real uiy, Real uiz,
                                   Real ujx, Real ujy, Real ujz, int ox, int oy, int oz, Real* u,
                                   bool* cap);
template <>
__device__ inline bool pair_energy<float>(const RepConst& C, const RepDyn& D, float uix, float uiy,
                                          float uiz, float ujx, float ujy, float ujz, int ox,
                                          int oy, int oz, float* u, bool* cap) {
    const double dx = (double)ox * D.s[0], dy = (double)oy * D.s[1], dz = (double)oz * D.s[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < C.rc2)) return false;
    *u = gb_pair_f32(C, uix, uiy, uiz, ujx, ujy, ujz, (float)dx, (float)dy, (float)dz, r2, cap);
    return true;
}
template <>
__device__ inline bool pair_energy<double>(const RepConst& C, const RepDyn& D, double uix,
                                           double uiy, double uiz, double ujx, double ujy,
                                           double ujz, int ox, int oy, int oz, double* u, bool* cap) {
    const double dx = (double)ox * D.s[0], dy = (double)oy * D.s[1], dz = (double)oz * D.s[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < C.rc2)) return false;
    *u = gb_pair_f64(C, uix, uiy, uiz, ujx, ujy, ujz, dx, dy, dz, r2, cap);
    return true;
}

// Per-lane accumulator of a finite energy difference: compensated in FP32, plain in FP64.
template <typename Real> struct DeltaAcc;
template <> struct DeltaAcc<float> {
    float hi = 0.0f, lo = 0.0f;
    __device__ void add(float v) { two_sum(hi, lo, v); }
    __device__ double reduce() {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            float oh = __shfl_xor(hi, off), ol = __shfl_xor(lo, off);
            two_sum(hi, lo, oh);
            lo += ol;
        }
        return bcast((double)hi + (double)lo);
    }
};
template <> struct DeltaAcc<double> {
    double s = 0.0;
    __device__ void add(double v) { s += v; }
    __device__ double reduce() { return wave_sum(s); }
};

// Flat stencil index e -> global particle slot (segments of the 27 neighbour cells).
__device__ inline int stencil_slot(const int* s_off, const int* s_start, int part_off, int e) {
    int lo = 0, hi = 27;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s_off[mid] <= e) lo = mid; else hi = mid;
    }
    return part_off + s_start[lo] + (e - s_off[lo]);
}

// One sublattice pass: a wavefront per active cell makes the cell's k trial moves in sequence
// (the TODO body of MCnptColloid.c:683-700 filled with attempt_monte_carlo_move, :2445-2604).
// The 3x3x3 stencil is staged in LDS as SoA; each move tests every stencil particle against the
// old and trial positions, compacts those inside the cutoff of either, and evaluates both GB
// energies per listed pair across the 64 lanes.  A stencil larger than the staging capacity is
// read from global memory instead (same arithmetic, no staging).
template <typename Real>
__global__ __launch_bounds__(64) void k_pass(PassArgs P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = blockIdx.x;
    const int r = P.blk_rep[b];
    const RepConst& C = P.RC[r];
    RepDyn& D = P.RD[r];
    const int a = b - C.act_off;
    const int nact = D.nact;
    if (a >= nact) return;
    const int lane = threadIdx.x;
    const int cap = P.cap;
    Real* __restrict__ Ux = (Real*)P.ux;
    Real* __restrict__ Uy = (Real*)P.uy;
    Real* __restrict__ Uz = (Real*)P.uz;

    int* s_start = (int*)smem;            // 32
    int* s_off = s_start + 32;            // 32 (28 used)
    uint32_t* sqx = (uint32_t*)(smem + 256);
    uint32_t* sqy = sqx + cap;
    uint32_t* sqz = sqy + cap;
    Real* sux = (Real*)(sqz + cap);
    Real* suy = sux + cap;
    Real* suz = suy + cap;
    unsigned short* slist = (unsigned short*)(suz + cap);
    unsigned char* dp = (unsigned char*)(slist + cap);
    dp = (unsigned char*)(((uintptr_t)dp + 15) & ~(uintptr_t)15);
    double* sthr = (double*)dp;                       // 64
    Real* svx = (Real*)(sthr + 64);                   // 3 x 64
    Real* svy = svx + 64;
    Real* svz = svy + 64;
    int* sdqx = (int*)(svz + 64);                     // 3 x 64
    int* sdqy = sdqx + 64;
    int* sdqz = sdqy + 64;
    uint32_t* sj = (uint32_t*)(sdqz + 64);            // 64

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
    __syncthreads();
    const int S = s_off[27];
    const int coff = s_off[13];
    const int nc = s_off[14] - coff;
    const int cstart = C.part_off + s_start[13];
    const bool staged = S <= cap;
    if (!staged && lane == 0) atomicOr(&D.overflow, 1);
    if (staged) {
        for (int e = lane; e < S; e += 64) {
            int gi = stencil_slot(s_off, s_start, C.part_off, e);
            sqx[e] = P.qx[gi]; sqy[e] = P.qy[gi]; sqz[e] = P.qz[gi];
            sux[e] = Ux[gi]; suy[e] = Uy[gi]; suz[e] = Uz[gi];
        }
    }

    const int k = D.k;
    const uint32_t t = D.t;
    const size_t tbase = ((size_t)P.pass * (size_t)nact + (size_t)a) * (size_t)k;
    int n_att = 0, n_acc = 0, n_left = 0, n_hard = 0, n_pairs = 0, n_coll = 0;
    int dcap_sum = 0, dccap_sum = 0;
    long long n_cand = 0;
    double dfin_sum = 0.0, moved = 0.0;
    const float sx = (float)D.s[0], sy = (float)D.s[1], sz = (float)D.s[2];
    const float rc2 = C.f_rc2_pre;       // pre-filter; the cutoff itself is decided in FP64

    if (nc == 0) {
        if (P.trace)
            for (int m = lane; m < k; m += 64) {
                mcgb_move_trace rec = {-1, MCGB_MOVE_EMPTY, 0, 0, 0.0, 0.0};
                P.trace[tbase + m] = rec;
            }
    } else {
        for (int m0 = 0; m0 < k; m0 += 64) {
            __syncthreads();
            {   // lane-parallel draws of up to 64 moves
                const int m = m0 + lane;
                if (m < k) {
                    uint32_t A[4], B[4];
                    philox4x32_10((uint32_t)m, (uint32_t)cell, t, (uint32_t)P.pass << 8, C.key0, C.key1, A);
                    philox4x32_10((uint32_t)m, (uint32_t)cell, t, ((uint32_t)P.pass << 8) | 1u, C.key0, C.key1, B);
                    sj[lane] = A[0];
                    sdqx[lane] = step_q(A[1], D.K[0]);
                    sdqy[lane] = step_q(A[2], D.K[1]);
                    sdqz[lane] = step_q(A[3], D.K[2]);
                    double zz = 1.0 - 2.0 * u01(B[0]);
                    double sxy = sqrt(fmax(0.0, 1.0 - zz * zz));
                    double sn, cn;
                    sincospi(2.0 * u01(B[1]), &sn, &cn);
                    svx[lane] = (Real)(sxy * cn);
                    svy[lane] = (Real)(sxy * sn);
                    svz[lane] = (Real)zz;
                    sthr[lane] = -log(u01(B[2])) * C.temperature;
                }
            }
            __syncthreads();
            const int mend = min(k, m0 + 64);
            for (int m = m0; m < mend; ++m) {
                const int ml = m - m0;
                const int jsel = (int)(((uint64_t)sj[ml] * (uint32_t)nc) >> 32);
                const int li = coff + jsel;          // flat stencil index of the moved particle
                const int gsel = cstart + jsel;      // its global slot
                uint32_t qix, qiy, qiz;
                Real uix, uiy, uiz;
                if (staged) {
                    qix = sqx[li]; qiy = sqy[li]; qiz = sqz[li];
                    uix = sux[li]; uiy = suy[li]; uiz = suz[li];
                } else {
                    qix = P.qx[gsel]; qiy = P.qy[gsel]; qiz = P.qz[gsel];
                    uix = Ux[gsel]; uiy = Uy[gsel]; uiz = Uz[gsel];
                }
                const int dqx = sdqx[ml], dqy = sdqy[ml], dqz = sdqz[ml];
                const uint32_t qnx = qix + (uint32_t)dqx, qny = qiy + (uint32_t)dqy,
                               qnz = qiz + (uint32_t)dqz;
                int outcome;
                int capd = 0, ccd = 0;
                double dfin = 0.0, scale = 0.0;
                ++n_att;
                if (cell_coord(qnx, gx, (uint32_t)nx) != (uint32_t)cx ||
                    cell_coord(qny, gy, (uint32_t)ny) != (uint32_t)cy ||
                    cell_coord(qnz, gz, (uint32_t)nz) != (uint32_t)cz) {
                    outcome = MCGB_MOVE_LEFT;       // moves are confined to the active cell
                    ++n_left;
                } else {
                    const double rnx = (double)(int32_t)(qnx - 0x80000000u) * D.s[0];
                    const double rny = (double)(int32_t)(qny - 0x80000000u) * D.s[1];
                    const double rnz = (double)(int32_t)(qnz - 0x80000000u) * D.s[2];
                    if (C.colloid_on && rnx * rnx + rny * rny + rnz * rnz <= C.R2) {
                        outcome = MCGB_MOVE_HARD;   // MCnptColloid.c:2516-2520
                        ++n_hard;
                    } else {
                        // trial orientation normalize(u + d_rot v) (MCnptColloid.c:2526-2550)
                        double tx = (double)uix + C.drot * (double)svx[ml];
                        double ty = (double)uiy + C.drot * (double)svy[ml];
                        double tz = (double)uiz + C.drot * (double)svz[ml];
                        double n2 = tx * tx + ty * ty + tz * tz;
                        Real unx, uny, unz;
                        if (n2 <= 1e-12) {
                            unx = uix; uny = uiy; unz = uiz;
                        } else {
                            double nr = sqrt(n2);
                            unx = (Real)(tx / nr); uny = (Real)(ty / nr); unz = (Real)(tz / nr);
                        }
                        DeltaAcc<Real> acc;
                        Real scl = 0;
                        int pr = 0;
                        if (staged) {
                            // candidate test against both states, wave-level compaction
                            int base = 0;
                            for (int e0 = 0; e0 < S; e0 += 64) {
                                const int e = e0 + lane;
                                bool keep = false;
                                if (e < S && e != li) {
                                    const int ox = (int)(sqx[e] - qix), oy = (int)(sqy[e] - qiy),
                                              oz = (int)(sqz[e] - qiz);
                                    const float fx = (float)ox * sx, fy = (float)oy * sy, fz = (float)oz * sz;
                                    const float hx2 = (float)(ox - dqx) * sx, hy2 = (float)(oy - dqy) * sy,
                                                hz2 = (float)(oz - dqz) * sz;
                                    const float r2o = fx * fx + fy * fy + fz * fz;
                                    const float r2n = hx2 * hx2 + hy2 * hy2 + hz2 * hz2;
                                    keep = (r2o < rc2) | (r2n < rc2);
                                }
                                const unsigned long long mask = __ballot(keep);
                                if (keep) slist[base + __popcll(mask & __lanemask_lt())] = (unsigned short)e;
                                base += __popcll(mask);
                            }
                            __syncthreads();
                            const int items = 2 * base;
                            for (int it = lane; it < items; it += 64) {
                                const int e = slist[it >> 1];
                                const bool nw = it & 1;
                                const uint32_t px = nw ? qnx : qix, py = nw ? qny : qiy, pz = nw ? qnz : qiz;
                                bool cp;
                                Real u;
                                if (pair_energy<Real>(C, D, nw ? unx : uix, nw ? uny : uiy, nw ? unz : uiz,
                                                      sux[e], suy[e], suz[e], (int)(sqx[e] - px),
                                                      (int)(sqy[e] - py), (int)(sqz[e] - pz), &u, &cp)) {
                                    ++pr;
                                    if (cp) capd += nw ? 1 : -1;
                                    else {
                                        acc.add(nw ? u : -u);
                                        scl += u < 0 ? -u : u;
                                    }
                                }
                            }
                        } else {
                            // global-memory path: both states per candidate, no compaction
                            for (int e = lane; e < S; e += 64) {
                                if (e == li) continue;
                                const int gj = stencil_slot(s_off, s_start, C.part_off, e);
                                const uint32_t jx = P.qx[gj], jy = P.qy[gj], jz = P.qz[gj];
                                const Real ujx = Ux[gj], ujy = Uy[gj], ujz = Uz[gj];
#pragma unroll
                                for (int st = 0; st < 2; ++st) {
                                    const bool nw = st == 1;
                                    bool cp;
                                    Real u;
                                    if (pair_energy<Real>(C, D, nw ? unx : uix, nw ? uny : uiy, nw ? unz : uiz,
                                                          ujx, ujy, ujz, (int)(jx - (nw ? qnx : qix)),
                                                          (int)(jy - (nw ? qny : qiy)),
                                                          (int)(jz - (nw ? qnz : qiz)), &u, &cp)) {
                                        ++pr;
                                        if (cp) capd += nw ? 1 : -1;
                                        else {
                                            acc.add(nw ? u : -u);
                                            scl += u < 0 ? -u : u;
                                        }
                                    }
                                }
                            }
                        }
                        // colloid-rod term of both states (lanes 0 and 1), FP64
                        double cacc = 0.0, cscl = 0.0;
                        if (C.colloid_on && lane < 2) {
                            const bool nw = lane == 1;
                            const double rx = (double)(int32_t)((nw ? qnx : qix) - 0x80000000u) * D.s[0];
                            const double ry = (double)(int32_t)((nw ? qny : qiy) - 0x80000000u) * D.s[1];
                            const double rz = (double)(int32_t)((nw ? qnz : qiz) - 0x80000000u) * D.s[2];
                            bool cp;
                            const double u = colloid_f64(C, rx, ry, rz, (double)(nw ? unx : uix),
                                                         (double)(nw ? uny : uiy), (double)(nw ? unz : uiz), &cp);
                            if (cp) ccd = nw ? 1 : -1;
                            else {
                                cacc = nw ? u : -u;
                                cscl = fabs(u);
                            }
                        }
                        dfin = acc.reduce() + wave_sum(cacc);
                        capd = wave_sum(capd);
                        ccd = wave_sum(ccd);
                        pr = wave_sum(pr);
                        scale = wave_sum((double)scl + cscl);
                        n_pairs += pr;
                        n_cand += 2LL * (S - 1);
                        if (C.colloid_on) n_coll += 2;
                        const double dE = (double)capd * 1.0e9 + (double)ccd * 1.0e8 + dfin;
                        if (dE <= sthr[ml]) {                 // log U <= -beta dE (MCnptColloid.c:2574)
                            outcome = MCGB_MOVE_ACCEPT;
                            ++n_acc;
                            dcap_sum += capd;
                            dccap_sum += ccd;
                            dfin_sum += dfin;
                            moved += fabs(dfin);
                            if (lane == 0) {
                                if (staged) {
                                    sqx[li] = qnx; sqy[li] = qny; sqz[li] = qnz;
                                    sux[li] = unx; suy[li] = uny; suz[li] = unz;
                                }
                                P.qx[gsel] = qnx; P.qy[gsel] = qny; P.qz[gsel] = qnz;
                                Ux[gsel] = unx; Uy[gsel] = uny; Uz[gsel] = unz;
                            }
                        } else {
                            outcome = MCGB_MOVE_REJECT;
                        }
                    }
                }
                if (P.trace && lane == 0) {
                    mcgb_move_trace rec;
                    rec.pid = P.pid[gsel];
                    rec.outcome = outcome;
                    rec.dcap = capd;
                    rec.dccap = ccd;
                    rec.dfin = dfin;
                    rec.scale = scale;
                    P.trace[tbase + m] = rec;
                }
                __syncthreads();
            }
        }
    }
    if (lane == 0) {
        CellAcc ca;
        ca.att = n_att; ca.acc = n_acc; ca.left = n_left; ca.hard = n_hard;
        ca.dcap = dcap_sum; ca.dccap = dccap_sum; ca.pairs = n_pairs; ca.coll = n_coll;
        ca.cand = n_cand; ca.dfin = dfin_sum; ca.moved_abs = moved; ca.pad = 0.0;
        P.acc[C.cell_off + cell] = ca;
    }
}

// ------------------------------------------------------------------ end of the passes

// Folds the per-cell accumulators of the sweep into the running energy and counters (fixed
// reduction order: deterministic) and flags an FP64 re-summation when needed.
__global__ __launch_bounds__(256) void k_end_moves(const RepConst* __restrict__ RC,
                                                   RepDyn* __restrict__ RD,
                                                   const CellAcc* __restrict__ acc) {
    __shared__ long long si[256][9];
    __shared__ double sd[256][2];
    const int r = blockIdx.x;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    const int tid = threadIdx.x;
    long long v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double f = 0.0, mv = 0.0;
    for (int c = tid; c < D.ncell; c += 256) {
        const CellAcc& A = acc[C.cell_off + c];
        v[0] += A.att; v[1] += A.acc; v[2] += A.left; v[3] += A.hard;
        v[4] += A.cand; v[5] += A.pairs; v[6] += A.coll; v[7] += A.dcap; v[8] += A.dccap;
        f += A.dfin;
        mv += A.moved_abs;
    }
    for (int q = 0; q < 9; ++q) si[tid][q] = v[q];
    sd[tid][0] = f;
    sd[tid][1] = mv;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (tid < off) {
            for (int q = 0; q < 9; ++q) si[tid][q] += si[tid + off][q];
            sd[tid][0] += sd[tid + off][0];
            sd[tid][1] += sd[tid + off][1];
        }
        __syncthreads();
    }
    if (tid == 0) {
        D.attempted += si[0][0]; D.accepted += si[0][1]; D.left += si[0][2]; D.hard += si[0][3];
        D.cand += si[0][4]; D.pairs += si[0][5]; D.coll += si[0][6];
        D.cap_pairs += si[0][7];
        D.cap_coll += si[0][8];
        D.e_fin += sd[0][0];
        D.moved_abs = sd[0][1];
        if (sd[0][1] > 1.0e10) D.resum = 1;
    }
}

// ------------------------------------------------------------------ FP64 full energy

// Per-particle FP64 energy over a cell list (each unordered pair once: the partner must come
// later in the list), colloid term and colloid-core overlap flag.  mode: 0 always; 1 only when
// the replica flagged a re-summation; 2 the NPT trial box (indirect list, zero shift).
// tot_ene, MCnptColloid.c:1598-1622.
template <typename Real>
__global__ __launch_bounds__(128) void k_energy(
    const RepConst* __restrict__ RC, const RepDyn* __restrict__ RD, int mode,
    const uint32_t* __restrict__ qx, const uint32_t* __restrict__ qy, const uint32_t* __restrict__ qz,
    const Real* __restrict__ ux, const Real* __restrict__ uy, const Real* __restrict__ uz,
    const int* __restrict__ cell_start, const int* __restrict__ list,
    double* __restrict__ efin, int* __restrict__ ecap, int* __restrict__ eccap, int* __restrict__ ehard) {
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    if (mode == 1 && !D.resum) return;
    if (mode == 2 && !D.trial_ok) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.N) return;
    const bool ind = mode == 2;
    const int n0 = ind ? D.nt[0] : D.n[0], n1 = ind ? D.nt[1] : D.n[1], n2 = ind ? D.nt[2] : D.n[2];
    const double L0 = ind ? D.Lt[0] : D.L[0], L1 = ind ? D.Lt[1] : D.L[1], L2 = ind ? D.Lt[2] : D.L[2];
    const double s0 = L0 / kTwo32, s1 = L1 / kTwo32, s2 = L2 / kTwo32;
    const uint32_t g0 = ind ? 0u : D.g[0], g1 = ind ? 0u : D.g[1], g2 = ind ? 0u : D.g[2];
    const int po = C.part_off;
    const int gi = po + (ind ? list[po + i] : i);
    const uint32_t xi = qx[gi], yi = qy[gi], zi = qz[gi];
    const double uix = ux[gi], uiy = uy[gi], uiz = uz[gi];
    const int cx = (int)cell_coord(xi, g0, n0), cy = (int)cell_coord(yi, g1, n1),
              cz = (int)cell_coord(zi, g2, n2);
    const int* cs = cell_start + C.cs_off;
    double f = 0.0;
    int cp = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int ex = cx + dx, ey = cy + dy, ez = cz + dz;
                ex += ex < 0 ? n0 : 0; ex -= ex >= n0 ? n0 : 0;
                ey += ey < 0 ? n1 : 0; ey -= ey >= n1 ? n1 : 0;
                ez += ez < 0 ? n2 : 0; ez -= ez >= n2 ? n2 : 0;
                const int nb = ex + n0 * (ey + n1 * ez);
                const int b0 = cs[nb], b1 = cs[nb + 1];
                for (int kk = b0; kk < b1; ++kk) {
                    if (kk <= i) continue;
                    const int gj = po + (ind ? list[po + kk] : kk);
                    const double ddx = (double)(int32_t)(qx[gj] - xi) * s0;
                    const double ddy = (double)(int32_t)(qy[gj] - yi) * s1;
                    const double ddz = (double)(int32_t)(qz[gj] - zi) * s2;
                    const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (r2 < C.rc2) {
                        bool c;
                        const double u = gb_pair_f64(C, uix, uiy, uiz, (double)ux[gj], (double)uy[gj],
                                                     (double)uz[gj], ddx, ddy, ddz, r2, &c);
                        if (c) ++cp; else f += u;
                    }
                }
            }
    int ccp = 0, hard = 0;
    if (C.colloid_on) {
        const double rx = (double)(int32_t)(xi - 0x80000000u) * s0;
        const double ry = (double)(int32_t)(yi - 0x80000000u) * s1;
        const double rz = (double)(int32_t)(zi - 0x80000000u) * s2;
        bool c;
        const double u = colloid_f64(C, rx, ry, rz, uix, uiy, uiz, &c);
        if (c) ccp = 1; else f += u;
        hard = (rx * rx + ry * ry + rz * rz <= C.R2) ? 1 : 0;
    }
    efin[po + i] = f;
    ecap[po + i] = cp;
    eccap[po + i] = ccp;
    ehard[po + i] = hard;
}

// Deterministic per-replica reduction of k_energy's output into D.e_res, then by mode:
// 0 replace the running energy; 1 the same when a re-summation was flagged (records the drift);
// 2 the NPT ln V acceptance test (absent from the reference, whose NPT is a label only,
// MCnptColloid.c:1197, :1291); 3 only D.e_res.
__global__ __launch_bounds__(256) void k_energy_reduce(const RepConst* __restrict__ RC,
                                                       RepDyn* __restrict__ RD, int mode,
                                                       const double* __restrict__ efin,
                                                       const int* __restrict__ ecap,
                                                       const int* __restrict__ eccap,
                                                       const int* __restrict__ ehard) {
    __shared__ double sf[256];
    __shared__ long long sc[256][3];
    const int r = blockIdx.x;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    if (mode == 1 && !D.resum) return;
    if (mode == 2 && !D.trial_ok) return;
    const int tid = threadIdx.x;
    double f = 0.0;
    long long a = 0, b = 0, h = 0;
    for (int i = tid; i < C.N; i += 256) {
        f += efin[C.part_off + i];
        a += ecap[C.part_off + i];
        b += eccap[C.part_off + i];
        h += ehard[C.part_off + i];
    }
    sf[tid] = f; sc[tid][0] = a; sc[tid][1] = b; sc[tid][2] = h;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (tid < off) {
            sf[tid] += sf[tid + off];
            for (int q = 0; q < 3; ++q) sc[tid][q] += sc[tid + off][q];
        }
        __syncthreads();
    }
    if (tid != 0) return;
    D.e_res[0] = (double)sc[0][0]; D.e_res[1] = (double)sc[0][1];
    D.e_res[2] = sf[0]; D.e_res[3] = (double)sc[0][2];
    if (mode == 3) return;                             // compute only
    if (mode == 0 || mode == 1) {
        if (mode == 1) {
            double run = (double)D.cap_pairs * 1e9 + (double)D.cap_coll * 1e8 + D.e_fin;
            double now = sc[0][0] * 1e9 + sc[0][1] * 1e8 + sf[0];
            double dr = fabs(run - now);
            if (dr > D.max_drift && D.moved_abs <= 1.0e10) D.max_drift = dr;
        }
        D.cap_pairs = sc[0][0];
        D.cap_coll = sc[0][1];
        D.e_fin = sf[0];
        return;
    }
    // mode 2: accept or reject the trial box
    if (sc[0][2] > 0) return;                          // a rod inside the colloid core
    const double V = D.L[0] * D.L[1] * D.L[2], Vn = D.Lt[0] * D.Lt[1] * D.Lt[2];
    const double dU = ((double)sc[0][0] - (double)D.cap_pairs) * 1.0e9 +
                      ((double)sc[0][1] - (double)D.cap_coll) * 1.0e8 + (sf[0] - D.e_fin);
    const double dlnv = C.npt_max_dlnv * (2.0 * u01(D.vol[0]) - 1.0);
    const double arg = -C.beta * (dU + C.pressure * (Vn - V)) + (double)(C.N + 1) * dlnv;
    if (log(u01(D.vol[1])) <= arg) {
        D.vol_acc += 1;
        for (int d = 0; d < 3; ++d) D.L[d] = D.Lt[d];
        D.cap_pairs = sc[0][0];
        D.cap_coll = sc[0][1];
        D.e_fin = sf[0];
    }
}

// NPT: trial box L' = L exp(d lnV / 3) and its cell grid (trial_ok = 0 when too small).
__global__ void k_npt_trial(const RepConst* __restrict__ RC, RepDyn* __restrict__ RD, int nrep) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    D.trial_ok = 0;
    if (!C.npt) return;
    D.vol_att += 1;
    const double dlnv = C.npt_max_dlnv * (2.0 * u01(D.vol[0]) - 1.0);
    const double f = exp(dlnv / 3.0);
    int ok = 1;
    for (int d = 0; d < 3; ++d) {
        D.Lt[d] = D.L[d] * f;
        int n = cell_dim(D.Lt[d], C.rc);      // wider cells than needed are still exact
        if (n == 0) ok = 0;
        D.nt[d] = n > C.ncap[d] ? C.ncap[d] : n;
    }
    D.trial_ok = ok;
}

// ------------------------------------------------------------------ observables

// Partial FP64 sums of the Q tensor, 64 blocks per replica
// (compute_liquid_crystal_order_parameters, MCnptColloid.c:2149-2167).
template <typename Real>
__global__ __launch_bounds__(256) void k_observe(const RepConst* __restrict__ RC,
                                                 const Real* __restrict__ ux,
                                                 const Real* __restrict__ uy,
                                                 const Real* __restrict__ uz,
                                                 double* __restrict__ part) {
    __shared__ double sq[256][6];
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    const int tid = threadIdx.x;
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * 256 + tid; i < C.N; i += 64 * 256) {
        const double x = ux[C.part_off + i], y = uy[C.part_off + i], z = uz[C.part_off + i];
        q[0] += x * x; q[1] += y * y; q[2] += z * z; q[3] += x * y; q[4] += x * z; q[5] += y * z;
    }
    for (int k = 0; k < 6; ++k) sq[tid][k] = q[k];
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (tid < off)
            for (int k = 0; k < 6; ++k) sq[tid][k] += sq[tid + off][k];
        __syncthreads();
    }
    if (tid == 0)
        for (int k = 0; k < 6; ++k) part[((size_t)r * 64 + blockIdx.x) * 6 + k] = sq[0][k];
}

constexpr int kSeriesWidth = 12;   // E/N, Q(6)/N, density, L(3), sweep index

// Appends the sweep's observables to the series ring and advances the sweep index.
__global__ void k_observe_reduce(const RepConst* __restrict__ RC, RepDyn* __restrict__ RD, int nrep,
                                 const double* __restrict__ part, double* __restrict__ series,
                                 int series_cap, int advance) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < 64; ++b)
        for (int k = 0; k < 6; ++k) q[k] += part[((size_t)r * 64 + b) * 6 + k];
    if (series_cap > 0) {
        double* e = series + ((size_t)r * series_cap + (size_t)(D.series_count % series_cap)) * kSeriesWidth;
        const double E = (double)D.cap_pairs * 1e9 + (double)D.cap_coll * 1e8 + D.e_fin;
        e[0] = E / C.N;
        for (int k = 0; k < 6; ++k) e[1 + k] = q[k] / C.N;
        const double V = D.L[0] * D.L[1] * D.L[2];
        e[7] = C.N / (V - C.vcoll);
        e[8] = D.L[0]; e[9] = D.L[1]; e[10] = D.L[2];
        e[11] = (double)D.t;
        D.series_count += 1;
    }
    if (advance) D.t += 1;
}

// ------------------------------------------------------------------ batch kernels (parity hooks)

__global__ void k_pair_batch(RepConst C, int n, int prec, const double* __restrict__ ei,
                             const double* __restrict__ ej, const double* __restrict__ dr,
                             double* __restrict__ u, int* __restrict__ capped) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    bool cp;
    double v;
    if (prec == 64) {
        double dx = dr[3 * k], dy = dr[3 * k + 1], dz = dr[3 * k + 2];
        v = gb_pair_f64(C, ei[3 * k], ei[3 * k + 1], ei[3 * k + 2], ej[3 * k], ej[3 * k + 1],
                        ej[3 * k + 2], dx, dy, dz, dx * dx + dy * dy + dz * dz, &cp);
    } else {
        // the FP32 kernel takes unit vectors (the sweep stores normalised orientations)
        double ni = 1.0 / sqrt(ei[3 * k] * ei[3 * k] + ei[3 * k + 1] * ei[3 * k + 1] + ei[3 * k + 2] * ei[3 * k + 2]);
        double nj = 1.0 / sqrt(ej[3 * k] * ej[3 * k] + ej[3 * k + 1] * ej[3 * k + 1] + ej[3 * k + 2] * ej[3 * k + 2]);
        const double ddx = dr[3 * k], ddy = dr[3 * k + 1], ddz = dr[3 * k + 2];
        v = (double)gb_pair_f32(C, (float)(ei[3 * k] * ni), (float)(ei[3 * k + 1] * ni),
                                (float)(ei[3 * k + 2] * ni), (float)(ej[3 * k] * nj),
                                (float)(ej[3 * k + 1] * nj), (float)(ej[3 * k + 2] * nj), (float)ddx,
                                (float)ddy, (float)ddz, ddx * ddx + ddy * ddy + ddz * ddz, &cp);
    }
    capped[k] = cp ? 1 : 0;
    u[k] = cp ? 1.0e9 : v;
}

__global__ void k_colloid_batch(RepConst C, int n, const double* __restrict__ rel,
                                const double* __restrict__ e, double* __restrict__ u,
                                int* __restrict__ capped) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    bool cp;
    double v = colloid_f64(C, rel[3 * k], rel[3 * k + 1], rel[3 * k + 2], e[3 * k], e[3 * k + 1],
                           e[3 * k + 2], &cp);
    capped[k] = cp ? 1 : 0;
    u[k] = cp ? 1.0e8 : v;
}

__global__ void k_philox_batch(int n, const uint32_t* __restrict__ ctr, const uint32_t* __restrict__ key,
                               uint32_t* __restrict__ out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t o[4];
    philox4x32_10(ctr[4 * k], ctr[4 * k + 1], ctr[4 * k + 2], ctr[4 * k + 3], key[2 * k], key[2 * k + 1], o);
    for (int q = 0; q < 4; ++q) out[4 * k + q] = o[q];
}

}  // namespace mcgb
