// mcgb_kernels.hpp — HIP kernels of the checkerboard Gay-Berne sweep (gfx950, wave64).
//
// One sweep of every replica is the launch sequence
//   k_begin -> k_count_draws -> k_scan -> k_scatter -> k_sort_gather   (draws, cell-list rebuild)
//   -> k_pass x P (sublattice passes, P = 8 or 27) -> k_end (sweep totals, orientation sums)
//   -> k_energy/k_energy_reduce (re-sum, only when flagged; the sweep's record)
//   [NPT: -> trial cell list, k_energy, volume test -> k_record].
// Nothing in it synchronises with the host.
#pragma once

#include "mcgb_device.hpp"
#include "mcgb_pass.hpp"
#include "mcgb_spec.hpp"
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

__device__ inline long long bcast(long long v) {
    unsigned long long u = (unsigned long long)v;
    int lo = __builtin_amdgcn_readfirstlane((int)(u & 0xFFFFFFFFull));
    int hi = __builtin_amdgcn_readfirstlane((int)(u >> 32));
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return bcast(v);
}

// ------------------------------------------------------------------ per-sweep set-up

// Fixed-point step amplitude round(d_max 2^32 / L), held below 2^31 so that step_q's product
// stays within 63 bits (d_max < L/2 is enforced where the step sizes are set).
__device__ inline long long step_amplitude(float dmax, double L) {
    const long long K = (long long)floor((double)dmax * kTwo32 / L + 0.5);
    return K < 2147483647LL ? K : 2147483647LL;
}

// Step-size controller of the reference's serial driver (MCnptColloid.c:987-1031) on the
// acceptance acc/att of the moves since its last action.  The step sizes are floats and the
// constants doubles there, so each update is a double expression rounded to float; the
// explicit-rounding builtins keep the compiler from fusing them into an FMA.  An increase
// that would reach half the smallest box edge is not made (the reference has no such bound;
// fixed-point steps need it).
__device__ inline void adapt_steps(float& dmax, float& drot, unsigned long long acc,
                                   unsigned long long att, float min_thick, double half_box) {
    const float ratio = (float)acc / (float)att;
    const double d = (double)dmax, mt = (double)min_thick;
    if ((double)ratio < 0.3) {
        float nd = (float)__dsub_rn(d, __dmul_rn(d, 0.05));
        if ((double)nd < __dmul_rn(0.0001, mt)) {
            nd = (float)__dmul_rn(0.01, mt);
            if ((double)drot > 0.000001) drot = (float)__dsub_rn((double)drot, __dmul_rn(0.5, (double)drot));
        }
        dmax = nd;
    } else if ((double)ratio > 0.55) {
        float nd = (float)__dadd_rn(d, __dmul_rn(d, 0.05));
        float nr = drot;
        if ((double)nd > __dmul_rn(1.5, mt)) {
            nd = (float)__dmul_rn(0.75, mt);
            if ((double)drot < 0.7) nr = (float)__dadd_rn((double)drot, __dmul_rn(0.5, (double)drot));
        }
        if ((double)nd < half_box) {
            dmax = nd;
            drot = nr;
        }
    }
}

// Draws of sweep t (grid shift, sublattice order, NPT draws), the cell grid and step sizes.
__global__ __launch_bounds__(64) void k_begin(const RepConst* __restrict__ RC,
                                              RepDyn* __restrict__ RD, int nrep,
                                              int resum_interval, int* __restrict__ ovf_count) {
    __shared__ uint32_t sdr[64][26];                 // shuffle draws and order of each replica:
    __shared__ uint32_t sord[64][27];                // data-dependent indices stay in LDS
    if (ovf_count && blockIdx.x == 0 && threadIdx.x < 27) ovf_count[threadIdx.x] = 0;
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    const uint32_t t = D.t;
    const int P = C.col * C.col * C.col;
    uint32_t s[8][4];
#pragma unroll
    for (int j = 0; j < 8; ++j) philox4x32_10((uint32_t)j, 0u, t, kSetupTag, C.key0, C.key1, s[j]);
    D.g[0] = s[0][0]; D.g[1] = s[0][1]; D.g[2] = s[0][2];
    D.vol[0] = s[0][3]; D.vol[1] = s[2][3];
    // shuffle_int_array (MCnptColloid.c:3721-3737) with Philox draws: S_1[0..3], S_2[0..2],
    // then S_3..S_7 in order (S_7[3] unused)
    uint32_t* dr = sdr[threadIdx.x];
    uint32_t* ord = sord[threadIdx.x];
#pragma unroll
    for (int w = 0; w < 4; ++w) dr[w] = s[1][w];
#pragma unroll
    for (int w = 0; w < 3; ++w) dr[4 + w] = s[2][w];
#pragma unroll
    for (int j = 3; j < 8; ++j)
#pragma unroll
        for (int w = 0; w < 4; ++w)
            if (7 + 4 * (j - 3) + w < 26) dr[7 + 4 * (j - 3) + w] = s[j][w];
    for (int i = 0; i < P; ++i) ord[i] = (uint32_t)i;
    for (int i = P - 1, d = 0; i >= 1; --i, ++d) {
        uint32_t j = (uint32_t)(((uint64_t)dr[d] * (uint64_t)(i + 1)) >> 32);
        uint32_t tmp = ord[i]; ord[i] = ord[j]; ord[j] = tmp;
    }
    for (int i = 0; i < P; ++i) D.order[i] = ord[i];
    int nn[3];
    for (int d = 0; d < 3; ++d) nn[d] = grid_dim(D.L[d], C.rc, C.col);
    fit_grid(nn, C.ncap[0] * C.ncap[1] * C.ncap[2], C.col);
    int nc = 1;
    for (int d = 0; d < 3; ++d) {
        D.n[d] = nn[d];
        nc *= nn[d];
        D.s[d] = D.L[d] / kTwo32;
        D.K[d] = step_amplitude(D.dmax, D.L[d]);
    }
    D.ncell = nc;
    D.nact = nc / P;
    D.k = (C.N + nc - 1) / nc;
    D.moved_abs = 0.0;
    for (int k = 0; k < 8; ++k) D.moved_q[k] = 0ull;
    D.resum = (resum_interval > 0 && ((t + 1) % (uint32_t)resum_interval) == 0) ? 1 : 0;
}


// Host-requested cell list with zero shift (set_state, explicit recompute): grid from D.L.
__global__ void k_grid_g0(const RepConst* __restrict__ RC, RepDyn* __restrict__ RD, int nrep) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    int nn[3];
    for (int d = 0; d < 3; ++d) nn[d] = grid_dim(D.L[d], C.rc, C.col);
    fit_grid(nn, C.ncap[0] * C.ncap[1] * C.ncap[2], C.col);
    int nc = 1;
    for (int d = 0; d < 3; ++d) {
        D.n[d] = nn[d];
        nc *= nn[d];
        D.s[d] = D.L[d] / kTwo32;
        D.K[d] = step_amplitude(D.dmax, D.L[d]);
        D.g[d] = 0u;
    }
    D.ncell = nc;
    D.nact = nc / (C.col * C.col * C.col);
    D.k = (C.N + nc - 1) / nc;
}


// ------------------------------------------------------------------ cell-list rebuild

// Cell of every particle (get_owner_block_index/get_owner_cell_index, MCnptColloid.c:4371-4399,
// as one exact integer multiply-shift per axis) and cell occupancy counts.
// trial != 0: the NPT trial grid (D.nt, zero shift).
// One atomic per distinct cell of a wavefront: lane counts of equal cells are summed first.
// The particles arrive in the previous sweep's cell order, so a wavefront spans a few cells.
// Returns the lane's slot: the counter's old value plus its rank among the lanes of its cell
// (the slot order within a cell is arbitrary; k_sort_gather orders every cell by id).
__device__ inline int wave_cell_add(int* __restrict__ ctr, int c, bool valid) {
    const int lane = threadIdx.x & 63;
    // the distinct cells first (no memory traffic): each lane learns its cell's leader (lowest
    // lane) and its rank among the cell's lanes, each leader the cell's lane count
    int leader = lane, rank = 0, n = 0;
    unsigned long long pending = __builtin_amdgcn_ballot_w64(valid);
    while (pending) {                                  // uniform: a ballot
        const int ld = __builtin_ctzll(pending);
        const int lc = __builtin_amdgcn_readlane(c, ld);
        const unsigned long long m = __builtin_amdgcn_ballot_w64(valid && c == lc);
        if ((m >> lane) & 1ull) {
            leader = ld;
            rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        }
        if (lane == ld) n = __popcll(m);
        pending &= ~m;
    }
    // then one atomic per distinct cell, all in flight together (the leaders' lanes)
    int old = 0;
    if (valid && leader == lane) old = atomicAdd(&ctr[c], n);
    old = __shfl(old, leader);
    return valid ? old + rank : 0;
}

// Cell and slot of particle i of replica r (every lane of the wavefront calls it: the counting
// is a wavefront-wide operation; valid = i < N).  The slot is the particle's place among the
// members of its cell, so the scatter after the scan needs no atomics.
__device__ inline void count_particle(const RepConst& C, const RepDyn& D, int i,
                                      const uint32_t* __restrict__ qx, const uint32_t* __restrict__ qy,
                                      const uint32_t* __restrict__ qz, int* __restrict__ cell_id,
                                      int* __restrict__ slot, int* __restrict__ count, int trial) {
    const bool valid = i < C.N;
    int c = 0;
    const int gi = C.part_off + i;
    if (valid) {
        const uint32_t nx = trial ? D.nt[0] : D.n[0], ny = trial ? D.nt[1] : D.n[1],
                       nz = trial ? D.nt[2] : D.n[2];
        const uint32_t gx = trial ? 0u : D.g[0], gy = trial ? 0u : D.g[1], gz = trial ? 0u : D.g[2];
        const uint32_t cx = cell_coord(qx[gi], gx, nx), cy = cell_coord(qy[gi], gy, ny),
                       cz = cell_coord(qz[gi], gz, nz);
        c = (int)(cx + nx * (cy + ny * cz));
    }
    const int s = wave_cell_add(count + C.cell_off, c, valid);
    if (valid) {
        cell_id[gi] = c;
        slot[gi] = s;
    }
}

__global__ void k_count(const RepConst* __restrict__ RC, const RepDyn* __restrict__ RD,
                        const uint32_t* __restrict__ qx, const uint32_t* __restrict__ qy,
                        const uint32_t* __restrict__ qz, int* __restrict__ cell_id,
                        int* __restrict__ slot, int* __restrict__ count, int trial) {
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    if (trial && !D.trial_ok) return;                  // uniform: the whole grid row
    count_particle(C, D, blockIdx.x * blockDim.x + threadIdx.x, qx, qy, qz, cell_id, slot, count,
                   trial);
}

constexpr int kScanTile = 1024;    // cells per tile of the count scan

__device__ inline int scan_ncell(const RepDyn& D, int trial) {
    return trial ? D.nt[0] * D.nt[1] * D.nt[2] : D.ncell;
}

// Exclusive scan of the counts into cell_start in one launch over tiles of kScanTile cells,
// with decoupled look-back (Merrill & Garland, "Single-pass parallel prefix scan with decoupled
// look-back", 2016): tile t publishes its aggregate, then its inclusive prefix once it has
// looked back through the published words of tiles t-1, t-2, ... (lower tiles are dispatched
// first, so every wait ends).  A status word is (kind << 32) | value, kind 1 = aggregate,
// 2 = inclusive prefix, 0 = not yet published; k_scatter, which runs after every scan, clears
// them for the next one.  The counts are zeroed for reuse as fill pointers; trial != 0: the NPT
// trial grid.
__global__ __launch_bounds__(256) void k_scan(const RepConst* __restrict__ RC,
                                              const RepDyn* __restrict__ RD,
                                              int* __restrict__ count,
                                              unsigned long long* __restrict__ status, int tmax,
                                              int* __restrict__ cell_start, int trial) {
    __shared__ int ws[4];
    __shared__ int s_excl;
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    if (trial && !D.trial_ok) return;
    const int ncell = scan_ncell(D, trial);
    const int t = blockIdx.x;
    if (t * kScanTile >= ncell) return;                     // uniform over the block
    int* cnt = count + C.cell_off;
    int* cs = cell_start + C.cs_off;
    unsigned long long* st = status + (size_t)r * tmax;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c0 = t * kScanTile + threadIdx.x * 4;
    int v[4], sum = 0;
    for (int j = 0; j < 4; ++j) {
        v[j] = c0 + j < ncell ? cnt[c0 + j] : 0;
        sum += v[j];
    }
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int agg = ws[0] + ws[1] + ws[2] + ws[3];
        int excl = 0;
        if (t == 0) {
            __hip_atomic_store(&st[0], (2ull << 32) | (unsigned)agg, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&st[t], (1ull << 32) | (unsigned)agg, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            for (int j = t - 1; j >= 0;) {
                const unsigned long long x =
                    __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned kind = (unsigned)(x >> 32);
                if (kind == 0u) continue;                   // not yet published: poll again
                excl += (int)(unsigned)x;
                if (kind == 2u) break;
                --j;
            }
            __hip_atomic_store(&st[t], (2ull << 32) | (unsigned)(excl + agg), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_excl = excl;
    }
    __syncthreads();
    int run = s_excl + incl - sum;
    for (int q = 0; q < w; ++q) run += ws[q];
    for (int j = 0; j < 4; ++j) {
        if (c0 + j < ncell) {
            cs[c0 + j] = run;
            cnt[c0 + j] = 0;
        }
        run += v[j];
    }
    if (t == 0 && threadIdx.x == 0) cs[ncell] = C.N;
}

// Scatter of particle slots and ids into their cell segments (the slot within the cell came
// with the count); block 0 of each replica row clears the scan's status words for the next
// scan.
__global__ void k_scatter(const RepConst* __restrict__ RC, const RepDyn* __restrict__ RD,
                          const int* __restrict__ cell_id, const int* __restrict__ slot,
                          const int* __restrict__ cell_start, const int* __restrict__ pid,
                          int* __restrict__ tmp, int* __restrict__ tmp_pid,
                          unsigned long long* __restrict__ status, int tmax, int trial) {
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    if (trial && !RD[r].trial_ok) return;              // uniform: the whole grid row
    if (blockIdx.x == 0)
        for (int t = threadIdx.x; t < tmax; t += blockDim.x) status[(size_t)r * tmax + t] = 0ull;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.N) return;
    const int gi = C.part_off + i;
    const int dst = C.part_off + cell_start[C.cs_off + cell_id[gi]] + slot[gi];
    tmp[dst] = i;
    tmp_pid[dst] = pid[gi];
}

// Orders every cell segment by particle id (a deterministic cell list: dist_batch_to_cells,
// MCnptColloid.c:4942-4964, in a fixed order).  Gather mode copies the particle SoA from
// buffer `src` into cell order in buffer `dst`; index mode (NPT trial list) writes the ordered
// slot indices to list_out.  One wavefront per cell, four cells per workgroup.
template <typename Real>
__global__ __launch_bounds__(256) void k_sort_gather(
    const RepConst* __restrict__ RC, const RepDyn* __restrict__ RD, const int* __restrict__ cell_start,
    const int* __restrict__ tmp, const int* __restrict__ tmp_pid, int index_only, int trial,
    int* __restrict__ list_out, int* __restrict__ count,
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
    // the fill pointers of k_scatter are not read again: back to zero for the next rebuild
    if (valid && lane == 0) count[C.cell_off + c] = 0;
    const int* seg = tmp + C.part_off + s;
    const int* pids = tmp_pid + C.part_off + s;          // the segment's ids, in slot order
    const bool in_lds = nc <= 256;
    if (in_lds)
        for (int k = lane; k < nc; k += 64) spid[w][k] = pids[k];
    __syncthreads();
    for (int k = lane; k < nc; k += 64) {
        const int idx = seg[k];
        const int p = in_lds ? spid[w][k] : pids[k];
        int rank = 0;
        for (int j = 0; j < nc; ++j) rank += (in_lds ? spid[w][j] : pids[j]) < p;
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

// ------------------------------------------------------------------ per-move draws

// Draws of every trial move of the sweep (all 8 passes): record p*nact*k + a*k + m of the
// replica holds move m of active cell a in pass p: particle selector, symmetric fixed-point step
// (MCnptColloid.c:2502-2509), rotation vector on the unit sphere (:2526-2550) and the
// Metropolis threshold -log(U) T (:2566-2574), from A = philox({m, cell, t, p<<8}) and
// B = philox({m, cell, t, (p<<8)|1}).
__device__ inline void draw_move(const RepConst& C, const RepDyn& D, MoveDraw* __restrict__ rec,
                                 int idx) {
    const int nact = D.nact, k = D.k;
    const int col = C.col;
    const int total = col * col * col * nact * k;
    if (idx >= total) return;
    const int p = idx / (nact * k);
    const int rem = idx - p * nact * k;
    const int a = rem / k;
    const int m = rem - a * k;
    const int nx = D.n[0], ny = D.n[1];
    const int hx = nx / col, hy = ny / col;
    const uint32_t sig = D.order[p];
    const int cx = col * (a % hx) + (int)(sig % (uint32_t)col),
              cy = col * ((a / hx) % hy) + (int)((sig / (uint32_t)col) % (uint32_t)col),
              cz = col * (a / (hx * hy)) + (int)(sig / (uint32_t)(col * col));
    const uint32_t cell = (uint32_t)(cx + nx * (cy + ny * cz));
    uint32_t A[4], B[4];
    philox4x32_10((uint32_t)m, cell, D.t, (uint32_t)p << 8, C.key0, C.key1, A);
    philox4x32_10((uint32_t)m, cell, D.t, ((uint32_t)p << 8) | 1u, C.key0, C.key1, B);
    MoveDraw d;
    d.j = A[0];
    d.dq[0] = step_q(A[1], D.K[0]);
    d.dq[1] = step_q(A[2], D.K[1]);
    d.dq[2] = step_q(A[3], D.K[2]);
    const double zz = 1.0 - 2.0 * u01(B[0]);
    const double sxy = sqrt(fmax(0.0, 1.0 - zz * zz));
    double sn, cn;
    sincospi(2.0 * u01(B[1]), &sn, &cn);
    d.v[0] = sxy * cn;
    d.v[1] = sxy * sn;
    d.v[2] = zz;
    d.thr = -log(u01(B[2])) * C.temperature;
    d.thr1 = -log(u01(B[3])) * C.temperature;
    rec[C.rec_off + idx] = d;
}

// The sweep's move draws and the cell counts of the rebuild in one launch (both follow k_begin
// and neither reads the other's output): blocks [0, nbc) count particles, the others draw
// moves, so the counting's atomic round trips and the draws' arithmetic run side by side.
__global__ __launch_bounds__(256) void k_count_draws(const RepConst* __restrict__ RC,
                                                     const RepDyn* __restrict__ RD,
                                                     MoveDraw* __restrict__ rec,
                                                     const uint32_t* __restrict__ qx,
                                                     const uint32_t* __restrict__ qy,
                                                     const uint32_t* __restrict__ qz,
                                                     int* __restrict__ cell_id, int* __restrict__ slot,
                                                     int* __restrict__ count, int nbc) {
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    if ((int)blockIdx.x < nbc) {                       // uniform over the block
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (blockIdx.x * blockDim.x < (unsigned)C.N)   // wavefronts holding particles
            count_particle(C, D, i, qx, qy, qz, cell_id, slot, count, 0);
    } else {
        draw_move(C, D, rec, (blockIdx.x - nbc) * blockDim.x + threadIdx.x);
    }
}

// ------------------------------------------------------------------ end of the passes

// End of the passes, one launch (k_end): block b of replica r folds the per-cell accumulators
// of cell chunk b and the orientation products u_a u_b of particles b*256 + tid + j*kEndParts*256
// into a partial, and adds its moved |dE|, rounded to whole energy units, to the replica's
// integer total (the re-summation test of k_energy: an integer sum is the same in any order).
// The next launches fold the partials (end_fold in k_energy_reduce): no fence or ticket here,
// the launch boundary orders them.  Every sum runs in a fixed order (four particles' loads
// issued together, lane butterflies, then the wavefronts in order): deterministic.
constexpr int kEndParts = 256;
constexpr int kEndCnt = 13;        // 11 move counters, the cells' ticks and moves
constexpr int kEndVals = kEndCnt + 8;   // and finite dE, moved |dE|, 6 orientation sums
struct EndPartial {
    long long v[kEndCnt];
    double d[8];                   // finite dE, moved |dE|, orientation sums
};
// a sweep whose accepted moves moved more than this much finite energy is re-summed in FP64
constexpr unsigned long long kForcedMoved = 10000000000ull;     // 1e10

template <typename Real>
__global__ __launch_bounds__(256) void k_end(const RepConst* __restrict__ RC,
                                             RepDyn* __restrict__ RD,
                                             const CellAcc* __restrict__ acc,
                                             const Real* __restrict__ ux, const Real* __restrict__ uy,
                                             const Real* __restrict__ uz, EndPartial* part) {
    // the 21 sums travel as doubles: the counters are integers far below 2^53, exact in FP64,
    // so one fixed-order DPP reduction (result in lane 63) serves them all
    __shared__ double sv[4][kEndVals];
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    const int ncell = D.ncell;
    const int chunk = (ncell + kEndParts - 1) / kEndParts;
    const int c0 = blockIdx.x * chunk, c1 = min(ncell, c0 + chunk);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    long long v[kEndCnt] = {};
    double d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = c0 + tid; c < c1; c += 256) {
        const CellAcc& A = acc[C.cell_off + c];
        v[0] += A.att; v[1] += A.acc; v[2] += A.left; v[3] += A.hard;
        v[4] += A.cand; v[5] += A.pairs; v[6] += A.coll; v[7] += A.dcap; v[8] += A.dccap;
        v[9] += A.s1rej; v[10] += A.s2rej;
        v[11] += A.ticks; v[12] += A.tmoves;
        d[0] += A.dfin;
        d[1] += A.moved_abs;
    }
    constexpr int kStride = kEndParts * 256;
    for (int i0 = blockIdx.x * 256 + tid; i0 < C.N; i0 += 4 * kStride) {
        double x[4], y[4], z[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kStride;
            const bool in = i < C.N;
            x[u] = in ? (double)ux[C.part_off + i] : 0.0;
            y[u] = in ? (double)uy[C.part_off + i] : 0.0;
            z[u] = in ? (double)uz[C.part_off + i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            d[2] += x[u] * x[u]; d[3] += y[u] * y[u]; d[4] += z[u] * z[u];
            d[5] += x[u] * y[u]; d[6] += x[u] * z[u]; d[7] += y[u] * z[u];
        }
    }
    double e[kEndVals];
#pragma unroll
    for (int k = 0; k < kEndCnt; ++k) e[k] = wave_sum_d63((double)v[k]);
#pragma unroll
    for (int k = 0; k < 8; ++k) e[kEndCnt + k] = wave_sum_d63(d[k]);
    if (lane == 63)
        for (int k = 0; k < kEndVals; ++k) sv[w][k] = e[k];
    __syncthreads();
    if (w != 0) return;                                  // wavefront 0 writes the partial
    // one sum per lane, the wavefronts in order
    EndPartial& P = part[(size_t)r * kEndParts + blockIdx.x];
    if (lane < kEndCnt)
        P.v[lane] = (long long)sv[0][lane] + (long long)sv[1][lane] + (long long)sv[2][lane] +
                    (long long)sv[3][lane];
    else if (lane < kEndVals)
        P.d[lane - kEndCnt] = ((sv[0][lane] + sv[1][lane]) + sv[2][lane]) + sv[3][lane];
    if (lane == 0) {
        const int k = kEndCnt + 1;
        const double mv = ((sv[0][k] + sv[1][k]) + sv[2][k]) + sv[3][k];
        // clamped (near-cap terms reach 1e36; anything above the test's 1e10 is enough)
        const unsigned long long q = (unsigned long long)llrint(fmin(mv, 1.0e15));
        if (q) atomicAdd(&D.moved_q[blockIdx.x & 7], q);
    }
}

// The fold of k_end's partials (wavefront 0 of k_energy_reduce's block for replica r): partials
// lane, lane + 64, ... in order per lane, then one DPP reduction per sum; lane 63 updates the
// running energy and counters, runs the step-size controller, flags an FP64 re-summation when
// the sweep moved more than kForcedMoved, keeps the orientation sums for the sweep's record and
// sets the pace of the next sweep's sequential passes (pace_pct percent of this sweep's mean
// cell time per move; 0: no pacing).
__device__ inline void end_fold(const RepConst& C, RepDyn& D, const EndPartial* __restrict__ part,
                                int pace_pct) {
    static_assert(kEndParts % 64 == 0, "whole partials per lane");
    const int lane = threadIdx.x & 63;
    long long v[kEndCnt] = {};
    double d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = lane; j < kEndParts; j += 64) {
        const EndPartial& P = part[j];
        for (int k = 0; k < kEndCnt; ++k) v[k] += P.v[k];
        for (int k = 0; k < 8; ++k) d[k] += P.d[k];
    }
    double e[kEndVals];
#pragma unroll
    for (int k = 0; k < kEndCnt; ++k) e[k] = wave_sum_d63((double)v[k]);
#pragma unroll
    for (int k = 0; k < 8; ++k) e[kEndCnt + k] = wave_sum_d63(d[k]);
    if (lane != 63) return;
    for (int k = 0; k < kEndCnt; ++k) v[k] = (long long)e[k];
    D.attempted += v[0]; D.accepted += v[1]; D.left += v[2]; D.hard += v[3];
    D.cand += v[4]; D.pairs += v[5]; D.coll += v[6];
    D.cap_pairs += v[7];
    D.cap_coll += v[8];
    D.da_s1 += v[9];
    D.da_s2 += v[10];
    D.e_fin += e[kEndCnt];
    D.moved_abs = e[kEndCnt + 1];
    if (moved_total(D) > kForcedMoved) D.resum = 1;
    D.pace = (pace_pct > 0 && v[12] > 0) ? (int)(v[11] * pace_pct / (100 * v[12])) : 0;
    if (C.adapt_moves > 0) {
        D.ad_att += (unsigned long long)v[0];
        D.ad_acc += (unsigned long long)v[1];
        if (D.ad_att >= (unsigned long long)C.adapt_moves) {
            const double half_box = 0.5 * fmin(D.L[0], fmin(D.L[1], D.L[2]));
            adapt_steps(D.dmax, D.drot, D.ad_acc, D.ad_att, C.min_thick, half_box);
            D.ad_att = D.ad_acc = 0;
            D.adapts += 1;
        }
    }
    for (int k = 0; k < 6; ++k) D.qsum[k] = e[kEndCnt + 2 + k];
}

// ------------------------------------------------------------------ FP64 full energy

// FP64 energy over a cell list (each unordered pair once: the partner must come later in the
// list), colloid term and colloid-core overlap count; one output record per wavefront of 64
// consecutive list positions.  mode: 0 always; 1 only when the replica flagged a
// re-summation; 2 the NPT trial box (indirect list, zero shift).  tot_ene,
// MCnptColloid.c:1598-1622.
//
// Each lane owns one particle and walks its (2R+1)^3 neighbour cells; in-cutoff partners are
// compacted (ballot) into a wavefront-private ring in LDS and evaluated 64 at a time with every
// lane busy, so the FP64 pair body never runs under divergence.  Every lane's running sums and
// the final shuffle tree are in a fixed order: the result is deterministic.
template <typename Real>
__global__ __launch_bounds__(128) void k_energy(
    const RepConst* __restrict__ RC, const RepDyn* __restrict__ RD, int mode,
    const uint32_t* __restrict__ qx, const uint32_t* __restrict__ qy, const uint32_t* __restrict__ qz,
    const Real* __restrict__ ux, const Real* __restrict__ uy, const Real* __restrict__ uz,
    const int* __restrict__ cell_start, const int* __restrict__ list,
    double* __restrict__ efin, int* __restrict__ ecap, int* __restrict__ eccap, int* __restrict__ ehard) {
    constexpr int QN = 128;                       // ring capacity (>= 63 pending + 64 appended)
    __shared__ int s_j[2][QN];
    __shared__ unsigned char s_l[2][QN];
    __shared__ uint32_t s_q[2][3][64];
    __shared__ double s_u[2][3][64];
    const int r = blockIdx.y;
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    if (mode == 1 && !D.resum && moved_total(D) <= kForcedMoved) return;   // k_end's moved energy
    if (mode == 2 && !D.trial_ok) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = (C.N + 63) >> 6;
    const int wi = blockIdx.x * 2 + w;
    if (wi >= nw) return;                         // whole wavefront (no workgroup barriers below)
    const int i = wi * 64 + lane;
    const bool act = i < C.N;
    const bool ind = mode == 2;
    const int n0 = ind ? D.nt[0] : D.n[0], n1 = ind ? D.nt[1] : D.n[1], n2 = ind ? D.nt[2] : D.n[2];
    const double L0 = ind ? D.Lt[0] : D.L[0], L1 = ind ? D.Lt[1] : D.L[1], L2 = ind ? D.Lt[2] : D.L[2];
    const double s0 = L0 / kTwo32, s1 = L1 / kTwo32, s2 = L2 / kTwo32;
    const uint32_t g0 = ind ? 0u : D.g[0], g1 = ind ? 0u : D.g[1], g2 = ind ? 0u : D.g[2];
    const int po = C.part_off;
    const double rc2 = C.rc2;
    uint32_t xi = 0u, yi = 0u, zi = 0u;
    double uix = 0.0, uiy = 0.0, uiz = 1.0;
    int cx = 0, cy = 0, cz = 0;
    if (act) {
        const int gi = po + (ind ? list[po + i] : i);
        xi = qx[gi]; yi = qy[gi]; zi = qz[gi];
        uix = ux[gi]; uiy = uy[gi]; uiz = uz[gi];
        cx = (int)cell_coord(xi, g0, n0); cy = (int)cell_coord(yi, g1, n1);
        cz = (int)cell_coord(zi, g2, n2);
    }
    s_q[w][0][lane] = xi; s_q[w][1][lane] = yi; s_q[w][2][lane] = zi;
    s_u[w][0][lane] = uix; s_u[w][1][lane] = uiy; s_u[w][2][lane] = uiz;
    const int* cs = cell_start + C.cs_off;
    const GBd gd = make_gbd(C);
    double f = 0.0;
    int cp = 0;
    int head = 0, tail = 0;                       // wavefront-uniform ring cursors

    // evaluate the n oldest queued pairs (n <= 64), pair (owner lane l, partner gj) on lane k
    auto flush = [&](int n) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < n) {
            const int slot = (head + lane) & (QN - 1);
            const int gj = s_j[w][slot];
            const int l = s_l[w][slot];
            const double ddx = (double)(int32_t)(qx[gj] - s_q[w][0][l]) * s0;
            const double ddy = (double)(int32_t)(qy[gj] - s_q[w][1][l]) * s1;
            const double ddz = (double)(int32_t)(qz[gj] - s_q[w][2][l]) * s2;
            const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            bool c;
            const double u = gb_pair_f64(gd, s_u[w][0][l], s_u[w][1][l], s_u[w][2][l], (double)ux[gj],
                                         (double)uy[gj], (double)uz[gj], ddx, ddy, ddz, r2, &c);
            if (c) ++cp; else f += u;
        }
        head += n;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    // stencil reach: +-1 cell on grids of cells >= r_c (8 colours, the NPT trial grid), +-2 on
    // the 27-colour grid of cells >= r_c/2
    const int R = ind ? 1 : C.col - 1;
    const int S1 = 2 * R + 1, NS = S1 * S1 * S1;
    for (int c = 0; c < NS; ++c) {
        int st = 0, len = 0;
        if (act) {
            int ex = cx + (c % S1) - R, ey = cy + ((c / S1) % S1) - R, ez = cz + c / (S1 * S1) - R;
            ex += ex < 0 ? n0 : 0; ex -= ex >= n0 ? n0 : 0;
            ey += ey < 0 ? n1 : 0; ey -= ey >= n1 ? n1 : 0;
            ez += ez < 0 ? n2 : 0; ez -= ez >= n2 ? n2 : 0;
            const int nb = ex + n0 * (ey + n1 * ez);
            const int b0 = cs[nb], b1 = cs[nb + 1];
            st = b0 > i + 1 ? b0 : i + 1;
            len = b1 - st;
            len = len > 0 ? len : 0;
        }
        int mx = len;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const int o = __shfl_xor(mx, off);
            mx = o > mx ? o : mx;
        }
        mx = bcast(mx);
        for (int t = 0; t < mx; ++t) {
            bool keep = false;
            int gj = 0;
            if (t < len) {
                const int kk = st + t;
                gj = po + (ind ? list[po + kk] : kk);
                const double ddx = (double)(int32_t)(qx[gj] - xi) * s0;
                const double ddy = (double)(int32_t)(qy[gj] - yi) * s1;
                const double ddz = (double)(int32_t)(qz[gj] - zi) * s2;
                keep = ddx * ddx + ddy * ddy + ddz * ddz < rc2;
            }
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const int slot = (tail + __popcll(m & __lanemask_lt())) & (QN - 1);
                s_j[w][slot] = gj;
                s_l[w][slot] = (unsigned char)lane;
            }
            tail += __popcll(m);
            if (tail - head >= 64) flush(64);
        }
    }
    while (tail > head) flush(tail - head < 64 ? tail - head : 64);

    int ccp = 0, hard = 0;
    if (C.colloid_on && act) {
        const double rx = (double)(int32_t)(xi - 0x80000000u) * s0;
        const double ry = (double)(int32_t)(yi - 0x80000000u) * s1;
        const double rz = (double)(int32_t)(zi - 0x80000000u) * s2;
        bool c;
        const double u = colloid_f64(C, rx, ry, rz, uix, uiy, uiz, &c);
        if (c) ccp = 1; else f += u;
        hard = (rx * rx + ry * ry + rz * rz <= C.R2) ? 1 : 0;
    }
    f = wave_sum(f);
    cp = wave_sum(cp);
    ccp = wave_sum(ccp);
    hard = wave_sum(hard);
    if (lane == 0) {
        efin[po + wi] = f;
        ecap[po + wi] = cp;
        eccap[po + wi] = ccp;
        ehard[po + wi] = hard;
    }
}

// Deterministic per-replica reduction of k_energy's output into D.e_res, then by mode:
// 0 replace the running energy; 1 the same when a re-summation was flagged (records the drift);
// 2 the NPT ln V acceptance test (absent from the reference, whose NPT is a label only,
// MCnptColloid.c:1197, :1291); 3 only D.e_res.
__device__ inline void energy_reduce(const RepConst& C, RepDyn& D, int mode,
                                     const double* __restrict__ efin, const int* __restrict__ ecap,
                                     const int* __restrict__ eccap, const int* __restrict__ ehard) {
    __shared__ double sf[256];
    __shared__ long long sc[256][3];
    const int tid = threadIdx.x;
    double f = 0.0;
    long long a = 0, b = 0, h = 0;
    const int nw = (C.N + 63) >> 6;                    // one record per k_energy wavefront
    for (int i = tid; i < nw; i += 256) {
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
            // every re-summation records its drift: periodic ones (and those the requested
            // interval and a large moved energy both ask for) in max_drift, the ones forced by
            // a sweep whose accepted moves moved > 1e10 of finite energy in max_drift_forced
            D.resums += 1;
            const double run = (double)D.cap_pairs * 1e9 + (double)D.cap_coll * 1e8 + D.e_fin;
            const double now = sc[0][0] * 1e9 + sc[0][1] * 1e8 + sf[0];
            const double dr = fabs(run - now);
            if (moved_total(D) <= kForcedMoved) {
                if (dr > D.max_drift) D.max_drift = dr;
            } else {
                D.resums_forced += 1;
                if (dr > D.max_drift_forced) D.max_drift_forced = dr;
            }
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

__device__ inline void series_record(const RepConst& C, RepDyn& D, int r, double* __restrict__ series,
                                     int series_cap, int advance);

// One block per replica; endp != null (the sweep's mode-1 call): first the fold of k_end's
// partials (end_fold); series != null: then the sweep's record (series_record, advancing the
// sweep index), after the energy.
__global__ __launch_bounds__(256) void k_energy_reduce(const RepConst* __restrict__ RC,
                                                       RepDyn* __restrict__ RD, int mode,
                                                       const double* __restrict__ efin,
                                                       const int* __restrict__ ecap,
                                                       const int* __restrict__ eccap,
                                                       const int* __restrict__ ehard,
                                                       double* __restrict__ series = nullptr,
                                                       int series_cap = 0,
                                                       const EndPartial* __restrict__ endp = nullptr,
                                                       int pace_pct = 0) {
    const int r = blockIdx.x;
    const RepConst& C = RC[r];
    RepDyn& D = RD[r];
    if (endp) {
        if (threadIdx.x < 64) end_fold(C, D, endp + (size_t)r * kEndParts, pace_pct);
        __syncthreads();                                // the fold's updates of D before the energy
    }
    const bool run = !(mode == 1 && !D.resum) && !(mode == 2 && !D.trial_ok);   // uniform
    if (run) energy_reduce(C, D, mode, efin, ecap, eccap, ehard);
    if (series && threadIdx.x == 0) series_record(C, D, r, series, series_cap, 1);
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
    int nn[3];
    for (int d = 0; d < 3; ++d) {
        D.Lt[d] = D.L[d] * f;
        nn[d] = cell_dim(D.Lt[d], C.rc);
        if (nn[d] == 0) ok = 0;
    }
    if (ok) fit_grid(nn, C.ncap[0] * C.ncap[1] * C.ncap[2], 2);
    for (int d = 0; d < 3; ++d) D.nt[d] = nn[d];
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

// Appends the sweep's observables (running energy, the orientation sums of k_end, box) of
// replica r to the series ring and advances the sweep index.
__device__ inline void series_record(const RepConst& C, RepDyn& D, int r, double* __restrict__ series,
                                     int series_cap, int advance) {
    if (series_cap > 0) {
        double* e = series + ((size_t)r * series_cap + (size_t)(D.series_count % series_cap)) * kSeriesWidth;
        const double E = (double)D.cap_pairs * 1e9 + (double)D.cap_coll * 1e8 + D.e_fin;
        e[0] = E / C.N;
        for (int k = 0; k < 6; ++k) e[1 + k] = D.qsum[k] / C.N;
        const double V = D.L[0] * D.L[1] * D.L[2];
        e[7] = C.N / (V - C.vcoll);
        e[8] = D.L[0]; e[9] = D.L[1]; e[10] = D.L[2];
        e[11] = (double)D.t;
        D.series_count += 1;
    }
    if (advance) D.t += 1;
}

__global__ void k_record(const RepConst* __restrict__ RC, RepDyn* __restrict__ RD, int nrep,
                         double* __restrict__ series, int series_cap, int advance) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    series_record(RC[r], RD[r], r, series, series_cap, advance);
}

// ------------------------------------------------------------------ snapshots

// Configuration of replica r in particle order, FP64, into a staging record
// {x[N], y[N], z[N], ux[N], uy[N], uz[N], L[3]} (the host conversion of mcgb_get_state, done
// where the data is: x = q L / 2^32, the same product as the host's, so the bits agree).
template <typename Real>
__global__ __launch_bounds__(256) void k_snapshot(const RepConst* __restrict__ RC,
                                                  const RepDyn* __restrict__ RD, int r,
                                                  const uint32_t* __restrict__ qx,
                                                  const uint32_t* __restrict__ qy,
                                                  const uint32_t* __restrict__ qz,
                                                  const Real* __restrict__ ux,
                                                  const Real* __restrict__ uy,
                                                  const Real* __restrict__ uz,
                                                  const int* __restrict__ pid,
                                                  double* __restrict__ out) {
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    const int N = C.N;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3) out[6 * (size_t)N + i] = D.L[i];
    if (i >= N) return;
    const int s = C.part_off + i;
    const int p = pid[s];
    if ((unsigned)p >= (unsigned)N) return;
    out[p] = (double)qx[s] * (D.L[0] / kTwo32);
    out[N + p] = (double)qy[s] * (D.L[1] / kTwo32);
    out[2 * (size_t)N + p] = (double)qz[s] * (D.L[2] / kTwo32);
    out[3 * (size_t)N + p] = (double)ux[s];
    out[4 * (size_t)N + p] = (double)uy[s];
    out[5 * (size_t)N + p] = (double)uz[s];
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
        v = gb_pair_f64(make_gbd(C), ei[3 * k], ei[3 * k + 1], ei[3 * k + 2], ej[3 * k], ej[3 * k + 1],
                        ej[3 * k + 2], dx, dy, dz, dx * dx + dy * dy + dz * dz, &cp);
    } else {
        // the FP32 kernel takes unit vectors (the sweep stores normalised orientations)
        double ni = 1.0 / sqrt(ei[3 * k] * ei[3 * k] + ei[3 * k + 1] * ei[3 * k + 1] + ei[3 * k + 2] * ei[3 * k + 2]);
        double nj = 1.0 / sqrt(ej[3 * k] * ej[3 * k] + ej[3 * k + 1] * ej[3 * k + 1] + ej[3 * k + 2] * ej[3 * k + 2]);
        const double ddx = dr[3 * k], ddy = dr[3 * k + 1], ddz = dr[3 * k + 2];
        v = (double)gb_pair_f32(make_gbf(C), (float)(ei[3 * k] * ni), (float)(ei[3 * k + 1] * ni),
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
