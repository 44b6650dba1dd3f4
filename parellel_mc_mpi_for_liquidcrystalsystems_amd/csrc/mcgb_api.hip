// mcgb_api.hip — the C ABI of libmcgb (include/mcgb.h): context, device memory, the launch
// sequence of a sweep, parity hooks.  Kernels live in mcgb_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "mcgb_internal.h"
#include "mcgb_kernels.hpp"
#include "mcgb_duo.hpp"

using namespace mcgb;

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            mcgb_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return MCGB_ERR_HIP;                                                               \
        }                                                                                      \
    } while (0)

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kSeriesCapDefault = 1 << 16;
// Sequential-pass pacing (k_pass): a cell that falls behind kPaceDefault percent of the last
// sweep's mean time per move (its time counted from the start of its workgroup) raises its
// wavefronts' issue priority.  At N = 256000 (1331 two-wavefront cells per pass, up to three
// wavefronts per SIMD): 103% +6%, 109% +9%, 115% +8% over no pacing; with at most two
// wavefronts per SIMD it changes nothing.
constexpr int kPaceDefault = 109;
constexpr int kEvSet = 27 + 3;     // events of one timed sweep: start, rebuild, passes, end

struct Events {
    std::vector<hipEvent_t> ev;
    int get(int i, hipEvent_t* out) {
        while ((int)ev.size() <= i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            ev.push_back(e);
        }
        *out = ev[i];
        return 0;
    }
    ~Events() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

}  // namespace

struct mcgb_ctx {
    int nrep = 0, prec = 32;
    size_t rsize = 4;
    std::vector<mcgb_params> P;
    std::vector<RepConst> hC;
    std::vector<int> blk_rep;
    std::vector<int> state_set;
    int nblk = 0, Ntot = 0, Nmax = 0, ncell_tot = 0, ncs_tot = 0, ncell_max = 0;
    int cap = 0;
    int cap_request = 0;
    int waves = 4;
    size_t lds = 0;
    size_t lds_spec = 0;
    size_t lds_duo = 0;
    bool spec = true;          // speculative pass schedule for untraced plain sweeps
    bool o4 = false;           // FP32 sequential pass built for four wavefronts per SIMD
    bool duo = false;          // two moves in flight per cell (two wavefronts), traced or not
    int resum_interval = 100;
    int pace_pct = kPaceDefault;   // sequential-pass pacing, percent of the mean move time (0: off)
    int series_cap = kSeriesCapDefault;
    bool any_npt = false;
    hipStream_t st = nullptr;
    int device = 0;
    RepConst* dC = nullptr;
    RepDyn* dD = nullptr;
    int* dBlk = nullptr;
    uint32_t* q[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    void* u[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int* pid[2] = {nullptr, nullptr};
    int cur = 0;
    int *cell_id = nullptr, *tmp = nullptr, *count = nullptr, *cell_start = nullptr;
    int *cell_start2 = nullptr, *list2 = nullptr;
    unsigned long long* scan_st = nullptr;   // [nrep][scan_tiles] look-back status of the scan
    int* slot = nullptr;                 // each particle's place within its cell (rebuild)
    int* tmp_pid = nullptr;              // particle ids scattered beside tmp (rebuild)
    EndPartial* endp = nullptr;          // [nrep][kEndParts] partial sweep totals
    int scan_tiles = 1;
    CellAcc* acc = nullptr;
    MoveDraw* draws = nullptr;
    int rec_max = 0;
    int* ovf_count = nullptr;
    int* ovf_list = nullptr;
    double* efin = nullptr;
    int *ecap = nullptr, *eccap = nullptr, *ehard = nullptr;
    double* obs = nullptr;
    double* series = nullptr;
    mcgb_move_trace* dtrace = nullptr;
    size_t dtrace_cap = 0;
    std::vector<void*> allocs;
    Events events;
    int col = 2, npass = 8;    // sublattice colours per axis and passes per sweep
    hipEvent_t tev[2 * kEvSet] = {};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // learned first stage + delayed acceptance
    float* sur_P = nullptr;            // FP32 master parameters (kNParam)
    MlpPacked* sur_M = nullptr;        // packed BF16 matrix-core operands
    bool sur_set = false;
    int da_on = 0, da_judge = 0;
    // automatic fallback from delayed acceptance to the plain kernel (mcgb_da_guard)
    int da_guard_window = 0;                 // sweeps per check, 0 = off
    double da_guard_s2_max = 0.5, da_guard_a1_min = 0.02;
    unsigned long long da_fallbacks = 0;
    double da_rates[2] = {0.0, 0.0};         // last window: stage-2 rejection, stage-1 acceptance
    unsigned long long* da_hist = nullptr;   // [nrep][kDaBins + 2]
    float* pairs = nullptr;            // emitted exact pair records (kPairRec floats each)
    long long pairs_cap = 0, pairs_n = 0;
    float *tr_grad = nullptr, *tr_m = nullptr, *tr_v = nullptr, *tr_partial = nullptr;
    double *tr_lpart = nullptr, *tr_loss = nullptr;
    int tr_nblk = 0, adam_t = 0;
    // asynchronous snapshots: one staging record and one pinned host record per replica
    double* snap_dev = nullptr;                // [6 Ntot + 3 nrep]
    std::vector<double*> snap_host;            // [replica][slot] hipHostMalloc, 6 N + 3 doubles
    std::vector<hipEvent_t> snap_ev;           // [replica][slot]; slot 1 serves mcgb_get_state
    std::vector<int> snap_pending;
    bool use_graph = true;
    bool timing = true;
    double ms_pass = 0, ms_rebuild = 0, ms_sweep = 0;
};

namespace {

int dalloc(mcgb_ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        mcgb_set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return MCGB_ERR_NOMEM;
    }
    c->allocs.push_back(*p);
    return MCGB_OK;
}

template <typename T>
int dalloc_t(mcgb_ctx* c, T** p, size_t n) {
    return dalloc(c, (void**)p, n * sizeof(T));
}

double box_length(const mcgb_params& p) {
    double R = p.colloid_wall_option == 1 ? p.colloid_radius : 0.0;
    double V = p.num_particles / p.system_density + (4.0 / 3.0) * kPi * R * R * R;
    return cbrt(V);
}

// Constants of one replica (get_system_const, MCnptColloid.c:2752-2783, dependency order).
void fill_consts(const mcgb_params& p, RepConst* C) {
    double dv[11];
    mcgb_params_derived(&p, dv);
    C->sigma0 = p.gb_sigma_sidetoside;
    C->four_eps0 = dv[1];
    C->chi = dv[4];
    C->chi2 = dv[5];
    C->chip = dv[7];
    C->chip2 = dv[7] * dv[7];
    C->mu = p.gb_mu_exponent;
    C->nu = p.gb_nu_exponent;
    C->rc = p.cutoff_radius;
    C->rc2 = dv[2];
    C->temperature = p.system_temperature;
    C->beta = dv[0];
    C->pressure = p.system_pressure;
    C->colloid_on = p.colloid_wall_option == 1 ? 1 : 0;
    C->R = p.colloid_radius;
    C->R2 = dv[10];
    C->colloid_cut = dv[9];
    C->kappa_prime = p.gb_kappa_prime;
    C->one_minus_chi = 1.0 - dv[4];
    C->one_minus_chip = 1.0 - dv[7];
    C->npt_max_dlnv = p.npt_max_dlnv;
    C->adapt_moves = p.adapt_moves;
    C->col = p.sublattice_colors == 3 ? 3 : 2;
    C->min_thick = (float)(dv[8] < p.gb_sigma_sidetoside ? dv[8] : p.gb_sigma_sidetoside);
    C->npt = p.mc_ensemble_type == 2 ? 1 : 0;
    C->vcoll = C->colloid_on ? (4.0 / 3.0) * kPi * p.colloid_radius * p.colloid_radius * p.colloid_radius : 0.0;
    C->f_sigma0 = (float)C->sigma0;
    C->f_four_eps0 = (float)C->four_eps0;
    C->f_chi = (float)C->chi;
    C->f_chi2 = (float)C->chi2;
    C->f_chip = (float)C->chip;
    C->f_chip2 = (float)C->chip2;
    C->f_mu = (float)C->mu;
    C->f_half_nu = (float)(0.5 * C->nu);
    C->f_rc2 = (float)C->rc2;
    C->f_g1c0 = (float)((1.0 - C->chi) * (1.0 - C->chi));
    C->f_g1c1 = (float)(C->chi * (1.0 - C->chi));
    C->f_g1d0 = (float)(1.0 - C->chi2);
    C->f_g2c0 = (float)((1.0 - C->chip) * (1.0 - C->chip));
    C->f_g2c1 = (float)(C->chip * (1.0 - C->chip));
    C->f_g2d0 = (float)(1.0 - C->chip2);
    C->f_rc2_pre = (float)(C->rc2 * (1.0 + 1e-5));
    const bool imu = C->mu == 1.0 || C->mu == 2.0, inu = C->nu == 1.0 || C->nu == 2.0;
    C->f_ipow = imu && inu ? (int)C->mu + 2 * ((int)C->nu - 1) : 0;
}

int check_params(const mcgb_params& p, int r) {
    // non-finite inputs would reach every energy as NaN (a NaN dE is never accepted, but the
    // running total would be lost): refused here
    const double vals[] = {p.system_density, p.system_temperature, p.system_pressure,
                           p.gb_sigma_sidetoside, p.gb_epsilon0, p.gb_kappa, p.gb_kappa_prime,
                           p.gb_mu_exponent, p.gb_nu_exponent, (double)p.max_displacement,
                           (double)p.max_rotation, p.cutoff_radius, p.colloid_radius,
                           p.npt_max_dlnv};
    for (double v : vals)
        if (!std::isfinite(v)) {
            mcgb_set_error("replica %d: a parameter is not finite", r);
            return MCGB_ERR_ARG;
        }
    if (p.num_particles <= 0 || p.system_density <= 0.0 || p.system_temperature <= 0.0 ||
        p.cutoff_radius <= 0.0 || p.gb_sigma_sidetoside <= 0.0 || p.gb_mu_exponent == 0.0 ||
        p.gb_kappa <= 0.0 || p.gb_kappa_prime <= 0.0) {
        mcgb_set_error("replica %d: num_particles, density, temperature, cutoff, sigma0, kappa, "
                       "kappa' must be > 0 and mu != 0", r);
        return MCGB_ERR_ARG;
    }
    if (p.max_displacement < 0.0f || p.max_rotation < 0.0f) {
        mcgb_set_error("replica %d: negative move size", r);
        return MCGB_ERR_ARG;
    }
    if (p.adapt_moves < 0) {
        mcgb_set_error("replica %d: adapt_moves must be >= 0", r);
        return MCGB_ERR_ARG;
    }
    if (p.sublattice_colors != 2 && p.sublattice_colors != 3) {
        mcgb_set_error("replica %d: sublattice_colors must be 2 or 3 (got %d)", r, p.sublattice_colors);
        return MCGB_ERR_ARG;
    }
    return MCGB_OK;
}

// Exclusive scan of the cell counts (three launches over tiles of kScanTile cells).
void enqueue_scan(mcgb_ctx* c, int* cell_start, int trial) {
    k_scan<<<dim3(c->scan_tiles, c->nrep), 256, 0, c->st>>>(c->dC, c->dD, c->count, c->scan_st,
                                                            c->scan_tiles, cell_start, trial);
}

// Cell-list rebuild of the current buffer into the other one.  draws: the sweep's move draws
// are made in the counting launch (k_count_draws).
template <typename Real>
int enqueue_rebuild(mcgb_ctx* c, bool draws = false) {
    const int src = c->cur, dst = 1 - c->cur;
    const dim3 gp((c->Nmax + 255) / 256, c->nrep);
    // counts start at zero: set once at creation, reset by k_sort_gather after every rebuild
    if (draws) {
        const int nbc = (c->Nmax + 255) / 256, nbd = (c->rec_max + 255) / 256;
        k_count_draws<<<dim3(nbc + nbd, c->nrep), 256, 0, c->st>>>(
            c->dC, c->dD, c->draws, c->q[src][0], c->q[src][1], c->q[src][2], c->cell_id, c->slot,
            c->count, nbc);
    } else {
        k_count<<<gp, 256, 0, c->st>>>(c->dC, c->dD, c->q[src][0], c->q[src][1], c->q[src][2],
                                       c->cell_id, c->slot, c->count, 0);
    }
    enqueue_scan(c, c->cell_start, 0);
    k_scatter<<<gp, 256, 0, c->st>>>(c->dC, c->dD, c->cell_id, c->slot, c->cell_start, c->pid[src],
                                     c->tmp, c->tmp_pid, c->scan_st, c->scan_tiles, 0);
    const dim3 gs((c->ncell_max + 3) / 4, c->nrep);
    k_sort_gather<Real><<<gs, 256, 0, c->st>>>(
        c->dC, c->dD, c->cell_start, c->tmp, c->tmp_pid, 0, 0, nullptr, c->count, c->q[src][0], c->q[src][1],
        c->q[src][2], (const Real*)c->u[src][0], (const Real*)c->u[src][1], (const Real*)c->u[src][2],
        c->q[dst][0], c->q[dst][1], c->q[dst][2], (Real*)c->u[dst][0], (Real*)c->u[dst][1],
        (Real*)c->u[dst][2], c->pid[dst]);
    c->cur = dst;
    HIPCHK(hipGetLastError());
    return MCGB_OK;
}

// FP64 energy of the current cell list; mode as in k_energy_reduce.
template <typename Real>
int enqueue_energy(mcgb_ctx* c, int mode, bool record = false, bool fold = false) {
    const int b = c->cur;
    const dim3 ge((c->Nmax + 127) / 128, c->nrep);
    const int* cs = mode == 2 ? c->cell_start2 : c->cell_start;
    k_energy<Real><<<ge, 128, 0, c->st>>>(c->dC, c->dD, mode, c->q[b][0], c->q[b][1], c->q[b][2],
                                          (const Real*)c->u[b][0], (const Real*)c->u[b][1],
                                          (const Real*)c->u[b][2], cs, c->list2, c->efin, c->ecap,
                                          c->eccap, c->ehard);
    k_energy_reduce<<<c->nrep, 256, 0, c->st>>>(c->dC, c->dD, mode, c->efin, c->ecap, c->eccap, c->ehard,
                                                record ? c->series : nullptr, c->series_cap,
                                                fold ? c->endp : nullptr, c->pace_pct);
    HIPCHK(hipGetLastError());
    return MCGB_OK;
}

template <typename Real>
int enqueue_npt(mcgb_ctx* c) {
    const int b = c->cur;
    k_npt_trial<<<(c->nrep + 63) / 64, 64, 0, c->st>>>(c->dC, c->dD, c->nrep);
    const dim3 gp((c->Nmax + 255) / 256, c->nrep);
    k_count<<<gp, 256, 0, c->st>>>(c->dC, c->dD, c->q[b][0], c->q[b][1], c->q[b][2], c->cell_id,
                                   c->slot, c->count, 1);
    enqueue_scan(c, c->cell_start2, 1);
    k_scatter<<<gp, 256, 0, c->st>>>(c->dC, c->dD, c->cell_id, c->slot, c->cell_start2, c->pid[b],
                                     c->tmp, c->tmp_pid, c->scan_st, c->scan_tiles, 1);
    const dim3 gs((c->ncell_max + 3) / 4, c->nrep);
    k_sort_gather<Real><<<gs, 256, 0, c->st>>>(
        c->dC, c->dD, c->cell_start2, c->tmp, c->tmp_pid, 1, 1, c->list2, c->count, nullptr, nullptr, nullptr,
        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    HIPCHK(hipGetLastError());
    return enqueue_energy<Real>(c, 2);
}

// Orientation products of the current state, per block (mcgb_order_parameter).
template <typename Real>
int enqueue_observe(mcgb_ctx* c) {
    const int b = c->cur;
    k_observe<Real><<<dim3(64, c->nrep), 256, 0, c->st>>>(c->dC, (const Real*)c->u[b][0],
                                                          (const Real*)c->u[b][1],
                                                          (const Real*)c->u[b][2], c->obs);
    HIPCHK(hipGetLastError());
    return MCGB_OK;
}

template <typename Real, bool TR>
void launch_pass_t(mcgb_ctx* c, const PassArgs& P) {
    if (!TR && c->da_on) {
        switch (c->waves) {
        case 1: k_pass_da<Real, 1><<<c->nblk, 64, c->lds, c->st>>>(P); break;
        case 2: k_pass_da<Real, 2><<<c->nblk, 128, c->lds, c->st>>>(P); break;
        case 3: k_pass_da<Real, 3><<<c->nblk, 192, c->lds, c->st>>>(P); break;
        case 8: k_pass_da<Real, 8><<<c->nblk, 512, c->lds, c->st>>>(P); break;
        default: k_pass_da<Real, 4><<<c->nblk, 256, c->lds, c->st>>>(P); break;
        }
    } else if (c->duo && c->waves == 2) {
        k_pass_duo<Real, TR><<<c->nblk, 128, c->lds_duo, c->st>>>(P);
    } else if (!TR && c->spec) {
        switch (c->waves) {
        case 1: k_pass_spec<Real, 1><<<c->nblk, 64, c->lds_spec, c->st>>>(P); break;
        case 2: k_pass_spec<Real, 2><<<c->nblk, 128, c->lds_spec, c->st>>>(P); break;
        case 3: k_pass_spec<Real, 3><<<c->nblk, 192, c->lds_spec, c->st>>>(P); break;
        case 8: k_pass_spec<Real, 8><<<c->nblk, 512, c->lds_spec, c->st>>>(P); break;
        default: k_pass_spec<Real, 4><<<c->nblk, 256, c->lds_spec, c->st>>>(P); break;
        }
    } else if (!TR && c->o4 && std::is_same<Real, float>::value) {
        switch (c->waves) {
        case 1: k_pass_o4<1><<<c->nblk, 64, c->lds, c->st>>>(P); break;
        case 2: k_pass_o4<2><<<c->nblk, 128, c->lds, c->st>>>(P); break;
        case 3: k_pass_o4<3><<<c->nblk, 192, c->lds, c->st>>>(P); break;
        case 8: k_pass_o4<8><<<c->nblk, 512, c->lds, c->st>>>(P); break;
        default: k_pass_o4<4><<<c->nblk, 256, c->lds, c->st>>>(P); break;
        }
    } else {
    switch (c->waves) {
    case 1: k_pass<Real, 1, false, TR><<<c->nblk, 64, c->lds, c->st>>>(P); break;
    case 2: k_pass<Real, 2, false, TR><<<c->nblk, 128, c->lds, c->st>>>(P); break;
    case 3: k_pass<Real, 3, false, TR><<<c->nblk, 192, c->lds, c->st>>>(P); break;
    case 8: k_pass<Real, 8, false, TR><<<c->nblk, 512, c->lds, c->st>>>(P); break;
    default: k_pass<Real, 4, false, TR><<<c->nblk, 256, c->lds, c->st>>>(P); break;
    }
    }
    // cells whose stencil exceeded the staging capacity: single-wavefront workgroups walk the
    // overflow list and read the stencil from HBM (with an empty list every workgroup exits at
    // once); only the stencil table, draws and partials live in LDS
    PassArgs Q = P;
    Q.cap = 32;
    k_pass<Real, 1, true, TR><<<std::min(c->nblk, 256), 64, PassLds<Real, 1>::bytes(32), c->st>>>(Q);
}

template <typename Real>
void launch_pass(mcgb_ctx* c, const PassArgs& P) {
    if (P.trace) launch_pass_t<Real, true>(c, P);
    else launch_pass_t<Real, false>(c, P);
}

template <typename Real>
size_t pass_lds_seq(int waves, int cap) {
    switch (waves) {
    case 1: return PassLds<Real, 1>::bytes(cap);
    case 2: return PassLds<Real, 2>::bytes(cap);
    case 3: return PassLds<Real, 3>::bytes(cap);
    case 8: return PassLds<Real, 8>::bytes(cap);
    default: return PassLds<Real, 4>::bytes(cap);
    }
}
template <typename Real>
size_t pass_lds_spec(int waves, int cap) {
    switch (waves) {
    case 1: return SpecLds<Real, 1>::bytes(cap);
    case 2: return SpecLds<Real, 2>::bytes(cap);
    case 3: return SpecLds<Real, 3>::bytes(cap);
    case 8: return SpecLds<Real, 8>::bytes(cap);
    default: return SpecLds<Real, 4>::bytes(cap);
    }
}
// LDS of the largest of the schedules' staging layouts (all use the same capacity; the
// two-moves schedule runs with two wavefronts only)
template <typename Real>
size_t pass_lds(int waves, int cap) {
    size_t b = std::max(pass_lds_seq<Real>(waves, cap), pass_lds_spec<Real>(waves, cap));
    return waves == 2 ? std::max(b, DuoLds<Real>::bytes(cap)) : b;
}

// One sweep of every replica.  ev (may be null): npass + 3 events recorded at the sweep start,
// after the rebuild, after each of the npass passes and at the end.
template <typename Real>
int enqueue_one(mcgb_ctx* c, mcgb_move_trace* trace, const hipEvent_t* ev) {
    int rc;
    if (ev) HIPCHK(hipEventRecord(ev[0], c->st));
    k_begin<<<(c->nrep + 63) / 64, 64, 0, c->st>>>(c->dC, c->dD, c->nrep, c->resum_interval,
                                                  c->ovf_count);
    if ((rc = enqueue_rebuild<Real>(c, true))) return rc;
    if (ev) HIPCHK(hipEventRecord(ev[1], c->st));
    PassArgs P;
    P.RC = c->dC; P.RD = c->dD; P.blk_rep = c->dBlk;
    const int b = c->cur;
    P.qx = c->q[b][0]; P.qy = c->q[b][1]; P.qz = c->q[b][2];
    P.ux = c->u[b][0]; P.uy = c->u[b][1]; P.uz = c->u[b][2];
    P.pid = c->pid[b];
    P.cell_start = c->cell_start;
    P.acc = c->acc;
    P.trace = trace;
    P.draws = c->draws;
    P.ovf_count = c->ovf_count;
    P.ovf_list = c->ovf_list;
    P.nblk = c->nblk;
    P.cap = c->cap;
    P.mlp = c->sur_M;
    P.da_hist = c->da_hist;
    P.da_judge = c->da_judge;
    for (int p = 0; p < c->npass; ++p) {
        P.pass = p;
        launch_pass<Real>(c, P);
        if (ev) HIPCHK(hipEventRecord(ev[2 + p], c->st));
    }
    HIPCHK(hipGetLastError());
    k_end<Real><<<dim3(kEndParts, c->nrep), 256, 0, c->st>>>(
        c->dC, c->dD, c->acc, (const Real*)c->u[b][0], (const Real*)c->u[b][1],
        (const Real*)c->u[b][2], c->endp);
    // the re-summation (when flagged), the NPT volume move, the sweep's record: the record goes
    // with the energy reduction when no volume move follows
    if ((rc = enqueue_energy<Real>(c, 1, !c->any_npt, true))) return rc;
    if (c->any_npt) {
        if ((rc = enqueue_npt<Real>(c))) return rc;
        k_record<<<(c->nrep + 63) / 64, 64, 0, c->st>>>(c->dC, c->dD, c->nrep, c->series,
                                                        c->series_cap, 1);
    }
    if (ev) HIPCHK(hipEventRecord(ev[2 + c->npass], c->st));
    return MCGB_OK;
}

// The sweep sequence of two consecutive sweeps (the particle buffers swap once per sweep, so
// two sweeps return to the starting buffer) captured as a HIP graph for the current buffer.
template <typename Real>
int build_graph(mcgb_ctx* c) {
    const int par = c->cur;
    if (c->graph[par]) return MCGB_OK;
    hipGraph_t g;
    HIPCHK(hipStreamSynchronize(c->st));      // capture starts on an idle stream
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_one<Real>(c, nullptr, nullptr);
    if (!rc) rc = enqueue_one<Real>(c, nullptr, nullptr);
    hipError_t ec = hipStreamEndCapture(c->st, &g);
    if (rc) return rc;
    if (ec != hipSuccess) {
        mcgb_set_error("graph capture failed: %s", hipGetErrorString(ec));
        return MCGB_ERR_HIP;
    }
    hipError_t ei = hipGraphInstantiate(&c->graph[par], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c->graph[par] = nullptr;
        mcgb_set_error("graph instantiation failed: %s", hipGetErrorString(ei));
        return MCGB_ERR_HIP;
    }
    return MCGB_OK;
}

template <typename Real>
int enqueue_sweeps(mcgb_ctx* c, int n_sweeps, mcgb_move_trace* trace, bool async = false) {
    int rc;
    const bool tm = c->timing && !async;
    int timed_sets = 0;   // event sets (npass + 3 events) holding the last sweeps, in c->tev
    if (n_sweeps <= 0) return MCGB_OK;
    // graph replay for the bulk; with timing on, the last two sweeps are launched one kernel
    // at a time between HIP events (events inside a graph cannot be timed)
    const int tail = tm ? std::min(n_sweeps, 2) : 0;
    int bulk = n_sweeps - tail;
    if (trace || !c->use_graph || bulk < 2) {
        for (int s = 0; s < bulk; ++s)
            if ((rc = enqueue_one<Real>(c, trace, nullptr))) return rc;
    } else {
        if (bulk & 1) {
            if ((rc = enqueue_one<Real>(c, nullptr, nullptr))) return rc;
            --bulk;
        }
        if ((rc = build_graph<Real>(c))) return rc;
        const int par = c->cur;
        for (int i = 0; i < bulk / 2; ++i) HIPCHK(hipGraphLaunch(c->graph[par], c->st));
    }
    for (int s = 0; s < tail; ++s)
        if ((rc = enqueue_one<Real>(c, trace, c->tev + kEvSet * s))) return rc;
    timed_sets = tail;
    if (async) return MCGB_OK;
    HIPCHK(hipStreamSynchronize(c->st));
    if (timed_sets > 0) {
        double tp = 0, tr = 0, tsw = 0;
        for (int k = 0; k < timed_sets; ++k) {
            const hipEvent_t* e = c->tev + kEvSet * k;
            float ms;
            HIPCHK(hipEventElapsedTime(&ms, e[0], e[1]));
            tr += ms;
            for (int p = 0; p < c->npass; ++p) {
                HIPCHK(hipEventElapsedTime(&ms, e[1 + p], e[2 + p]));
                tp += ms;
            }
            HIPCHK(hipEventElapsedTime(&ms, e[0], e[2 + c->npass]));
            tsw += ms;
        }
        c->ms_rebuild = tr / timed_sets;
        c->ms_pass = tp / ((double)c->npass * timed_sets);
        c->ms_sweep = tsw / timed_sets;
    }
    return MCGB_OK;
}

int read_dyn(mcgb_ctx* c, std::vector<RepDyn>& D) {
    D.resize(c->nrep);
    HIPCHK(hipMemcpyAsync(D.data(), c->dD, sizeof(RepDyn) * c->nrep, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return MCGB_OK;
}

int write_dyn(mcgb_ctx* c, const std::vector<RepDyn>& D) {
    HIPCHK(hipMemcpyAsync(c->dD, D.data(), sizeof(RepDyn) * c->nrep, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return MCGB_OK;
}

int check_ctx(mcgb_ctx* c, int replica) {
    if (!c) { mcgb_set_error("null context"); return MCGB_ERR_ARG; }
    if (replica < 0 || replica >= c->nrep) {
        mcgb_set_error("replica %d out of range [0, %d)", replica, c->nrep);
        return MCGB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(c->device));
    return MCGB_OK;
}

// Cell list with zero shift for the current boxes, then the FP64 energy (mode 0 or 3).
int recompute(mcgb_ctx* c, int mode) {
    k_grid_g0<<<(c->nrep + 63) / 64, 64, 0, c->st>>>(c->dC, c->dD, c->nrep);
    int rc = c->prec == 64 ? enqueue_rebuild<double>(c) : enqueue_rebuild<float>(c);
    if (rc) return rc;
    rc = c->prec == 64 ? enqueue_energy<double>(c, mode) : enqueue_energy<float>(c, mode);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return MCGB_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

int mcgb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void mcgb_destroy(mcgb_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (int k = 0; k < 2; ++k)
        if (c->graph[k]) (void)hipGraphExecDestroy(c->graph[k]);
    for (int k = 0; k < 2 * kEvSet; ++k)
        if (c->tev[k]) (void)hipEventDestroy(c->tev[k]);
    for (double* h : c->snap_host)
        if (h) (void)hipHostFree(h);
    for (hipEvent_t e : c->snap_ev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int mcgb_create(const mcgb_params* p, int n_replicas, mcgb_ctx** out) {
    if (!p || !out || n_replicas <= 0 || n_replicas > kMaxReplicas) {
        mcgb_set_error("mcgb_create: need 1..%d parameter blocks and an output pointer", kMaxReplicas);
        return MCGB_ERR_ARG;
    }
    *out = nullptr;
    const int prec = p[0].precision;
    if (prec != 32 && prec != 64) {
        mcgb_set_error("precision must be 32 or 64 (got %d)", prec);
        return MCGB_ERR_ARG;
    }
    for (int r = 0; r < n_replicas; ++r) {
        int rc = check_params(p[r], r);
        if (rc) return rc;
        if (p[r].precision != prec) {
            mcgb_set_error("replica %d: all replicas must share one precision", r);
            return MCGB_ERR_ARG;
        }
    }
    int ndev = 0;
    hipError_t he = hipGetDeviceCount(&ndev);
    if (he != hipSuccess || ndev == 0) {
        mcgb_set_error("no HIP device available (%s)", hipGetErrorString(he));
        return MCGB_ERR_HIP;
    }
    mcgb_ctx* c = new mcgb_ctx();
    c->nrep = n_replicas;
    c->prec = prec;
    c->rsize = prec == 64 ? 8 : 4;
    c->P.assign(p, p + n_replicas);
    c->resum_interval = p[0].energy_check_interval;
    HIPCHK(hipGetDevice(&c->device));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    for (int k = 0; k < 2 * kEvSet; ++k) HIPCHK(hipEventCreate(&c->tev[k]));
    // MCGB_NO_GRAPH=1: launch every sweep kernel by kernel (for profilers and debuggers that
    // do not follow graph launches); results are identical
    {
        const char* ng = getenv("MCGB_NO_GRAPH");
        if (ng && ng[0] && strcmp(ng, "0") != 0) c->use_graph = false;
    }

    // ---- layout of every replica
    c->hC.resize(n_replicas);
    c->state_set.assign(n_replicas, 0);
    std::vector<RepDyn> D(n_replicas);
    double occ_max = 0.0;
    for (int r = 0; r < n_replicas; ++r) {
        RepConst& C = c->hC[r];
        memset(&C, 0, sizeof(C));
        fill_consts(p[r], &C);
        const mcgb_params& pr = p[r];
        double L = box_length(pr);
        double Lcap = C.npt ? L * 1.3 : L;
        int ncell = 1;
        for (int d = 0; d < 3; ++d) {
            int n = grid_dim(Lcap, pr.cutoff_radius, C.col);
            if (n == 0 || mcgb_cell_dim(Lcap, pr.cutoff_radius) == 0) {
                mcgb_set_error("replica %d: box %.4f is below 4 cells of the cutoff %.4f (8-sublattice "
                               "grid impossible)", r, L, pr.cutoff_radius);
                mcgb_destroy(c);
                return MCGB_ERR_GEOMETRY;
            }
            C.ncap[d] = n;
            ncell *= n;
        }
        if ((double)pr.max_displacement >= 0.5 * L) {
            mcgb_set_error("replica %d: max_displacement must be below half the box", r);
            mcgb_destroy(c);
            return MCGB_ERR_GEOMETRY;
        }
        C.N = pr.num_particles;
        C.part_off = c->Ntot;
        C.cell_off = c->ncell_tot;
        C.cs_off = c->ncs_tot;
        C.act_off = c->nblk;
        C.nact_cap = ncell / (C.col * C.col * C.col);
        C.key0 = (uint32_t)pr.seed;
        C.key1 = (uint32_t)(pr.seed >> 32) + (uint32_t)r;
        C.rec_off = c->Ntot + c->ncell_tot;
        c->rec_max = std::max(c->rec_max, pr.num_particles + ncell);
        c->Ntot += C.N;
        c->Nmax = std::max(c->Nmax, C.N);
        c->ncell_tot += ncell;
        c->ncs_tot += ncell + 1;
        c->ncell_max = std::max(c->ncell_max, ncell);
        for (int b = 0; b < C.nact_cap; ++b) c->blk_rep.push_back(r);
        c->nblk += C.nact_cap;
        c->any_npt = c->any_npt || C.npt;
        int n0 = grid_dim(L, pr.cutoff_radius, C.col);
        double occ = (double)C.N / ((double)n0 * n0 * n0);
        occ_max = std::max(occ_max, occ);
        RepDyn& Dr = D[r];
        memset(&Dr, 0, sizeof(Dr));
        for (int d = 0; d < 3; ++d) Dr.L[d] = L;
        Dr.dmax = pr.max_displacement;
        Dr.drot = pr.max_rotation;
    }
    // sublattice colours (one value for the whole batch: the passes are shared launches)
    c->col = p[0].sublattice_colors == 3 ? 3 : 2;
    c->npass = c->col * c->col * c->col;
    for (int r = 1; r < n_replicas; ++r)
        if (p[r].sublattice_colors != p[0].sublattice_colors) {
            mcgb_set_error("replica %d: all replicas must share one sublattice_colors", r);
            mcgb_destroy(c);
            return MCGB_ERR_ARG;
        }
    // wavefronts per active cell and LDS staging capacity (1.10x the mean stencil, 27 or 125
    // cells, plus 32, within 64 KiB; a larger stencil takes the HBM path)
    c->waves = p[0].waves_per_cell;
    if (c->waves != 1 && c->waves != 2 && c->waves != 3 && c->waves != 4 && c->waves != 8) {
        mcgb_set_error("waves_per_cell must be 1, 2, 3, 4 or 8 (got %d)", c->waves);
        mcgb_destroy(c);
        return MCGB_ERR_ARG;
    }
    if (p[0].pass_schedule == 3) c->waves = 2;         // one move per wavefront, two in flight
    const double ncells_stencil = c->col == 3 ? 125.0 : 27.0;
    int cap = (int)(ncells_stencil * occ_max * 1.10 + 32.0);
    cap = (cap + 31) & ~31;
    if (p[0].stage_capacity > 0) cap = (p[0].stage_capacity + 31) & ~31;
    cap = std::max(32, cap);
    c->cap_request = cap;
    while (cap > 32 && (c->rsize == 8 ? pass_lds<double>(c->waves, cap) : pass_lds<float>(c->waves, cap)) > 65536)
        cap -= 32;
    c->cap = cap;
    c->lds = c->rsize == 8 ? pass_lds_seq<double>(c->waves, cap) : pass_lds_seq<float>(c->waves, cap);
    c->lds_spec = c->rsize == 8 ? pass_lds_spec<double>(c->waves, cap) : pass_lds_spec<float>(c->waves, cap);
    c->lds_duo = c->rsize == 8 ? DuoLds<double>::bytes(cap) : DuoLds<float>::bytes(cap);
    c->spec = p[0].pass_schedule == 1 && c->col == 2;    // the speculative pass: 8 sublattices
    c->o4 = p[0].pass_schedule == 2 && c->rsize == 4;    // four wavefronts per SIMD: FP32
    c->duo = p[0].pass_schedule == 3;                    // two moves in flight: two wavefronts

    // ---- device memory
    int rc = 0;
    const size_t N = c->Ntot;
    for (int b = 0; b < 2 && !rc; ++b) {
        for (int d = 0; d < 3 && !rc; ++d) rc = dalloc_t(c, &c->q[b][d], N);
        for (int d = 0; d < 3 && !rc; ++d) rc = dalloc(c, &c->u[b][d], N * c->rsize);
        if (!rc) rc = dalloc_t(c, &c->pid[b], N);
    }
    if (!rc) rc = dalloc_t(c, &c->cell_id, N);
    if (!rc) rc = dalloc_t(c, &c->tmp, N);
    if (!rc) rc = dalloc_t(c, &c->list2, N);
    if (!rc) rc = dalloc_t(c, &c->count, c->ncell_tot);
    if (!rc) rc = dalloc_t(c, &c->cell_start, c->ncs_tot);
    if (!rc) rc = dalloc_t(c, &c->cell_start2, c->ncs_tot);
    c->scan_tiles = (c->ncell_max + kScanTile - 1) / kScanTile;
    if (!rc) rc = dalloc_t(c, &c->scan_st, (size_t)c->nrep * c->scan_tiles);
    if (!rc) rc = dalloc_t(c, &c->slot, N);
    if (!rc) rc = dalloc_t(c, &c->tmp_pid, N);
    if (!rc) rc = dalloc_t(c, &c->endp, (size_t)c->nrep * kEndParts);
    if (!rc) rc = dalloc_t(c, &c->acc, c->ncell_tot);
    if (!rc) rc = dalloc_t(c, &c->draws, (size_t)c->Ntot + c->ncell_tot);
    if (!rc) rc = dalloc_t(c, &c->ovf_count, 27);
    if (!rc) rc = dalloc_t(c, &c->ovf_list, (size_t)c->npass * c->nblk);
    if (!rc) rc = dalloc_t(c, &c->efin, N);
    if (!rc) rc = dalloc_t(c, &c->ecap, N);
    if (!rc) rc = dalloc_t(c, &c->eccap, N);
    if (!rc) rc = dalloc_t(c, &c->ehard, N);
    if (!rc) rc = dalloc_t(c, &c->obs, (size_t)n_replicas * 64 * 6);
    if (!rc) rc = dalloc_t(c, &c->series, (size_t)n_replicas * c->series_cap * kSeriesWidth);
    if (!rc) rc = dalloc_t(c, &c->dC, n_replicas);
    if (!rc) rc = dalloc_t(c, &c->dD, n_replicas);
    if (!rc) rc = dalloc_t(c, &c->dBlk, c->nblk);
    if (rc) { mcgb_destroy(c); return rc; }
    // the cell counts are zero between rebuilds (each rebuild's k_sort_gather resets them)
    // and so are the scan's status words (cleared by k_scatter)
    if (hipMemset(c->count, 0, sizeof(int) * (size_t)c->ncell_tot) != hipSuccess ||
        hipMemset(c->scan_st, 0, sizeof(unsigned long long) * (size_t)c->nrep * c->scan_tiles) != hipSuccess) {
        mcgb_set_error("hipMemset of the cell counts failed");
        mcgb_destroy(c);
        return MCGB_ERR_HIP;
    }
    if (hipMemcpy(c->dC, c->hC.data(), sizeof(RepConst) * n_replicas, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dD, D.data(), sizeof(RepDyn) * n_replicas, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dBlk, c->blk_rep.data(), sizeof(int) * c->nblk, hipMemcpyHostToDevice) != hipSuccess) {
        mcgb_set_error("hipMemcpy of the replica tables failed");
        mcgb_destroy(c);
        return MCGB_ERR_HIP;
    }
    // placeholder configuration (spread points) until mcgb_set_state
    {
        std::vector<uint32_t> qq(N);
        std::vector<int> ids(N);
        const uint32_t mul[3] = {0x9E3779B9u, 0x85EBCA6Bu, 0xC2B2AE35u};
        for (int r = 0; r < n_replicas; ++r)
            for (int i = 0; i < c->hC[r].N; ++i) ids[c->hC[r].part_off + i] = i;
        for (int d = 0; d < 3; ++d) {
            for (size_t i = 0; i < N; ++i) qq[i] = (uint32_t)(ids[i] + 1) * mul[d];
            if (hipMemcpy(c->q[0][d], qq.data(), N * 4, hipMemcpyHostToDevice) != hipSuccess) rc = MCGB_ERR_HIP;
        }
        std::vector<char> ub(N * c->rsize, 0);
        for (size_t i = 0; i < N; ++i) {
            if (c->rsize == 8) ((double*)ub.data())[i] = 1.0;
            else ((float*)ub.data())[i] = 1.0f;
        }
        for (int d = 0; d < 3; ++d) {
            std::vector<char> z(N * c->rsize, 0);
            if (hipMemcpy(c->u[0][d], d == 2 ? ub.data() : z.data(), N * c->rsize, hipMemcpyHostToDevice) != hipSuccess)
                rc = MCGB_ERR_HIP;
        }
        if (hipMemcpy(c->pid[0], ids.data(), N * 4, hipMemcpyHostToDevice) != hipSuccess) rc = MCGB_ERR_HIP;
        if (rc) {
            mcgb_set_error("initial upload failed");
            mcgb_destroy(c);
            return rc;
        }
    }
    c->cur = 0;
    *out = c;
    return MCGB_OK;
}

// Uploads replica r's state.  rec = null: orientations are normalised, the sweep index, step
// sizes and adaptation count carry over and the other counters restart.  rec given (resume):
// orientations go up as they are and the record's counters are restored.
static int set_state_impl(mcgb_ctx* c, int r, int n, const uint32_t* qx, const uint32_t* qy,
                          const uint32_t* qz, const double* ux, const double* uy, const double* uz,
                          const double box[3], const mcgb_resume* rec = nullptr) {
    const RepConst& C = c->hC[r];
    if (n != C.N) {
        mcgb_set_error("replica %d holds %d particles, %d given", r, C.N, n);
        return MCGB_ERR_ARG;
    }
    for (int d = 0; d < 3; ++d) {
        if (!(box[d] > 0.0) || mcgb_cell_dim(box[d], c->P[r].cutoff_radius) == 0) {
            mcgb_set_error("replica %d: box %.4f is below 4 cells of the cutoff", r, box[d]);
            return MCGB_ERR_GEOMETRY;
        }
        if ((double)c->P[r].max_displacement >= 0.5 * box[d]) {
            mcgb_set_error("replica %d: max_displacement must be below half the box", r);
            return MCGB_ERR_GEOMETRY;
        }
    }
    // every check before the first upload: a rejected call leaves the replica as it was
    std::vector<double> un[3];
    for (int d = 0; d < 3; ++d) un[d].resize(n);
    for (int i = 0; i < n; ++i) {
        double nn = sqrt(ux[i] * ux[i] + uy[i] * uy[i] + uz[i] * uz[i]);
        if (!(nn > 0.0) || !std::isfinite(nn)) {
            mcgb_set_error("replica %d: orientation %d has zero or non-finite norm", r, i);
            return MCGB_ERR_STATE;
        }
        if (rec) nn = 1.0;                      // a resumed state goes up as it was stored
        un[0][i] = ux[i] / nn; un[1][i] = uy[i] / nn; un[2][i] = uz[i] / nn;
    }
    std::vector<RepDyn> D;
    int rc = read_dyn(c, D);
    if (rc) return rc;
    RepDyn& Dr = D[r];
    uint32_t t = Dr.t;
    int sc = Dr.series_count;
    const float dmax = rec ? rec->max_displacement : Dr.dmax, drot = rec ? rec->max_rotation : Dr.drot;
    const unsigned long long adapts = rec ? rec->adaptations : Dr.adapts;
    for (int d = 0; d < 3; ++d)
        if ((double)dmax >= 0.5 * box[d]) {
            mcgb_set_error("replica %d: the current max_displacement %g must be below half the box",
                           r, (double)dmax);
            return MCGB_ERR_GEOMETRY;
        }
    const int b = c->cur;
    const size_t off = C.part_off;
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i) ids[i] = i;
    HIPCHK(hipMemcpyAsync(c->q[b][0] + off, qx, n * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->q[b][1] + off, qy, n * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->q[b][2] + off, qz, n * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pid[b] + off, ids.data(), n * 4, hipMemcpyHostToDevice, c->st));
    std::vector<float> uf(c->rsize == 4 ? n : 0);
    for (int d = 0; d < 3; ++d) {
        if (c->rsize == 8) {
            HIPCHK(hipMemcpyAsync((double*)c->u[b][d] + off, un[d].data(), n * 8, hipMemcpyHostToDevice, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
        } else {
            for (int i = 0; i < n; ++i) uf[i] = (float)un[d][i];
            HIPCHK(hipMemcpyAsync((float*)c->u[b][d] + off, uf.data(), n * 4, hipMemcpyHostToDevice, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
        }
    }
    memset(&Dr, 0, sizeof(Dr));
    Dr.t = rec ? (uint32_t)rec->sweep_index : (c->state_set[r] ? t : 0u);
    Dr.series_count = sc;
    Dr.dmax = dmax;
    Dr.drot = drot;
    Dr.adapts = adapts;
    if (rec) {
        Dr.ad_att = rec->adapt_attempted;
        Dr.ad_acc = rec->adapt_accepted;
    }
    for (int d = 0; d < 3; ++d) Dr.L[d] = box[d];
    if ((rc = write_dyn(c, D))) return rc;
    // Full sum of every replica into e_res only; the running energy of replica r alone is set
    // (from the sum, or from the record), the other replicas' chains are left as they were.
    if ((rc = recompute(c, 3))) return rc;
    if ((rc = read_dyn(c, D))) return rc;
    RepDyn& Ds = D[r];
    Ds.cap_pairs = rec ? rec->capped_pairs : (long long)Ds.e_res[0];
    Ds.cap_coll = rec ? rec->capped_colloid : (long long)Ds.e_res[1];
    Ds.e_fin = rec ? rec->energy_finite : Ds.e_res[2];
    if ((rc = write_dyn(c, D))) return rc;
    c->state_set[r] = 1;
    return MCGB_OK;
}

int mcgb_get_resume(mcgb_ctx* c, int r, mcgb_resume* out) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (!out) {
        mcgb_set_error("mcgb_get_resume: null record");
        return MCGB_ERR_ARG;
    }
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    const RepDyn& d = D[r];
    out->seed = c->P[r].seed;
    out->sweep_index = d.t;
    out->max_displacement = d.dmax;
    out->max_rotation = d.drot;
    out->adapt_attempted = d.ad_att;
    out->adapt_accepted = d.ad_acc;
    out->adaptations = d.adapts;
    out->capped_pairs = d.cap_pairs;
    out->capped_colloid = d.cap_coll;
    out->energy_finite = d.e_fin;
    return MCGB_OK;
}

int mcgb_resume_state(mcgb_ctx* c, int r, int n, const uint32_t* qx, const uint32_t* qy,
                      const uint32_t* qz, const double* ux, const double* uy, const double* uz,
                      const double box[3], const mcgb_resume* rec) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (!qx || !qy || !qz || !ux || !uy || !uz || !box || !rec) {
        mcgb_set_error("mcgb_resume_state: null array or record");
        return MCGB_ERR_ARG;
    }
    if (rec->seed != c->P[r].seed) {
        mcgb_set_error("mcgb_resume_state: the record's seed %llu is not replica %d's (%llu)",
                       (unsigned long long)rec->seed, r, (unsigned long long)c->P[r].seed);
        return MCGB_ERR_ARG;
    }
    if (rec->sweep_index > 0xffffffffull || !(rec->max_displacement > 0.0f) ||
        !(rec->max_rotation >= 0.0f) || rec->capped_pairs < 0 || rec->capped_colloid < 0 ||
        !std::isfinite(rec->energy_finite)) {
        mcgb_set_error("mcgb_resume_state: sweep index, step sizes or energy out of range");
        return MCGB_ERR_ARG;
    }
    return set_state_impl(c, r, n, qx, qy, qz, ux, uy, uz, box, rec);
}

int mcgb_set_state_fixed(mcgb_ctx* c, int r, int n, const uint32_t* qx, const uint32_t* qy,
                         const uint32_t* qz, const double* ux, const double* uy, const double* uz,
                         const double box[3]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (!qx || !qy || !qz || !ux || !uy || !uz || !box) {
        mcgb_set_error("mcgb_set_state_fixed: null array");
        return MCGB_ERR_ARG;
    }
    return set_state_impl(c, r, n, qx, qy, qz, ux, uy, uz, box);
}

int mcgb_set_state(mcgb_ctx* c, int r, int n, const double* x, const double* y, const double* z,
                   const double* ux, const double* uy, const double* uz, const double box[3]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (!x || !y || !z || !ux || !uy || !uz || !box || n <= 0) {
        mcgb_set_error("mcgb_set_state: null array");
        return MCGB_ERR_ARG;
    }
    std::vector<uint32_t> q[3];
    const double* xs[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        q[d].resize(n);
        for (int i = 0; i < n; ++i) {
            if (!std::isfinite(xs[d][i])) {
                mcgb_set_error("replica %d: position %d is not finite", r, i);
                return MCGB_ERR_STATE;
            }
            q[d][i] = mcgb_to_fixed(xs[d][i], box[d]);
        }
    }
    return set_state_impl(c, r, n, q[0].data(), q[1].data(), q[2].data(), ux, uy, uz, box);
}

int mcgb_get_state_fixed(mcgb_ctx* c, int r, int n, uint32_t* qx, uint32_t* qy, uint32_t* qz,
                         double* ux, double* uy, double* uz, double box[3]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    const RepConst& C = c->hC[r];
    if (n != C.N) {
        mcgb_set_error("replica %d holds %d particles, %d requested", r, C.N, n);
        return MCGB_ERR_ARG;
    }
    const int b = c->cur;
    const size_t off = C.part_off;
    std::vector<uint32_t> q[3];
    std::vector<int> ids(n);
    std::vector<double> ud[3];
    std::vector<float> uf(n);
    for (int d = 0; d < 3; ++d) {
        q[d].resize(n);
        ud[d].resize(n);
        HIPCHK(hipMemcpyAsync(q[d].data(), c->q[b][d] + off, n * 4, hipMemcpyDeviceToHost, c->st));
        if (c->rsize == 8) {
            HIPCHK(hipMemcpyAsync(ud[d].data(), (double*)c->u[b][d] + off, n * 8, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
        } else {
            HIPCHK(hipMemcpyAsync(uf.data(), (float*)c->u[b][d] + off, n * 4, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            for (int i = 0; i < n; ++i) ud[d][i] = uf[i];
        }
    }
    HIPCHK(hipMemcpyAsync(ids.data(), c->pid[b] + off, n * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int i = 0; i < n; ++i) {
        int p = ids[i];
        if (p < 0 || p >= n) {
            mcgb_set_error("corrupt particle id %d at slot %d", p, i);
            return MCGB_ERR_STATE;
        }
        if (qx) qx[p] = q[0][i];
        if (qy) qy[p] = q[1][i];
        if (qz) qz[p] = q[2][i];
        if (ux) ux[p] = ud[0][i];
        if (uy) uy[p] = ud[1][i];
        if (uz) uz[p] = ud[2][i];
    }
    if (box) {
        std::vector<RepDyn> D;
        if ((rc = read_dyn(c, D))) return rc;
        for (int d = 0; d < 3; ++d) box[d] = D[r].L[d];
    }
    return MCGB_OK;
}

static int snapshot_slot(mcgb_ctx* c, int r, int slot, double** host, hipEvent_t* ev);
static int snapshot_enqueue(mcgb_ctx* c, int r, double* host, hipEvent_t ev);
static void snapshot_copy_out(const double* h, int N, double* x, double* y, double* z, double* ux,
                              double* uy, double* uz, double box[3]);

// The FP64 state in particle order: converted on the device (the same product q L / 2^32 as a
// host conversion) and copied into a pinned record of its own, so a snapshot in flight is kept.
int mcgb_get_state(mcgb_ctx* c, int r, int n, double* x, double* y, double* z, double* ux,
                   double* uy, double* uz, double box[3]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (n != c->hC[r].N) {
        mcgb_set_error("replica %d holds %d particles, %d requested", r, c->hC[r].N, n);
        return MCGB_ERR_ARG;
    }
    if (!c->state_set[r]) {
        mcgb_set_error("replica %d has no configuration", r);
        return MCGB_ERR_STATE;
    }
    double* host;
    hipEvent_t ev;
    if ((rc = snapshot_slot(c, r, 1, &host, &ev))) return rc;
    if ((rc = snapshot_enqueue(c, r, host, ev))) return rc;
    HIPCHK(hipEventSynchronize(ev));
    snapshot_copy_out(host, n, x, y, z, ux, uy, uz, box);
    return MCGB_OK;
}

namespace {
void drop_graphs(mcgb_ctx* c);
}  // namespace

// Delayed-acceptance counters summed over the replicas: stage-1 rejections, stage-2
// rejections, moves that reached the energy test.
static int da_totals(mcgb_ctx* c, unsigned long long out[3]) {
    std::vector<RepDyn> D;
    int rc = read_dyn(c, D);
    if (rc) return rc;
    out[0] = out[1] = out[2] = 0;
    for (const RepDyn& d : D) {
        out[0] += d.da_s1;
        out[1] += d.da_s2;
        out[2] += d.attempted - d.left - d.hard;
    }
    return MCGB_OK;
}

static int sweep_impl(mcgb_ctx* c, int n_sweeps, mcgb_move_trace* trace) {
    for (int r = 0; r < c->nrep; ++r)
        if (!c->state_set[r]) {
            mcgb_set_error("replica %d has no configuration (call mcgb_set_state first)", r);
            return MCGB_ERR_STATE;
        }
    int rc = MCGB_OK;
    for (int left = n_sweeps; left > 0;) {
        // with the guard on, delayed acceptance runs in windows; after a window whose surrogate
        // let through too many moves the exact stage rejects, or let almost none through, the
        // context falls back to the plain kernel for the rest of the call and after it
        const bool guard = c->da_on && c->da_guard_window > 0 && !trace;
        const int chunk = guard ? std::min(left, c->da_guard_window) : left;
        unsigned long long a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        if (guard && (rc = da_totals(c, a))) return rc;
        rc = c->prec == 64 ? enqueue_sweeps<double>(c, chunk, trace)
                           : enqueue_sweeps<float>(c, chunk, trace);
        if (rc) return rc;
        if (guard) {
            if ((rc = da_totals(c, b))) return rc;
            const unsigned long long s1 = b[0] - a[0], s2 = b[1] - a[1], ev = b[2] - a[2];
            if (ev > 0) {
                const double a1 = (double)(ev - s1) / (double)ev;
                const double r2 = ev > s1 ? (double)s2 / (double)(ev - s1) : 0.0;
                c->da_rates[0] = r2;
                c->da_rates[1] = a1;
                if (r2 > c->da_guard_s2_max || a1 < c->da_guard_a1_min) {
                    c->da_on = 0;
                    drop_graphs(c);
                    c->da_fallbacks += 1;
                }
            }
        }
        left -= chunk;
    }
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    for (int r = 0; r < c->nrep; ++r) {
        double e = (double)D[r].cap_pairs * 1e9 + (double)D[r].cap_coll * 1e8 + D[r].e_fin;
        if (!std::isfinite(e)) {
            mcgb_set_error("replica %d: running energy is not finite", r);
            return MCGB_ERR_STATE;
        }
    }
    return MCGB_OK;
}

int mcgb_sweep_async(mcgb_ctx* c, int n_sweeps) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (n_sweeps < 0) {
        mcgb_set_error("n_sweeps must be >= 0");
        return MCGB_ERR_ARG;
    }
    for (int r = 0; r < c->nrep; ++r)
        if (!c->state_set[r]) {
            mcgb_set_error("replica %d has no configuration (call mcgb_set_state first)", r);
            return MCGB_ERR_STATE;
        }
    return c->prec == 64 ? enqueue_sweeps<double>(c, n_sweeps, nullptr, true)
                         : enqueue_sweeps<float>(c, n_sweeps, nullptr, true);
}

int mcgb_sync(mcgb_ctx* c) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    for (int r = 0; r < c->nrep; ++r) {
        double e = (double)D[r].cap_pairs * 1e9 + (double)D[r].cap_coll * 1e8 + D[r].e_fin;
        if (!std::isfinite(e)) {
            mcgb_set_error("replica %d: running energy is not finite", r);
            return MCGB_ERR_STATE;
        }
    }
    return MCGB_OK;
}

// Device-side conversion of replica r to FP64 particle order and its copy into the pinned host
// record `host` (6 N + 3 doubles), behind event `ev`, after the work already enqueued.
static int snapshot_enqueue(mcgb_ctx* c, int r, double* host, hipEvent_t ev) {
    const RepConst& C = c->hC[r];
    const size_t rec = 6 * (size_t)C.N + 3;
    int rc;
    if (!c->snap_dev && (rc = dalloc_t(c, &c->snap_dev, 6 * (size_t)c->Ntot + 3 * (size_t)c->nrep)))
        return rc;
    double* stage = c->snap_dev + 6 * (size_t)C.part_off + 3 * (size_t)r;
    const int b = c->cur;
    const dim3 grid((C.N + 255) / 256);
    if (c->prec == 64)
        k_snapshot<double><<<grid, 256, 0, c->st>>>(
            c->dC, c->dD, r, c->q[b][0], c->q[b][1], c->q[b][2], (const double*)c->u[b][0],
            (const double*)c->u[b][1], (const double*)c->u[b][2], c->pid[b], stage);
    else
        k_snapshot<float><<<grid, 256, 0, c->st>>>(
            c->dC, c->dD, r, c->q[b][0], c->q[b][1], c->q[b][2], (const float*)c->u[b][0],
            (const float*)c->u[b][1], (const float*)c->u[b][2], c->pid[b], stage);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, stage, rec * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipEventRecord(ev, c->st));
    return MCGB_OK;
}

// Pinned host record and event of slot `slot` (0: user snapshots, 1: mcgb_get_state) of
// replica r, allocated on first use.
static int snapshot_slot(mcgb_ctx* c, int r, int slot, double** host, hipEvent_t* ev) {
    if (c->snap_host.empty()) {
        c->snap_host.assign(2 * (size_t)c->nrep, nullptr);
        c->snap_ev.assign(2 * (size_t)c->nrep, nullptr);
        c->snap_pending.assign(c->nrep, 0);
    }
    const size_t k = 2 * (size_t)r + slot;
    const size_t rec = 6 * (size_t)c->hC[r].N + 3;
    if (!c->snap_ev[k]) HIPCHK(hipEventCreateWithFlags(&c->snap_ev[k], hipEventDisableTiming));
    if (!c->snap_host[k]) {
        hipError_t e = hipHostMalloc((void**)&c->snap_host[k], rec * sizeof(double), hipHostMallocDefault);
        if (e != hipSuccess) {
            c->snap_host[k] = nullptr;
            mcgb_set_error("hipHostMalloc(%zu bytes) failed: %s", rec * sizeof(double), hipGetErrorString(e));
            return MCGB_ERR_NOMEM;
        }
    }
    *host = c->snap_host[k];
    *ev = c->snap_ev[k];
    return MCGB_OK;
}

static void snapshot_copy_out(const double* h, int N, double* x, double* y, double* z, double* ux,
                              double* uy, double* uz, double box[3]) {
    double* dst[6] = {x, y, z, ux, uy, uz};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) memcpy(dst[k], h + (size_t)k * N, (size_t)N * sizeof(double));
    if (box)
        for (int d = 0; d < 3; ++d) box[d] = h[6 * (size_t)N + d];
}

int mcgb_snapshot_begin(mcgb_ctx* c, int r) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (!c->state_set[r]) {
        mcgb_set_error("replica %d has no configuration", r);
        return MCGB_ERR_STATE;
    }
    double* host;
    hipEvent_t ev;
    if ((rc = snapshot_slot(c, r, 0, &host, &ev))) return rc;
    if ((rc = snapshot_enqueue(c, r, host, ev))) return rc;
    c->snap_pending[r] = 1;
    return MCGB_OK;
}

int mcgb_snapshot_wait(mcgb_ctx* c, int r, int n, double* x, double* y, double* z, double* ux,
                       double* uy, double* uz, double box[3]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (c->snap_pending.empty() || !c->snap_pending[r]) {
        mcgb_set_error("replica %d: no snapshot in flight (call mcgb_snapshot_begin first)", r);
        return MCGB_ERR_STATE;
    }
    const int N = c->hC[r].N;
    if (n != N) {
        mcgb_set_error("replica %d holds %d particles, %d requested", r, N, n);
        return MCGB_ERR_ARG;
    }
    HIPCHK(hipEventSynchronize(c->snap_ev[2 * (size_t)r]));
    c->snap_pending[r] = 0;
    snapshot_copy_out(c->snap_host[2 * (size_t)r], N, x, y, z, ux, uy, uz, box);
    return MCGB_OK;
}

int mcgb_sweep(mcgb_ctx* c, int n_sweeps) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (n_sweeps < 0) {
        mcgb_set_error("n_sweeps must be >= 0");
        return MCGB_ERR_ARG;
    }
    return sweep_impl(c, n_sweeps, nullptr);
}

int mcgb_sweep_traced(mcgb_ctx* c, mcgb_move_trace* trace, int capacity, uint32_t info[36]) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (!trace || capacity <= 0) {
        mcgb_set_error("mcgb_sweep_traced: need a trace buffer");
        return MCGB_ERR_ARG;
    }
    if (c->nrep != 1) {
        mcgb_set_error("mcgb_sweep_traced: the context must hold exactly one replica");
        return MCGB_ERR_ARG;
    }
    if (c->da_on) {
        mcgb_set_error("mcgb_sweep_traced: traces record plain Metropolis moves; disable delayed acceptance first");
        return MCGB_ERR_ARG;
    }
    const RepConst& C = c->hC[0];
    const int ncap = C.ncap[0] * C.ncap[1] * C.ncap[2];
    const size_t need = (size_t)C.N + (size_t)ncap;     // n_cells * ceil(N / n_cells) <= N + n_cells
    if (c->dtrace_cap < need) {
        if (c->dtrace) {
            (void)hipFree(c->dtrace);
            c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)c->dtrace), c->allocs.end());
        }
        c->dtrace = nullptr;
        if ((rc = dalloc_t(c, &c->dtrace, need))) return rc;
        c->dtrace_cap = need;
    }
    if ((rc = sweep_impl(c, 1, c->dtrace))) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    const size_t nrec = (size_t)D[0].ncell * (size_t)D[0].k;
    if ((size_t)capacity < nrec) {
        mcgb_set_error("trace capacity %d < %zu records", capacity, nrec);
        return MCGB_ERR_CAPACITY;
    }
    HIPCHK(hipMemcpy(trace, c->dtrace, nrec * sizeof(mcgb_move_trace), hipMemcpyDeviceToHost));
    if (info) {
        memset(info, 0, 36 * sizeof(uint32_t));
        for (int d = 0; d < 3; ++d) info[d] = D[0].g[d];
        info[3] = (uint32_t)D[0].k;
        info[4] = (uint32_t)D[0].ncell;
        info[5] = (uint32_t)c->npass;
        for (int p = 0; p < c->npass; ++p) info[6 + p] = D[0].order[p];
    }
    return MCGB_OK;
}

int mcgb_get_stats(mcgb_ctx* c, int r, mcgb_stats* out) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (!out) { mcgb_set_error("null stats"); return MCGB_ERR_ARG; }
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    const RepDyn& d = D[r];
    memset(out, 0, sizeof(*out));
    out->sweeps = d.t;
    out->attempted = d.attempted;
    out->accepted = d.accepted;
    out->rejected_left_cell = d.left;
    out->rejected_hard_core = d.hard;
    out->candidates = d.cand;
    out->pairs = d.pairs;
    out->colloid_evals = d.coll;
    out->volume_attempted = d.vol_att;
    out->volume_accepted = d.vol_acc;
    out->cap_pairs = d.cap_pairs;
    out->cap_colloid = d.cap_coll;
    out->energy_finite = d.e_fin;
    out->energy = (double)d.cap_pairs * 1e9 + (double)d.cap_coll * 1e8 + d.e_fin;
    for (int k = 0; k < 3; ++k) out->box[k] = d.L[k];
    out->volume = d.L[0] * d.L[1] * d.L[2];
    out->density = c->hC[r].N / (out->volume - c->hC[r].vcoll);
    out->max_drift = d.max_drift;
    out->resummations = d.resums;
    out->da_stage1_rejected = d.da_s1;
    out->da_stage2_rejected = d.da_s2;
    out->max_drift_forced = d.max_drift_forced;
    out->resummations_forced = d.resums_forced;
    return MCGB_OK;
}

int mcgb_get_step_sizes(mcgb_ctx* c, int r, float* dmax, float* drot, uint64_t* adaptations) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    if (dmax) *dmax = D[r].dmax;
    if (drot) *drot = D[r].drot;
    if (adaptations) *adaptations = D[r].adapts;
    return MCGB_OK;
}

int mcgb_set_step_sizes(mcgb_ctx* c, int r, float dmax, float drot) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    RepDyn& d = D[r];
    if (!(dmax >= 0.0f) || !(drot >= 0.0f)) {
        mcgb_set_error("mcgb_set_step_sizes: step sizes must be >= 0");
        return MCGB_ERR_ARG;
    }
    for (int k = 0; k < 3; ++k)
        if ((double)dmax >= 0.5 * d.L[k]) {
            mcgb_set_error("mcgb_set_step_sizes: max_displacement must be below half the box");
            return MCGB_ERR_GEOMETRY;
        }
    d.dmax = dmax;
    d.drot = drot;
    d.ad_att = d.ad_acc = 0;
    return write_dyn(c, D);
}

int mcgb_reset_counters(mcgb_ctx* c, int r) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    RepDyn& d = D[r];
    d.attempted = d.accepted = d.left = d.hard = d.cand = d.pairs = d.coll = d.vol_att = d.vol_acc = 0;
    d.max_drift = d.max_drift_forced = 0.0;
    d.overflow = 0;
    d.resums = d.resums_forced = 0;
    d.da_s1 = d.da_s2 = 0;
    return write_dyn(c, D);
}

int mcgb_total_energy(mcgb_ctx* c, int r, double out[4]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if ((rc = recompute(c, 3))) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    const double* e = D[r].e_res;
    out[0] = e[0] * 1e9 + e[1] * 1e8 + e[2];
    out[1] = e[0];
    out[2] = e[1];
    out[3] = e[2];
    return MCGB_OK;
}

int mcgb_resum_energy(mcgb_ctx* c, int r, double* drift) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    std::vector<RepDyn> D0;
    if ((rc = read_dyn(c, D0))) return rc;
    double run = (double)D0[r].cap_pairs * 1e9 + (double)D0[r].cap_coll * 1e8 + D0[r].e_fin;
    if ((rc = recompute(c, 0))) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    double now = (double)D[r].cap_pairs * 1e9 + (double)D[r].cap_coll * 1e8 + D[r].e_fin;
    if (drift) *drift = run - now;
    return MCGB_OK;
}

int mcgb_order_parameter(mcgb_ctx* c, int r, double* p2, double director[3]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    rc = c->prec == 64 ? enqueue_observe<double>(c) : enqueue_observe<float>(c);
    if (rc) return rc;
    std::vector<double> part((size_t)c->nrep * 64 * 6);
    HIPCHK(hipMemcpyAsync(part.data(), c->obs, part.size() * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < 64; ++b)
        for (int k = 0; k < 6; ++k) s[k] += part[((size_t)r * 64 + b) * 6 + k];
    const double N = c->hC[r].N;
    double q[6] = {s[0] / N - 1.0 / 3.0, s[1] / N - 1.0 / 3.0, s[2] / N - 1.0 / 3.0, s[3] / N, s[4] / N, s[5] / N};
    double ev[3], V[9];
    mcgb_eigen_sym3(q, ev, V);
    if (p2) *p2 = 1.5 * ev[2];
    if (director) { director[0] = V[2]; director[1] = V[5]; director[2] = V[8]; }
    return MCGB_OK;
}

int mcgb_series(mcgb_ctx* c, int r, int max, double* epp, double* p2, double* dens) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    int cnt = D[r].series_count;
    int avail = std::min(cnt, c->series_cap);
    int n = std::min(avail, max);
    if (n <= 0) return 0;
    std::vector<double> buf((size_t)c->series_cap * kSeriesWidth);
    HIPCHK(hipMemcpy(buf.data(), c->series + (size_t)r * c->series_cap * kSeriesWidth,
                     buf.size() * 8, hipMemcpyDeviceToHost));
    // oldest first: the last `avail` entries of the ring, of which the first n are returned
    int first = cnt - avail;
    for (int k = 0; k < n; ++k) {
        const double* e = &buf[(size_t)((first + k) % c->series_cap) * kSeriesWidth];
        if (epp) epp[k] = e[0];
        if (p2) {
            double q[6] = {e[1] - 1.0 / 3.0, e[2] - 1.0 / 3.0, e[3] - 1.0 / 3.0, e[4], e[5], e[6]};
            double ev[3];
            mcgb_eigen_sym3(q, ev, nullptr);
            p2[k] = 1.5 * ev[2];
        }
        if (dens) dens[k] = e[7];
    }
    return n;
}

int mcgb_series_clear(mcgb_ctx* c) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    for (auto& d : D) d.series_count = 0;
    return write_dyn(c, D);
}

// The pacing percentage is a kernel argument of k_energy_reduce: a captured sweep graph holds it.
static void set_pace_pct(mcgb_ctx* c, int pct) {
    if (pct == c->pace_pct) return;
    for (int k = 0; k < 2; ++k)
        if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
    c->pace_pct = pct;
}

static int set_waves(mcgb_ctx* c, int waves) {
    for (int k = 0; k < 2; ++k)
        if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
    c->waves = waves;
    int cap = c->cap_request;
    while (cap > 32 && (c->rsize == 8 ? pass_lds<double>(waves, cap) : pass_lds<float>(waves, cap)) > 65536)
        cap -= 32;
    c->cap = cap;
    c->lds = c->rsize == 8 ? pass_lds_seq<double>(waves, cap) : pass_lds_seq<float>(waves, cap);
    c->lds_spec = c->rsize == 8 ? pass_lds_spec<double>(waves, cap) : pass_lds_spec<float>(waves, cap);
    c->lds_duo = c->rsize == 8 ? DuoLds<double>::bytes(cap) : DuoLds<float>::bytes(cap);
    return MCGB_OK;
}

int mcgb_autotune(mcgb_ctx* c, int sweeps, int* best_waves, int* best_schedule,
                  double ms_per_pass[20]) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (sweeps <= 0) { mcgb_set_error("mcgb_autotune: sweeps must be > 0"); return MCGB_ERR_ARG; }
    const int cand[5] = {1, 2, 3, 4, 8};
    const bool tm = c->timing;
    c->timing = true;
    double best = 1e300;
    int bw = c->waves;
    int bs = c->duo ? 3 : (c->spec ? 1 : (c->o4 ? 2 : 0));
    for (int sched = 0; sched < 4; ++sched) {
        c->spec = sched == 1;
        c->o4 = sched == 2;
        c->duo = sched == 3;
        // the speculative pass runs 8 sublattices only, the four-wavefront build FP32 only
        if ((c->spec && c->col != 2) || (c->o4 && c->rsize != 4)) {
            if (ms_per_pass)
                for (int k = 0; k < 5; ++k) ms_per_pass[5 * sched + k] = 0.0;
            continue;
        }
        for (int k = 0; k < 5; ++k) {
            // the two-moves schedule runs two wavefronts per cell only
            if (c->duo && cand[k] != 2) {
                if (ms_per_pass) ms_per_pass[5 * sched + k] = 0.0;
                continue;
            }
            set_waves(c, cand[k]);
            if ((rc = sweep_impl(c, sweeps, nullptr))) { c->timing = tm; return rc; }
            if (ms_per_pass) ms_per_pass[5 * sched + k] = c->ms_pass;
            if (c->ms_pass < best) { best = c->ms_pass; bw = cand[k]; bs = sched; }
        }
    }
    c->spec = bs == 1;
    c->o4 = bs == 2;
    c->duo = bs == 3;
    set_waves(c, bw);
    // the sequential schedule's pacing: off, the default and one step above it, each after a
    // sweep that sets the pace from the cells' times under that setting
    if (bs == 0) {
        const int pcts[3] = {0, kPaceDefault, kPaceDefault + 3};
        double pbest = 1e300;
        int pb = c->pace_pct;
        for (int i = 0; i < 3; ++i) {
            set_pace_pct(c, pcts[i]);
            if ((rc = sweep_impl(c, 1, nullptr)) || (rc = sweep_impl(c, sweeps, nullptr))) {
                c->timing = tm;
                return rc;
            }
            if (c->ms_pass < pbest) { pbest = c->ms_pass; pb = pcts[i]; }
        }
        set_pace_pct(c, pb);
    }
    c->timing = tm;
    if (best_waves) *best_waves = bw;
    if (best_schedule) *best_schedule = bs;
    return MCGB_OK;
}

int mcgb_set_pace(mcgb_ctx* c, int percent) {
    if (!c) { mcgb_set_error("null context"); return MCGB_ERR_ARG; }
    if (percent < 0 || percent > 1000) {
        mcgb_set_error("mcgb_set_pace: percent must be in [0, 1000] (got %d)", percent);
        return MCGB_ERR_ARG;
    }
    set_pace_pct(c, percent);
    return MCGB_OK;
}

int mcgb_get_pace(mcgb_ctx* c, int* percent) {
    if (!c || !percent) { mcgb_set_error("null argument"); return MCGB_ERR_ARG; }
    *percent = c->pace_pct;
    return MCGB_OK;
}

int mcgb_set_kernel_timing(mcgb_ctx* c, int on) {
    if (!c) { mcgb_set_error("null context"); return MCGB_ERR_ARG; }
    c->timing = on != 0;
    return MCGB_OK;
}

int mcgb_kernel_times(mcgb_ctx* c, double* ms_pass, double* ms_rebuild, double* ms_sweep) {
    if (!c) { mcgb_set_error("null context"); return MCGB_ERR_ARG; }
    if (ms_pass) *ms_pass = c->ms_pass;
    if (ms_rebuild) *ms_rebuild = c->ms_rebuild;
    if (ms_sweep) *ms_sweep = c->ms_sweep;
    return MCGB_OK;
}

int mcgb_cell_list(mcgb_ctx* c, int r, uint32_t t, int* cell_start, int cell_start_len,
                   int* pid_sorted, int n, uint32_t shift[3], uint32_t order[27], int dims[3]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    const RepConst& C = c->hC[r];
    if (n != C.N) { mcgb_set_error("replica %d holds %d particles", r, C.N); return MCGB_ERR_ARG; }
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    const uint32_t t_saved = D[r].t;
    D[r].t = t;
    if ((rc = write_dyn(c, D))) return rc;
    k_begin<<<(c->nrep + 63) / 64, 64, 0, c->st>>>(c->dC, c->dD, c->nrep, 0, nullptr);
    rc = c->prec == 64 ? enqueue_rebuild<double>(c) : enqueue_rebuild<float>(c);
    if (rc) return rc;
    std::vector<RepDyn> D2;
    if ((rc = read_dyn(c, D2))) return rc;
    const int ncell = D2[r].ncell;
    if (cell_start_len < ncell + 1) {
        mcgb_set_error("cell_start needs %d entries", ncell + 1);
        return MCGB_ERR_CAPACITY;
    }
    HIPCHK(hipMemcpy(cell_start, c->cell_start + C.cs_off, (ncell + 1) * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pid_sorted, c->pid[c->cur] + C.part_off, n * sizeof(int), hipMemcpyDeviceToHost));
    if (shift) for (int d = 0; d < 3; ++d) shift[d] = D2[r].g[d];
    if (order) for (int p = 0; p < 27; ++p) order[p] = p < c->npass ? D2[r].order[p] : 0u;
    if (dims) for (int d = 0; d < 3; ++d) dims[d] = D2[r].n[d];
    D2[r].t = t_saved;
    return write_dyn(c, D2);
}

// ---------------------------------------------------------------- batch kernels

static int batch_consts(const mcgb_params* p, RepConst* C) {
    if (!p) { mcgb_set_error("null parameters"); return MCGB_ERR_ARG; }
    memset(C, 0, sizeof(*C));
    mcgb_params q = *p;
    if (q.system_temperature <= 0.0) q.system_temperature = 1.0;
    if (q.gb_mu_exponent == 0.0) { mcgb_set_error("mu must be nonzero"); return MCGB_ERR_ARG; }
    fill_consts(q, C);
    return MCGB_OK;
}

int mcgb_pair_energy_batch(const mcgb_params* p, int n, const double* ei, const double* ej,
                           const double* dr, int precision, double* u, int* capped) {
    RepConst C;
    int rc = batch_consts(p, &C);
    if (rc) return rc;
    if (n <= 0 || !ei || !ej || !dr || !u || !capped || (precision != 32 && precision != 64)) {
        mcgb_set_error("mcgb_pair_energy_batch: bad arguments");
        return MCGB_ERR_ARG;
    }
    double *d_ei, *d_ej, *d_dr, *d_u;
    int* d_c;
    HIPCHK(hipMalloc(&d_ei, n * 24));
    HIPCHK(hipMalloc(&d_ej, n * 24));
    HIPCHK(hipMalloc(&d_dr, n * 24));
    HIPCHK(hipMalloc(&d_u, n * 8));
    HIPCHK(hipMalloc(&d_c, n * 4));
    HIPCHK(hipMemcpy(d_ei, ei, n * 24, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_ej, ej, n * 24, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_dr, dr, n * 24, hipMemcpyHostToDevice));
    k_pair_batch<<<(n + 255) / 256, 256>>>(C, n, precision, d_ei, d_ej, d_dr, d_u, d_c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(u, d_u, n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(capped, d_c, n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipFree(d_ei)); HIPCHK(hipFree(d_ej)); HIPCHK(hipFree(d_dr));
    HIPCHK(hipFree(d_u)); HIPCHK(hipFree(d_c));
    return MCGB_OK;
}

}  // extern "C"

// Staging of mcgb_pair_energy_batch_noalloc: ei, ej, dr (3 doubles each), u (double) and the
// cap flag (int) of up to MCGB_NOALLOC_MAX_PAIRS pairs, in fixed regions.  One copy per GPU,
// held until the process ends.  In host code the array's name is only a stand-in: its device
// address (per device) comes from hipGetSymbolAddress.
static constexpr size_t kNoAllocMax = MCGB_NOALLOC_MAX_PAIRS;
static constexpr size_t kNoAllocBytes = kNoAllocMax * (3 * 24 + 8 + 4);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kNoAllocDevices] = {};

// The device address of `scratch` on the current device (looked up once per device).
int scratch_ptr(unsigned char** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) {
        mcgb_set_error("mcgb_pair_energy_batch_noalloc: device %d beyond the %d supported", dev,
                       kNoAllocDevices);
        return MCGB_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (unsigned char*)p;
    }
    *out = g_scratch[dev];
    return MCGB_OK;
}
}  // namespace

extern "C" {

int mcgb_pair_energy_batch_noalloc(const mcgb_params* p, int n, const double* ei,
                                   const double* ej, const double* dr, int precision, double* u,
                                   int* capped) {
    if (n > MCGB_NOALLOC_MAX_PAIRS) {
        mcgb_set_error("mcgb_pair_energy_batch_noalloc: %d pairs exceed the staging maximum of %d "
                       "(use mcgb_pair_energy_batch, or batches of at most %d)", n,
                       MCGB_NOALLOC_MAX_PAIRS, MCGB_NOALLOC_MAX_PAIRS);
        return MCGB_ERR_CAPACITY;
    }
    RepConst C;
    int rc = batch_consts(p, &C);
    if (rc) return rc;
    if (n <= 0 || !ei || !ej || !dr || !u || !capped || (precision != 32 && precision != 64)) {
        mcgb_set_error("mcgb_pair_energy_batch_noalloc: bad arguments");
        return MCGB_ERR_ARG;
    }
    unsigned char* base = nullptr;
    if ((rc = scratch_ptr(&base))) return rc;
    double* d_ei = (double*)base;
    double* d_ej = (double*)(base + kNoAllocMax * 24);
    double* d_dr = (double*)(base + 2 * kNoAllocMax * 24);
    double* d_u = (double*)(base + 3 * kNoAllocMax * 24);
    int* d_c = (int*)(base + 3 * kNoAllocMax * 24 + kNoAllocMax * 8);
    HIPCHK(hipMemcpy(d_ei, ei, (size_t)n * 24, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_ej, ej, (size_t)n * 24, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_dr, dr, (size_t)n * 24, hipMemcpyHostToDevice));
    k_pair_batch<<<(n + 255) / 256, 256>>>(C, n, precision, d_ei, d_ej, d_dr, d_u, d_c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(u, d_u, (size_t)n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(capped, d_c, (size_t)n * 4, hipMemcpyDeviceToHost));
    return MCGB_OK;
}

int mcgb_colloid_energy_batch(const mcgb_params* p, int n, const double* rel, const double* e,
                              int precision, double* u, int* capped) {
    RepConst C;
    int rc = batch_consts(p, &C);
    if (rc) return rc;
    (void)precision;   // the colloid term is evaluated in FP64 in both precisions
    if (n <= 0 || !rel || !e || !u || !capped) {
        mcgb_set_error("mcgb_colloid_energy_batch: bad arguments");
        return MCGB_ERR_ARG;
    }
    double *d_r, *d_e, *d_u;
    int* d_c;
    HIPCHK(hipMalloc(&d_r, n * 24));
    HIPCHK(hipMalloc(&d_e, n * 24));
    HIPCHK(hipMalloc(&d_u, n * 8));
    HIPCHK(hipMalloc(&d_c, n * 4));
    HIPCHK(hipMemcpy(d_r, rel, n * 24, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_e, e, n * 24, hipMemcpyHostToDevice));
    k_colloid_batch<<<(n + 255) / 256, 256>>>(C, n, d_r, d_e, d_u, d_c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(u, d_u, n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(capped, d_c, n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipFree(d_r)); HIPCHK(hipFree(d_e)); HIPCHK(hipFree(d_u)); HIPCHK(hipFree(d_c));
    return MCGB_OK;
}

int mcgb_philox_batch(int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    if (n <= 0 || !ctr || !key || !out) {
        mcgb_set_error("mcgb_philox_batch: bad arguments");
        return MCGB_ERR_ARG;
    }
    uint32_t *d_c, *d_k, *d_o;
    HIPCHK(hipMalloc(&d_c, n * 16));
    HIPCHK(hipMalloc(&d_k, n * 8));
    HIPCHK(hipMalloc(&d_o, n * 16));
    HIPCHK(hipMemcpy(d_c, ctr, n * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_k, key, n * 8, hipMemcpyHostToDevice));
    k_philox_batch<<<(n + 255) / 256, 256>>>(n, d_c, d_k, d_o);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, d_o, n * 16, hipMemcpyDeviceToHost));
    HIPCHK(hipFree(d_c)); HIPCHK(hipFree(d_k)); HIPCHK(hipFree(d_o));
    return MCGB_OK;
}

// The move draws the last sweep of replica r made (k_count_draws), in record order
// pass*(nact*k) + active_cell*k + move.  Returns the number of records.
int mcgb_sweep_draws(mcgb_ctx* c, int r, int capacity, uint32_t* j, int32_t* dq, double* v,
                     double* thr, uint32_t info[4]) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (!info || capacity < 0) {
        mcgb_set_error("mcgb_sweep_draws: null info or negative capacity");
        return MCGB_ERR_ARG;
    }
    std::vector<RepDyn> D;
    if ((rc = read_dyn(c, D))) return rc;
    const RepDyn& d = D[r];
    if (!c->state_set[r] || d.t == 0) {
        mcgb_set_error("mcgb_sweep_draws: replica %d has run no sweep", r);
        return MCGB_ERR_STATE;
    }
    const long long n = (long long)c->npass * d.nact * d.k;
    if (n > capacity) {
        mcgb_set_error("mcgb_sweep_draws: %lld records, capacity %d", n, capacity);
        return MCGB_ERR_CAPACITY;
    }
    std::vector<MoveDraw> h((size_t)n);
    HIPCHK(hipMemcpyAsync(h.data(), c->draws + c->hC[r].rec_off, sizeof(MoveDraw) * (size_t)n,
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (long long i = 0; i < n; ++i) {
        if (j) j[i] = h[i].j;
        for (int a = 0; a < 3; ++a) {
            if (dq) dq[3 * i + a] = h[i].dq[a];
            if (v) v[3 * i + a] = h[i].v[a];
        }
        if (thr) thr[i] = h[i].thr;
    }
    info[0] = d.t - 1u;
    info[1] = (uint32_t)d.k;
    info[2] = (uint32_t)d.nact;
    info[3] = (uint32_t)c->npass;
    return (int)n;
}

// ---------------------------------------------------------------- learned first stage

}  // extern "C"

namespace {

uint16_t f32_to_bf16(float f) {      // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// FP32 parameters -> the per-lane matrix-core operands of mlp_tile32.
void pack_mlp(const float* P, MlpPacked* M) {
    memset(M, 0, sizeof(*M));
    for (int lane = 0; lane < 64; ++lane) {
        const int o = lane & 31, g = lane >> 5;
        for (int t = 0; t < 2; ++t) {
            const int h = 32 * t + o;
            for (int e = 0; e < 8; ++e) {   // group 0: (W1, W1) against (x_hi, x_lo); group 1: b1
                const float v = g == 0 ? P[kOffW1 + 4 * h + (e & 3)] : (e == 0 ? P[kOffB1 + h] : 0.0f);
                M->w1[t][e][lane] = f32_to_bf16(v);
            }
            for (int m = 0; m < 4; ++m)
                for (int e = 0; e < 8; ++e) {
                    const int hin = 32 * (m >> 1) + mfma_row(8 * (m & 1) + e, g);
                    M->w2[t][m][e][lane] = f32_to_bf16(P[kOffW2 + kHid * h + hin]);
                }
            for (int r = 0; r < 16; ++r) {
                const int row = 32 * t + mfma_row(r, g);
                M->b2[t][lane][r] = P[kOffB2 + row];
                M->w3[t][lane][r] = P[kOffW3 + row];
            }
        }
    }
    M->b3 = P[kOffB3];
}

int sur_alloc(mcgb_ctx* c) {
    int rc = MCGB_OK;
    if (!c->sur_P) {
        if ((rc = dalloc_t(c, &c->sur_P, kNParam))) return rc;
        if ((rc = dalloc_t(c, &c->sur_M, 1))) return rc;
        if ((rc = dalloc_t(c, &c->tr_grad, kNParam))) return rc;
        if ((rc = dalloc_t(c, &c->tr_m, kNParam))) return rc;
        if ((rc = dalloc_t(c, &c->tr_v, kNParam))) return rc;
        if ((rc = dalloc_t(c, &c->tr_loss, 1))) return rc;
        if ((rc = dalloc_t(c, &c->da_hist, (size_t)c->nrep * (kDaBins + 2)))) return rc;
        HIPCHK(hipMemset(c->da_hist, 0, sizeof(unsigned long long) * c->nrep * (kDaBins + 2)));
        HIPCHK(hipMemset(c->sur_P, 0, sizeof(float) * kNParam));
        MlpPacked z;
        std::vector<float> zp(kNParam, 0.0f);
        pack_mlp(zp.data(), &z);
        HIPCHK(hipMemcpy(c->sur_M, &z, sizeof(z), hipMemcpyHostToDevice));
    }
    return rc;
}

int train_alloc(mcgb_ctx* c, int nblk) {
    if (nblk <= c->tr_nblk) return MCGB_OK;
    int rc;
    if ((rc = dalloc_t(c, &c->tr_partial, (size_t)nblk * kNParam))) return rc;
    if ((rc = dalloc_t(c, &c->tr_lpart, (size_t)nblk))) return rc;
    c->tr_nblk = nblk;
    return MCGB_OK;
}

void drop_graphs(mcgb_ctx* c) {
    for (int k = 0; k < 2; ++k)
        if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
}

// partial gradients of nblk workgroups -> c->tr_grad, loss sum -> c->tr_loss
int enqueue_grad(mcgb_ctx* c, const float* P, const float* data, long long ndata, const float* xs,
                 const float* ts, int nexp, uint32_t step, uint32_t seed, float clip, float wfloor,
                 int nblk, float inv_n) {
    k_mlp_grad<<<nblk, 64, 0, c->st>>>(P, data, ndata, xs, ts, nexp, step, seed, 0x6D6C7000u, clip,
                                       wfloor, inv_n, c->tr_partial, c->tr_lpart);
    k_mlp_reduce<<<(kNParam + 255) / 256, 256, 0, c->st>>>(c->tr_partial, nblk, c->tr_lpart,
                                                          c->tr_grad, c->tr_loss);
    HIPCHK(hipGetLastError());
    return MCGB_OK;
}

}  // namespace

extern "C" {

int mcgb_surrogate_set(mcgb_ctx* c, const float* params) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (!params) { mcgb_set_error("mcgb_surrogate_set: null parameters"); return MCGB_ERR_ARG; }
    for (int k = 0; k < kNParam; ++k)
        if (!std::isfinite(params[k])) {
            mcgb_set_error("mcgb_surrogate_set: parameter %d is not finite", k);
            return MCGB_ERR_ARG;
        }
    if ((rc = sur_alloc(c))) return rc;
    MlpPacked M;
    pack_mlp(params, &M);
    HIPCHK(hipMemcpy(c->sur_P, params, sizeof(float) * kNParam, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->sur_M, &M, sizeof(M), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->tr_m, 0, sizeof(float) * kNParam));
    HIPCHK(hipMemset(c->tr_v, 0, sizeof(float) * kNParam));
    c->adam_t = 0;
    c->sur_set = true;
    return MCGB_OK;
}

int mcgb_surrogate_get(mcgb_ctx* c, float* params) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (!params) { mcgb_set_error("mcgb_surrogate_get: null output"); return MCGB_ERR_ARG; }
    if ((rc = sur_alloc(c))) return rc;
    HIPCHK(hipMemcpy(params, c->sur_P, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
    return MCGB_OK;
}

int mcgb_surrogate_init(mcgb_ctx* c, uint32_t seed) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for every weight and bias of a layer (Philox draws)
    std::vector<float> P(kNParam);
    const uint32_t key[2] = {seed, 0x5EEDu};
    for (int k = 0; k < kNParam; ++k) {
        const uint32_t ctr[4] = {(uint32_t)k, 0u, 0u, 0x1A1A0000u};
        uint32_t o[4];
        mcgb_philox_host(ctr, key, o);
        const double u = (o[0] + 0.5) / 4294967296.0 * 2.0 - 1.0;
        const double fan = k < kOffW2 ? 4.0 : (k < kOffW3 ? 64.0 : 64.0);
        P[k] = (float)(u / sqrt(fan));
    }
    return mcgb_surrogate_set(c, P.data());
}

int mcgb_surrogate_eval(mcgb_ctx* c, int n, const float* x, int use_mfma, float* y) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (n <= 0 || !x || !y) { mcgb_set_error("mcgb_surrogate_eval: bad arguments"); return MCGB_ERR_ARG; }
    if ((rc = sur_alloc(c))) return rc;
    float *dx, *dy;
    HIPCHK(hipMalloc(&dx, sizeof(float) * 4 * (size_t)n));
    HIPCHK(hipMalloc(&dy, sizeof(float) * (size_t)n));
    HIPCHK(hipMemcpy(dx, x, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
    const int per = use_mfma ? 32 : 64;
    k_mlp_eval<<<(n + per - 1) / per, 64, 0, c->st>>>(c->sur_P, c->sur_M, n, dx, use_mfma ? 1 : 0, dy);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(y, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    HIPCHK(hipFree(dx));
    HIPCHK(hipFree(dy));
    return MCGB_OK;
}

int mcgb_surrogate_grad(mcgb_ctx* c, const float* params, int n, const float* x, const float* t,
                        float clip, float* grad, double* loss) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (n <= 0 || !x || !t || !grad) { mcgb_set_error("mcgb_surrogate_grad: bad arguments"); return MCGB_ERR_ARG; }
    if ((rc = sur_alloc(c))) return rc;
    const int nblk = (n + 63) / 64;
    if ((rc = train_alloc(c, nblk))) return rc;
    float *dP = c->sur_P, *dx, *dt;
    if (params) {
        HIPCHK(hipMalloc(&dP, sizeof(float) * kNParam));
        HIPCHK(hipMemcpy(dP, params, sizeof(float) * kNParam, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&dx, sizeof(float) * 4 * (size_t)n));
    HIPCHK(hipMalloc(&dt, sizeof(float) * (size_t)n));
    HIPCHK(hipMemcpy(dx, x, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dt, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    if ((rc = enqueue_grad(c, dP, nullptr, 0, dx, dt, n, 0, 0, clip, 0.0f, nblk, 1.0f / (float)n))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(grad, c->tr_grad, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
    double l;
    HIPCHK(hipMemcpy(&l, c->tr_loss, sizeof(double), hipMemcpyDeviceToHost));
    if (loss) *loss = l / n;
    if (params) HIPCHK(hipFree(dP));
    HIPCHK(hipFree(dx));
    HIPCHK(hipFree(dt));
    return MCGB_OK;
}

int mcgb_surrogate_train(mcgb_ctx* c, int steps, int batch, float lr, float clip, float wfloor,
                         uint32_t seed, float* losses) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (steps <= 0 || batch < 64 || !(lr > 0.0f) || !(clip > 0.0f)) {
        mcgb_set_error("mcgb_surrogate_train: need steps > 0, batch >= 64, lr > 0, clip > 0");
        return MCGB_ERR_ARG;
    }
    if (c->pairs_n <= 0) {
        mcgb_set_error("mcgb_surrogate_train: no training pairs (mcgb_emit_pairs)");
        return MCGB_ERR_STATE;
    }
    if ((rc = sur_alloc(c))) return rc;
    const int nblk = (batch + 63) / 64;
    if ((rc = train_alloc(c, nblk))) return rc;
    double* dl = nullptr;
    if (losses) HIPCHK(hipMalloc(&dl, sizeof(double) * (size_t)steps));
    const float inv = 1.0f / (float)(nblk * 64);
    for (int k = 0; k < steps; ++k) {
        if ((rc = enqueue_grad(c, c->sur_P, c->pairs, c->pairs_n, nullptr, nullptr, 0, (uint32_t)k,
                               seed, clip, wfloor, nblk, inv)))
            return rc;
        ++c->adam_t;
        // cosine decay from lr to lr/20 over the call
        const float lr_k = (float)(lr / 20.0 + 0.5 * (lr - lr / 20.0) * (1.0 + cos(kPi * k / steps)));
        k_mlp_adam<<<(kNParam + 255) / 256, 256, 0, c->st>>>(c->sur_P, c->tr_grad, c->tr_m, c->tr_v,
                                                             c->adam_t, lr_k, 0.9f, 0.999f, 1e-8f);
        if (dl) HIPCHK(hipMemcpyAsync(dl + k, c->tr_loss, sizeof(double), hipMemcpyDeviceToDevice, c->st));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    std::vector<float> P(kNParam);
    HIPCHK(hipMemcpy(P.data(), c->sur_P, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
    MlpPacked M;
    pack_mlp(P.data(), &M);
    HIPCHK(hipMemcpy(c->sur_M, &M, sizeof(M), hipMemcpyHostToDevice));
    c->sur_set = true;
    if (dl) {
        std::vector<double> h(steps);
        HIPCHK(hipMemcpy(h.data(), dl, sizeof(double) * (size_t)steps, hipMemcpyDeviceToHost));
        for (int k = 0; k < steps; ++k) losses[k] = (float)(h[k] / (nblk * 64));
        HIPCHK(hipFree(dl));
    }
    return MCGB_OK;
}

int mcgb_pairs_reserve(mcgb_ctx* c, long long capacity) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (capacity <= 0) { mcgb_set_error("mcgb_pairs_reserve: capacity must be positive"); return MCGB_ERR_ARG; }
    if (capacity <= c->pairs_cap) return MCGB_OK;
    float* np = nullptr;
    if ((rc = dalloc_t(c, &np, (size_t)capacity * kPairRec))) return rc;
    if (c->pairs) {
        HIPCHK(hipMemcpy(np, c->pairs, sizeof(float) * kPairRec * (size_t)c->pairs_n,
                         hipMemcpyDeviceToDevice));
        auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)c->pairs);
        if (it != c->allocs.end()) c->allocs.erase(it);
        HIPCHK(hipFree(c->pairs));
    }
    c->pairs = np;
    c->pairs_cap = capacity;
    return MCGB_OK;
}

int mcgb_pairs_clear(mcgb_ctx* c) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    c->pairs_n = 0;
    return MCGB_OK;
}

int mcgb_pairs_count(mcgb_ctx* c, long long* n) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (!n) { mcgb_set_error("mcgb_pairs_count: null output"); return MCGB_ERR_ARG; }
    *n = c->pairs_n;
    return MCGB_OK;
}

int mcgb_pairs_get(mcgb_ctx* c, long long first, int n, float* out) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (!out || n < 0 || first < 0 || first + n > c->pairs_n) {
        mcgb_set_error("mcgb_pairs_get: range [%lld, %lld) outside the %lld stored pairs", first,
                       first + n, c->pairs_n);
        return MCGB_ERR_ARG;
    }
    if (n == 0) return MCGB_OK;
    HIPCHK(hipMemcpy(out, c->pairs + (size_t)first * kPairRec, sizeof(float) * kPairRec * (size_t)n,
                     hipMemcpyDeviceToHost));
    return MCGB_OK;
}

int mcgb_pairs_put(mcgb_ctx* c, int n, const float* rec) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !rec)) { mcgb_set_error("mcgb_pairs_put: bad arguments"); return MCGB_ERR_ARG; }
    if (c->pairs_n + n > c->pairs_cap) {
        mcgb_set_error("mcgb_pairs_put: %d more pairs exceed the reserved %lld (%lld stored)", n,
                       c->pairs_cap, c->pairs_n);
        return MCGB_ERR_CAPACITY;
    }
    if (n == 0) return MCGB_OK;
    HIPCHK(hipMemcpy(c->pairs + (size_t)c->pairs_n * kPairRec, rec, sizeof(float) * kPairRec * (size_t)n,
                     hipMemcpyHostToDevice));
    c->pairs_n += n;
    return MCGB_OK;
}

int mcgb_emit_pairs(mcgb_ctx* c, int r, int stride, int offset, long long* added) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if (stride <= 0 || offset < 0) { mcgb_set_error("mcgb_emit_pairs: need stride > 0, offset >= 0"); return MCGB_ERR_ARG; }
    if (!c->pairs) { mcgb_set_error("mcgb_emit_pairs: reserve storage first (mcgb_pairs_reserve)"); return MCGB_ERR_STATE; }
    const int N = c->hC[r].N;
    const int M = offset < N ? (N - offset + stride - 1) / stride : 0;
    if (M == 0) { if (added) *added = 0; return MCGB_OK; }
    // a cell list of the current box (shift 0)
    k_grid_g0<<<(c->nrep + 63) / 64, 64, 0, c->st>>>(c->dC, c->dD, c->nrep);
    rc = c->prec == 64 ? enqueue_rebuild<double>(c) : enqueue_rebuild<float>(c);
    if (rc) return rc;
    int* dcnt;
    long long* doff;
    HIPCHK(hipMalloc(&dcnt, sizeof(int) * (size_t)M));
    HIPCHK(hipMalloc(&doff, sizeof(long long) * (size_t)M));
    const int b = c->cur;
    const int g = (M + 127) / 128;
    if (c->prec == 64)
        k_emit<double><<<g, 128, 0, c->st>>>(c->dC, c->dD, r, 0, c->q[b][0], c->q[b][1], c->q[b][2],
                                             (const double*)c->u[b][0], (const double*)c->u[b][1],
                                             (const double*)c->u[b][2], c->cell_start, stride, offset,
                                             M, dcnt, doff, nullptr);
    else
        k_emit<float><<<g, 128, 0, c->st>>>(c->dC, c->dD, r, 0, c->q[b][0], c->q[b][1], c->q[b][2],
                                            (const float*)c->u[b][0], (const float*)c->u[b][1],
                                            (const float*)c->u[b][2], c->cell_start, stride, offset,
                                            M, dcnt, doff, nullptr);
    HIPCHK(hipGetLastError());
    std::vector<int> cnt(M);
    HIPCHK(hipMemcpyAsync(cnt.data(), dcnt, sizeof(int) * (size_t)M, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    std::vector<long long> off(M);
    long long tot = 0;
    for (int m = 0; m < M; ++m) { off[m] = c->pairs_n + tot; tot += cnt[m]; }
    if (c->pairs_n + tot > c->pairs_cap) {
        HIPCHK(hipFree(dcnt));
        HIPCHK(hipFree(doff));
        mcgb_set_error("mcgb_emit_pairs: %lld more pairs exceed the reserved %lld (%lld stored)", tot,
                       c->pairs_cap, c->pairs_n);
        return MCGB_ERR_CAPACITY;
    }
    HIPCHK(hipMemcpy(doff, off.data(), sizeof(long long) * (size_t)M, hipMemcpyHostToDevice));
    if (c->prec == 64)
        k_emit<double><<<g, 128, 0, c->st>>>(c->dC, c->dD, r, 1, c->q[b][0], c->q[b][1], c->q[b][2],
                                             (const double*)c->u[b][0], (const double*)c->u[b][1],
                                             (const double*)c->u[b][2], c->cell_start, stride, offset,
                                             M, dcnt, doff, c->pairs);
    else
        k_emit<float><<<g, 128, 0, c->st>>>(c->dC, c->dD, r, 1, c->q[b][0], c->q[b][1], c->q[b][2],
                                            (const float*)c->u[b][0], (const float*)c->u[b][1],
                                            (const float*)c->u[b][2], c->cell_start, stride, offset,
                                            M, dcnt, doff, c->pairs);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipFree(dcnt));
    HIPCHK(hipFree(doff));
    c->pairs_n += tot;
    if (added) *added = tot;
    return MCGB_OK;
}

int mcgb_da_enable(mcgb_ctx* c, int on, int judge) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (on && !c->sur_set) {
        mcgb_set_error("mcgb_da_enable: no surrogate parameters (mcgb_surrogate_set/_init/_train)");
        return MCGB_ERR_STATE;
    }
    if ((rc = sur_alloc(c))) return rc;
    drop_graphs(c);
    c->da_on = on ? 1 : 0;
    c->da_judge = judge ? 1 : 0;
    return MCGB_OK;
}

int mcgb_da_guard(mcgb_ctx* c, int window_sweeps, double max_stage2_rejection,
                  double min_stage1_acceptance) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (window_sweeps < 0 || !(max_stage2_rejection >= 0.0) || !(min_stage1_acceptance >= 0.0)) {
        mcgb_set_error("mcgb_da_guard: window >= 0 and non-negative bounds required");
        return MCGB_ERR_ARG;
    }
    c->da_guard_window = window_sweeps;
    c->da_guard_s2_max = max_stage2_rejection;
    c->da_guard_a1_min = min_stage1_acceptance;
    return MCGB_OK;
}

int mcgb_da_status(mcgb_ctx* c, int* on, uint64_t* fallbacks, double rates[2]) {
    int rc = check_ctx(c, 0);
    if (rc) return rc;
    if (on) *on = c->da_on;
    if (fallbacks) *fallbacks = c->da_fallbacks;
    if (rates) { rates[0] = c->da_rates[0]; rates[1] = c->da_rates[1]; }
    return MCGB_OK;
}

int mcgb_da_histogram(mcgb_ctx* c, int r, uint64_t* counts, int clear) {
    int rc = check_ctx(c, r);
    if (rc) return rc;
    if ((rc = sur_alloc(c))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    if (counts)
        HIPCHK(hipMemcpy(counts, c->da_hist + (size_t)r * (kDaBins + 2),
                         sizeof(unsigned long long) * (kDaBins + 2), hipMemcpyDeviceToHost));
    if (clear)
        HIPCHK(hipMemset(c->da_hist + (size_t)r * (kDaBins + 2), 0,
                         sizeof(unsigned long long) * (kDaBins + 2)));
    return MCGB_OK;
}

}  // extern "C"
