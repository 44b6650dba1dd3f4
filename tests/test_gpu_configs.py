"""The benchmark configurations with a colloid, run through the C ABI against the FP64 oracle.

* the benchmark's colloid workload (plain, N = 256000, rho* = 0.32, T* = 1.0, colloid R = 5, the
  library's jittered lattice start with every site within 5.5 of the colloid removed):
  - the same geometry at N = 20000: FP64 sweeps move by move against the oracle, and an FP32
    sweep replayed move by move in FP64 with the distribution of its |dE|-relative errors;
  - at full size: the cell list bit-exact against the oracle, and after six sweeps the capped
    pair and colloid counts exact and the finite running energy against the FP64 recompute;
  - at eight times the full size (N = 2048000, layered start), FP32 and FP64: the same checks.
* delayed acceptance with the R = 5 colloid at T* = 0.8, at reduced N: a zero network gives
  the plain chain bit for bit with the colloid term in stage 1; a poor network samples the same
  <E>/N and P2 as the plain sweep (2 sigma, autocorrelation-corrected).
"""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.replay import replay_sweep
from tests.systems import cond_tol, layered_system

pytestmark = pytest.mark.gpu

RHO, R, EXCL, JITTER = 0.32, 5.0, 5.5, 0.05
NEAR_CAP = 1e6                 # sum |pair terms| of a move holding a near-cap pair


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def colloid_lattice(N, precision, T=1.0, seed=20261203, **kw):
    """Context holding the colloid lattice start (library generator, as the benchmark), its fixed-point
    state, and the oracle constants."""
    p = m.gb_params(N, RHO, T, colloid_radius=R, precision=precision, seed=seed,
                    lattice_jitter=JITTER, colloid_exclusion_radius=EXCL,
                    energy_check_interval=0, **kw)
    pos, u, box = m.init_lattice(N, RHO, colloid_radius=R, exclusion_radius=EXCL, jitter=JITTER,
                                 seed=seed)
    ctx = m.Context(p)
    ctx.set_state(pos, u, box)
    c = oc.consts(temperature=T, colloid_radius=R, colloid_on=1)
    return ctx, c, seed


def test_colloid_lattice_n20000_fp64_sweeps_match_oracle_move_by_move():
    ctx, c, seed = colloid_lattice(20000, 64)
    with ctx:
        q, u, L = ctx.get_state_fixed()
        s = oc.System(c, q, u, L, seed=seed, nthreads=8)
        st = ctx.stats()
        assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]
        assert s.energy[0] > 0                # the lattice start holds capped overlaps
        for sweep in range(2):
            print(f"colloid lattice N=20000 FP64: traced sweep {sweep}", flush=True)
            tr_g, info_g = ctx.sweep_traced()
            tr_o, info_o, _ = s.sweep(trace=True)
            assert np.array_equal(info_g, info_o)
            assert np.array_equal(tr_g["pid"], tr_o["pid"])
            assert np.array_equal(tr_g["outcome"], tr_o["outcome"]), f"sweep {sweep}"
            assert np.array_equal(tr_g["dcap"], tr_o["dcap"])
            assert np.array_equal(tr_g["dccap"], tr_o["dccap"])
            assert np.all(np.abs(tr_g["dfin"] - tr_o["dfin"]) <= cond_tol(1e-12, tr_o["scale"]))
            oc_ = np.bincount(tr_g["outcome"], minlength=5)
            assert oc_[1] > 0 and oc_[3] > 0 and oc_[4] > 1000
        qg, ug, _ = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qg[d], s.q[d])
            np.testing.assert_allclose(ug[d], s.u[d], rtol=0, atol=1e-12)
        st = ctx.stats()
        assert st["accepted"] == s.counters[1] and st["rejected_hard_core"] == s.counters[3]
        assert st["colloid_evals"] == s.counters[6] > 0
        assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]


def test_colloid_lattice_n20000_fp32_replay_and_relative_error_distribution():
    """Every move of an FP32 sweep replayed in FP64 from the same state: decisions follow, cap
    counts are exact, and the distribution of the |dE|-relative error is reported against the
    move's cancellation factor kappa = sum|terms| / |dE|.

    The lane sums and reductions are FP64, so a move's error is that of its FP32 pair terms:
    each term carries the few-ulp FP32 rounding of sigma, amplified by ~12 sigma/(r - sigma +
    sigma0) at contact.  The bound held for every move is 1e-5 sum|terms|, i.e. an error of at
    most 1e-5 kappa relative to |dE| itself: 1e-5 relative wherever old and trial energies do
    not cancel (kappa <= 1); moves with kappa > 1 are the cancellation cases, excluded from the
    plain 1e-5 |dE|-relative bound and held to 1e-5 kappa."""
    ctx, c, seed = colloid_lattice(20000, 32)
    with ctx:
        ctx.sweep(40)                          # relax the lattice start a little
        q0, u0, L = ctx.get_state_fixed()
        t = ctx.stats()["sweeps"]
        tr, info = ctx.sweep_traced()
        got = []
        print("colloid lattice N=20000 FP32: replaying one sweep", flush=True)
        qf, uf, stt = replay_sweep(c, q0, u0, L, seed, 0, t, tr, info, check_tol=1e-5,
                                   fp32=True, collect=got, max_scale=NEAR_CAP)
        q1, u1, _ = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q1[d], qf[d])
    a = np.array(got)
    # near-cap moves: a pair within ~0.05 sigma0 of the 1e9 cap (MCnptColloid.c:2670), energy
    # up to ~1e36, where r - sigma + sigma0 amplifies the FP32 rounding of sigma by up to ~5000
    # (the lattice start has them; no move of an equilibrium state does): held to the
    # condition-aware bound with the FP32 rounding of sigma, ~1e-7
    far = a[:, 2] > NEAR_CAP
    assert np.all(np.abs(a[far, 0] - a[far, 1]) <= cond_tol(1e-7, a[far, 2]))
    print(f"colloid lattice FP32: {far.sum()} near-cap moves within the condition-aware bound", flush=True)
    a = a[~far]
    dg, dr, scale = a[:, 0], a[:, 1], a[:, 2]
    err = np.abs(dg - dr)
    kappa = scale / np.maximum(np.abs(dr), 1e-300)
    good = kappa <= 100.0
    rel = err[good] / np.abs(dr[good])
    pct = np.percentile(rel, [50, 90, 99, 100])
    print(f"colloid lattice FP32 |dE|-relative error over {good.sum()} of {len(a)} moves (kappa <= 100): "
          f"median {pct[0]:.2e}, 90% {pct[1]:.2e}, 99% {pct[2]:.2e}, max {pct[3]:.2e}; "
          f"excluded (kappa > 100): {(~good).sum()}", flush=True)
    for lo, hi in ((0, 10), (10, 100), (100, 1e3), (1e3, 1e4), (1e4, np.inf)):
        b = (kappa > lo) & (kappa <= hi)
        if b.any():
            r = err[b] / np.abs(dr[b])
            print(f"  kappa ({lo:g}, {hi:g}]: {b.sum()} moves, |dE|-relative error median "
                  f"{np.median(r):.2e} max {r.max():.2e}; sum|terms|-relative max "
                  f"{(err[b] / np.maximum(scale[b], 1.0)).max():.2e}", flush=True)
    assert stt["checked"] > 10000 and good.sum() > 1000
    assert np.all(err <= 1e-5 * np.maximum(scale, 1.0))      # 1e-5 kappa |dE| for every move
    nc = kappa <= 1.0
    assert np.all(err[nc] <= 1e-5 * np.abs(dr[nc]))           # no cancellation: 1e-5 |dE|
    assert pct[0] <= 1e-5                                     # typical move (kappa <= 100)


def test_colloid_lattice_n256000_cell_list_caps_and_energy():
    N = 256000
    ctx, c, seed = colloid_lattice(N, 32)
    with ctx:
        q, u, L = ctx.get_state_fixed()
        n = oc.cell_dim(L[0], 4.0)
        assert n == 22
        for t in (0, 5):
            cs, ids, shift, order, dims = ctx.cell_list(t)
            g, o, _ = oc.sweep_setup(seed, 0, t)
            assert list(shift) == list(g) and list(order) == list(o) and list(dims) == [n] * 3
            cs_o, so_o = oc.cell_list(q, np.array([n] * 3, dtype=np.int32), g)
            assert np.array_equal(cs, cs_o) and np.array_equal(ids, so_o)
        e0 = ctx.total_energy()
        st = ctx.stats()
        assert st["cap_pairs"] == e0[1] and st["cap_colloid"] == e0[2] and e0[1] > 0
        ctx.sweep(6)
        st = ctx.stats()
        tot = ctx.total_energy()
        # capped terms are counted, not summed: exact
        assert st["cap_pairs"] == tot[1] and st["cap_colloid"] == tot[2]
        # the energy falls, though not by resolving caps: inside the cap (r - sigma + sigma0 <
        # 0.001) a pair costs a flat 1e9, just outside it up to ~1e36, so Metropolis moves deep
        # overlaps into the cap (the reference's potential, MCnptColloid.c:2670)
        assert st["energy"] < 1e-3 * e0[0]
        # the finite part: sums of FP32 pair terms (FP64 accumulation), each move within 1e-7
        # of its sum |terms|; the near-cap terms of the lattice start make the scale ~|E|
        assert abs(st["energy_finite"] - tot[3]) <= 1e-6 * (abs(tot[3]) + N)
        q1, u1, _ = ctx.get_state_fixed()
        nrm = np.sqrt(u1[0] ** 2 + u1[1] ** 2 + u1[2] ** 2)
        assert np.abs(nrm - 1.0).max() < 1e-6
        # the same cell list again after the sweeps (ids in cell order, counts)
        cs, ids, shift, order, dims = ctx.cell_list(6)
        g, o, _ = oc.sweep_setup(seed, 0, 6)
        cs_o, so_o = oc.cell_list(q1, np.array([n] * 3, dtype=np.int32), g)
        assert np.array_equal(cs, cs_o) and np.array_equal(ids, so_o)


@pytest.mark.parametrize("precision", [32, 64])
def test_eight_times_the_full_size(precision):
    """N = 2048000 at the benchmark's density with the R = 5 colloid (a 46^3 grid, 12167 active cells
    per pass, 8x the benchmark's state), from a layered start: the cell lists bit-exact against
    the oracle, and after three sweeps of the two-wavefront sequential kernel the cap counts
    exact and the running energy equal to the FP64 recompute."""
    N = 2048000
    print(f"N = {N}: FP{precision}", flush=True)
    q, u, L = layered_system(N, RHO, colloid_radius=R, seed=77)
    p = m.gb_params(N, RHO, 1.0, colloid_radius=R, precision=precision, seed=4077,
                    energy_check_interval=0, pass_schedule=0, waves_per_cell=2)
    with m.Context(p) as ctx:
        ctx.set_state_fixed(q, u, L)
        n = oc.cell_dim(L[0], 4.0)
        assert n == 46
        cs, ids, shift, order, dims = ctx.cell_list(0)
        g, _, _ = oc.sweep_setup(4077, 0, 0)
        cs_o, so_o = oc.cell_list(q, np.array([n] * 3, dtype=np.int32), g)
        assert list(dims) == [n] * 3
        assert np.array_equal(cs, cs_o) and np.array_equal(ids, so_o)
        ctx.sweep(3)
        st = ctx.stats()
        assert st["attempted"] > 2 * N and 0 < st["accepted"] < st["attempted"]
        tot = ctx.total_energy()
        assert st["cap_pairs"] == tot[1] and st["cap_colloid"] == tot[2]
        tol = 1e-9 if precision == 64 else 1e-6
        assert abs(st["energy_finite"] - tot[3]) <= tol * (abs(tot[3]) + N)
        q1, _, _ = ctx.get_state_fixed()
        cs, ids, shift, _, _ = ctx.cell_list(3)
        cs_o, so_o = oc.cell_list(q1, np.array([n] * 3, dtype=np.int32), shift)
        assert np.array_equal(cs, cs_o) and np.array_equal(ids, so_o)


# ------------------------------------------------------------------ delayed acceptance, colloid, T* = 0.8

def da_colloid_pair(N, precision, seed=5, pseed=91):
    T = 0.8
    q, u, L = layered_system(N, RHO, colloid_radius=R, seed=seed)
    out = []
    for _ in range(2):
        p = m.gb_params(N, RHO, T, colloid_radius=R, precision=precision, seed=pseed,
                        energy_check_interval=0)
        ctx = m.Context(p)
        ctx.set_state_fixed(q, u, L)
        out.append(ctx)
    return out


@pytest.mark.parametrize("precision", [64, 32])
def test_da_colloid_zero_network_is_the_plain_chain(precision):
    a, b = da_colloid_pair(6000, precision)
    with a, b:
        b.surrogate_set(np.zeros(m.SUR_NPARAM, dtype=np.float32))
        b.da_enable(True)
        a.sweep(4)
        b.sweep(4)
        qa, ua, _ = a.get_state_fixed()
        qb, ub, _ = b.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qa[d], qb[d]) and np.array_equal(ua[d], ub[d])
        sa, sb = a.stats(), b.stats()
        assert sa["accepted"] == sb["accepted"] and sb["da_stage1_rejected"] == 0
        assert sa["colloid_evals"] > 0 and sa["rejected_hard_core"] == sb["rejected_hard_core"]
        # same decisions; the delayed-acceptance kernel's own pair loop may round the running
        # sum differently in the last bits
        assert sb["energy"] == pytest.approx(sa["energy"], rel=1e-12 if precision == 64 else 1e-7)


def series_stats(ctx, sweeps):
    ctx.series_clear()
    ctx.sweep(sweeps)
    e, p2, _ = ctx.series(max_n=sweeps + 16)
    out = []
    for x in (e, p2):
        tau, _ = m.tau_int(x)
        out.append((float(x.mean()), float(x.std() * np.sqrt(max(tau, 1.0) / len(x))), tau))
    return out


def test_da_colloid_poor_network_samples_the_plain_distribution():
    """An untrained network (random weights) as stage 1 changes only the efficiency: <E>/N and
    P2 of the delayed-acceptance chain agree with the plain chain's within 2 combined standard
    errors (tau_int-corrected), with the colloid at T* = 0.8."""
    a, b = da_colloid_pair(4000, 32)
    with a, b:
        b.surrogate_init(21)
        b.da_enable(True)
        for ctx in (a, b):
            ctx.sweep(1000)
        (ea, sea, ta), (pa, spa, tpa) = series_stats(a, 6000)
        (eb, seb, tb), (pb, spb, tpb) = series_stats(b, 6000)
        zE = (ea - eb) / np.hypot(sea, seb)
        zP = (pa - pb) / np.hypot(spa, spb)
        print(f"delayed acceptance, colloid: plain E/N {ea:.4f}+-{sea:.4f} (tau {ta:.0f}) P2 {pa:.4f}+-{spa:.4f} "
              f"(tau {tpa:.0f}); DA E/N {eb:.4f}+-{seb:.4f} (tau {tb:.0f}) P2 {pb:.4f}+-{spb:.4f} "
              f"(tau {tpb:.0f}); z {zE:+.2f} {zP:+.2f}", flush=True)
        assert abs(zE) < 2.0 and abs(zP) < 2.0
        s = b.stats()
        assert s["da_stage1_rejected"] > 0 and s["da_stage2_rejected"] > 0
