"""HIP kernels (through the C ABI of libmcgb.so) against the FP64 oracle and the reference's
vectors.  Tolerances: FP64 path 1e-12 relative; FP32 path 1e-5 relative, measured against the
magnitude of the summed terms (for a per-move dE: the sum of |pair energies| of both states)."""
import os

import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.replay import replay_sweep
from tests.systems import cond_tol, lattice_system, layered_system

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PHYS = np.load(os.path.join(GOLD, "ref_physics.npz"))
SETS = {"gb3521": (1.0, 1.0, 3.0, 5.0, 2.0, 1.0), "gb44_20_1_1": (1.0, 1.0, 4.4, 20.0, 1.0, 1.0),
        "lj_limit": (1.0, 1.0, 1.0, 1.0, 2.0, 1.0), "gb3512_s12": (1.2, 0.7, 3.0, 5.0, 1.0, 2.0)}
# the FP32 pair takes eps through one reciprocal square root for integer exponents mu, nu in
# {1, 2} and through log2/exp2 otherwise: every case, against the FP64 kernel
SETS_FP32 = dict(SETS, gb3522=(1.0, 1.0, 3.0, 5.0, 2.0, 2.0),
                 gb35_frac=(1.0, 1.0, 3.0, 5.0, 1.5, 0.5))


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def params_for(name, R=0.0, precision=64):
    s0, e0, k, kp, mu, nu = SETS_FP32[name]
    p = m.gb_params(1000, 0.3, 1.0, colloid_radius=R, precision=precision, kappa=k,
                    kappa_prime=kp, mu=mu, nu=nu)
    p.gb_sigma_sidetoside, p.gb_epsilon0 = s0, e0
    return p


def test_philox_known_answers_and_oracle():
    kat_ctr = [[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]]
    kat_key = [[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]]
    want = [[0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8],
            [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]]
    out = m.philox_batch(kat_ctr, kat_key)
    assert out.tolist() == want
    rng = np.random.default_rng(5)
    ctr = rng.integers(0, 2 ** 32, size=(500, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, size=(500, 2), dtype=np.uint64).astype(np.uint32)
    out = m.philox_batch(ctr, key)
    for k in range(0, 500, 7):
        assert list(out[k]) == list(oc.philox(ctr[k], key[k]))


@pytest.mark.parametrize("name", sorted(SETS))
def test_pair_fp64_matches_reference_vectors(name):
    # 1e-12 relative, times the pair's condition number 12 (r + sigma) rho / sigma0 near the cap
    p = params_for(name)
    s0, e0, k, kp, mu, nu = SETS[name]
    c = oc.consts(sigma0=s0, epsilon0=e0, kappa=k, kappa_prime=kp, mu=mu, nu=nu)
    ei, ej, dr = PHYS[f"{name}/pair_ei"], PHYS[f"{name}/pair_ej"], PHYS[f"{name}/pair_dr"]
    ref = PHYS[f"{name}/pair_u"]
    u, cap = m.pair_energy_batch(p, ei, ej, dr, 64)
    assert np.array_equal(cap == 1, ref == 1e9)
    mag, rho = _terms(c, ei, ej, dr)          # error measured against 4 eps rho^6 (rho^6 + 1)
    r = np.linalg.norm(dr, axis=1)
    cond = 1.0 + 12.0 * (r + k * s0) * np.abs(rho) / s0
    fin = cap == 0
    err = np.abs(u - ref)[fin] / mag[fin]
    assert np.all(err <= 1e-12 * cond[fin])
    assert np.all(err[rho[fin] < 1.25] <= 1e-12)


def _terms(c, ei, ej, dr):
    """|4 eps rho^6| (rho^6 + 1): the magnitude the FP32 pair error is measured against, and rho."""
    ei = ei / np.linalg.norm(ei, axis=1, keepdims=True)
    ej = ej / np.linalg.norm(ej, axis=1, keepdims=True)
    r = np.linalg.norm(dr, axis=1)
    a, b, cc = (ei * dr).sum(1) / r, (ej * dr).sum(1) / r, (ei * ej).sum(1)
    g = lambda w: 1 - 0.5 * w * ((a + b) ** 2 / (1 + w * cc) + (a - b) ** 2 / (1 - w * cc))
    eps = (1 - c.chi2 * cc * cc) ** (-c.nu / 2) * g(c.chi_prime) ** c.mu
    rho = c.sigma0 / (r - c.sigma0 / np.sqrt(g(c.chi)) + c.sigma0)
    return c.four_eps0 * eps * rho ** 6 * (rho ** 6 + 1), rho


@pytest.mark.parametrize("name", sorted(SETS_FP32))
def test_pair_fp32_within_1e5(name):
    p = params_for(name, precision=32)
    s0, e0, k, kp, mu, nu = SETS_FP32[name]
    c = oc.consts(sigma0=s0, epsilon0=e0, kappa=k, kappa_prime=kp, mu=mu, nu=nu)
    rng = np.random.default_rng(11)
    n = 20000
    ei = rng.normal(size=(n, 3))
    ej = rng.normal(size=(n, 3))
    ei /= np.linalg.norm(ei, axis=1, keepdims=True)
    ej /= np.linalg.norm(ej, axis=1, keepdims=True)
    dr = rng.normal(size=(n, 3))
    dr *= (rng.uniform(0.5, 4.0, size=n) * s0 / np.linalg.norm(dr, axis=1))[:, None]
    # inputs exactly representable in FP32, so both sides see the same configuration
    ei, ej, dr = (a.astype(np.float32).astype(np.float64) for a in (ei, ej, dr))
    u64, cap64 = m.pair_energy_batch(p, ei, ej, dr, 64)
    u32, cap32 = m.pair_energy_batch(p, ei, ej, dr, 32)
    mag, rho = _terms(c, ei, ej, dr)
    # pairs up to rho = 1.25 (U up to ~70 eps0): every energy a thermal configuration holds
    ok = (cap64 == 0) & (rho < 1.25)
    assert ok.sum() > 5000
    rel = np.abs(u32 - u64)[ok] / mag[ok]
    assert rel.max() <= 1e-5, rel.max()
    # deeper inside the core the FP32 error grows with the condition number (~12 rho r)
    deep = (cap64 == 0) & (rho >= 1.25) & (rho < 2.0)
    rel = np.abs(u32 - u64)[deep] / mag[deep]
    assert rel.max() <= 5e-5, rel.max()


@pytest.mark.parametrize("name", ["gb3521", "gb3512_s12"])
@pytest.mark.parametrize("R", [5, 3])
def test_colloid_matches_reference_vectors(name, R):
    p = params_for(name, R=float(R))
    rel, e = PHYS[f"{name}/coll_R{R}_rel"], PHYS[f"{name}/coll_R{R}_e"]
    ref = PHYS[f"{name}/coll_R{R}_u"]
    rel = (50.0 + rel) - 50.0
    u, cap = m.colloid_energy_batch(p, rel, e)
    assert np.array_equal(cap == 1, ref == 1e8)
    np.testing.assert_allclose(u, ref, rtol=1e-12, atol=1e-12)


def make_pair(N=1300, rho=0.25, R=0.0, T=1.0, precision=64, seed=5, resum=0, nthreads=4,
              stage_capacity=0, start="layered", colors=2, cutoff=4.0, **kw):
    if start == "layered":
        q, u, L = layered_system(N, rho, colloid_radius=R, seed=seed)
    else:
        q, u, L = lattice_system(N, rho, colloid_radius=R, seed=seed)
    p = m.gb_params(N, rho, T, colloid_radius=R, precision=precision, seed=900 + seed,
                    energy_check_interval=resum, stage_capacity=stage_capacity,
                    sublattice_colors=colors, cutoff=cutoff, **kw)
    c = oc.consts(temperature=T, colloid_radius=R, colloid_on=1 if R > 0 else 0, colors=colors,
                  cutoff=cutoff)
    s = oc.System(c, q, u, L, seed=900 + seed, nthreads=nthreads, resum_interval=resum)
    ctx = m.Context(p)
    ctx.set_state_fixed(q, u, L)
    return ctx, s, c


@pytest.mark.parametrize("N,R", [(1300, 0.0), (1300, 3.0), (9000, 5.0)])
def test_cell_list_bit_exact(N, R):
    ctx, s, c = make_pair(N=N, R=R)
    with ctx:
        for t in (0, 7, 12345):
            cs, ids, shift, order, dims = ctx.cell_list(t)
            g, o, _ = oc.sweep_setup(s.seed, 0, t)
            assert list(shift) == list(g) and list(order) == list(o)
            n = [oc.cell_dim(s.L[d], c.cutoff) for d in range(3)]
            assert list(dims) == n
            cs_o, so_o = oc.cell_list(s.q, np.array(n, dtype=np.int32), g)
            assert np.array_equal(cs, cs_o)
            assert np.array_equal(ids, so_o)


def test_initial_energy_matches_oracle():
    for R in (0.0, 3.0):
        ctx, s, c = make_pair(R=R, start="lattice", rho=0.3)
        with ctx:
            st = ctx.stats()
            assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]
            # the lattice start holds near-cap pairs with energies ~1e30: condition-aware bound
            tol = cond_tol(1e-12, abs(s.energy[2]))
            assert abs(st["energy_finite"] - s.energy[2]) <= tol
            tot = ctx.total_energy()
            assert tot[1] == s.energy[0] and abs(tot[3] - s.energy[2]) <= tol


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp64_sweeps_match_oracle_move_by_move(R):
    ctx, s, c = make_pair(R=R, precision=64)
    with ctx:
        for sweep in range(4):
            tr_g, info_g = ctx.sweep_traced()
            tr_o, info_o, _ = s.sweep(trace=True)
            assert np.array_equal(info_g, info_o)
            assert np.array_equal(tr_g["pid"], tr_o["pid"])
            assert np.array_equal(tr_g["outcome"], tr_o["outcome"]), f"sweep {sweep}"
            assert np.array_equal(tr_g["dcap"], tr_o["dcap"]) and np.array_equal(tr_g["dccap"], tr_o["dccap"])
            err = np.abs(tr_g["dfin"] - tr_o["dfin"]) / np.maximum(tr_o["scale"], 1.0)
            assert err.max() <= 1e-12
            assert (tr_g["outcome"] == 4).sum() > 100
        q, u, L = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q[d], s.q[d])
            np.testing.assert_allclose(u[d], s.u[d], rtol=0, atol=1e-12)
        st = ctx.stats()
        assert st["attempted"] == s.counters[0] and st["accepted"] == s.counters[1]
        assert st["rejected_left_cell"] == s.counters[2] and st["rejected_hard_core"] == s.counters[3]
        assert st["candidates"] == s.counters[4] and st["pairs"] == s.counters[5]
        assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]
        assert st["energy_finite"] == pytest.approx(s.energy[2], rel=1e-10, abs=1e-6)


def test_fp64_dense_cells_near_colloid_match_oracle():
    # cutoff 6: a 4^3 grid of cells 6.8 wide holding ~90 particles each.  Cells near the
    # colloid with more than 64 particles evaluate both colloid terms per move, the others keep
    # them in the LDS cache; stencils beyond the staging capacity go to the HBM pass.  Decisions
    # and states stay exact against the oracle.
    ctx, s, c = make_pair(N=6000, rho=0.3, R=3.0, precision=64, seed=31, cutoff=6.0)
    with ctx:
        cs, _, _, _, dims = ctx.cell_list(0)
        counts = np.diff(cs)
        assert tuple(dims) == (4, 4, 4) and counts.max() > 64 and counts.min() <= 64
        for sweep in range(2):
            print(f"dense cells: sweep {sweep}", flush=True)
            tr_g, info_g = ctx.sweep_traced()
            tr_o, info_o, _ = s.sweep(trace=True)
            assert np.array_equal(info_g, info_o)
            assert np.array_equal(tr_g["outcome"], tr_o["outcome"]), f"sweep {sweep}"
            assert np.array_equal(tr_g["dcap"], tr_o["dcap"]) and np.array_equal(tr_g["dccap"], tr_o["dccap"])
            err = np.abs(tr_g["dfin"] - tr_o["dfin"]) / np.maximum(tr_o["scale"], 1.0)
            assert err.max() <= 1e-12
        q, u, L = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q[d], s.q[d])
            np.testing.assert_allclose(u[d], s.u[d], rtol=0, atol=1e-12)
        # the same moves untraced (the kernel the sweeps run) land on the same state
        ctx.sweep(2)
        for _ in range(2):
            s.sweep()
        q, u, L = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q[d], s.q[d])
            np.testing.assert_allclose(u[d], s.u[d], rtol=0, atol=1e-12)
        assert ctx.stats()["accepted"] == s.counters[1]


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp64_lattice_start_with_caps(R):
    # random-orientation lattice start: capped and near-cap pairs (finite energies to ~1e36);
    # decisions and cap counts exact, dE within the condition-aware bound
    ctx, s, c = make_pair(R=R, rho=0.3, precision=64, start="lattice")
    with ctx:
        for sweep in range(3):
            tr_g, info_g = ctx.sweep_traced()
            tr_o, info_o, _ = s.sweep(trace=True)
            assert np.array_equal(tr_g["outcome"], tr_o["outcome"]), f"sweep {sweep}"
            assert np.array_equal(tr_g["dcap"], tr_o["dcap"]) and np.array_equal(tr_g["dccap"], tr_o["dccap"])
            assert np.all(np.abs(tr_g["dfin"] - tr_o["dfin"]) <= cond_tol(1e-12, tr_o["scale"]))
        assert s.energy[0] > 0                    # caps were present and tracked
        q, u, L = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q[d], s.q[d])


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp32_sweep_replayed_within_1e5(R):
    ctx, s, c = make_pair(R=R, precision=32, seed=8)
    with ctx:
        ctx.sweep(2)
        q0, u0, L = ctx.get_state_fixed()
        t = ctx.stats()["sweeps"]
        tr, info = ctx.sweep_traced()
        qf, uf, st = replay_sweep(c, q0, u0, L, s.seed, 0, t, tr, info, check_tol=1e-5, fp32=True)
        assert st["checked"] > 500
        q1, u1, _ = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q1[d], qf[d])
            np.testing.assert_allclose(u1[d], uf[d], rtol=0, atol=1e-6)


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp32_four_wave_build_matches_sequential(R):
    # schedule 2 runs the FP32 move code of schedule 0 compiled for four wavefronts per SIMD:
    # only register allocation and instruction order differ, so at the same wavefront count
    # (which fixes the summation order) the trajectories agree bit for bit
    for w in (1, 2, 4):
        outs = []
        for sched in (0, 2):
            print(f"four-wave build: R {R} W {w} schedule {sched}", flush=True)
            ctx, s, c = make_pair(N=9000, R=R, precision=32, seed=21, pass_schedule=sched,
                                  waves_per_cell=w)
            with ctx:
                ctx.sweep(6)
                outs.append((ctx.get_state_fixed(), ctx.stats()))
        ((q0, u0, _), st0), ((q, u, _), st) = outs
        assert st["accepted"] == st0["accepted"] and st["energy"] == st0["energy"], w
        for d in range(3):
            assert np.array_equal(q[d], q0[d]) and np.array_equal(u[d], u0[d]), w


@pytest.mark.parametrize("precision", [32, 64])
def test_running_energy_vs_fp64_recompute(precision):
    ctx, s, c = make_pair(N=9000, R=5.0, precision=precision, seed=3)
    with ctx:
        ctx.sweep(6)
        assert ctx.stats()["max_drift"] == 0.0      # no periodic re-summation requested
        st = ctx.stats()
        tot = ctx.total_energy()
        assert st["cap_pairs"] == tot[1] and st["cap_colloid"] == tot[2]
        tol = 1e-9 if precision == 64 else 1e-6
        assert abs(st["energy_finite"] - tot[3]) <= tol * (abs(tot[3]) + 9000)
        p2, d = ctx.order_parameter()
        _, u, _ = ctx.get_state_fixed()
        U = np.stack(u, 1)
        Q = U.T @ U / U.shape[0] - np.eye(3) / 3
        assert p2 == pytest.approx(1.5 * np.linalg.eigvalsh(Q)[-1], abs=1e-10)
        e, p2s, rho = ctx.series()
        assert len(e) == 6 and p2s[-1] == pytest.approx(p2, abs=1e-10)
        assert e[-1] == pytest.approx(st["energy"] / 9000, rel=1e-12)


@pytest.mark.parametrize("precision", [32, 64])
def test_stencil_beyond_lds_capacity_uses_global_path(precision):
    # staging capacity forced below the stencil size: every cell takes the global-memory path
    ctx, s, c = make_pair(N=1300, precision=64, seed=12, stage_capacity=64)
    with ctx:
        tr_g, info_g = ctx.sweep_traced()
        tr_o, info_o, _ = s.sweep(trace=True)
        assert np.array_equal(tr_g["outcome"], tr_o["outcome"])
        err = np.abs(tr_g["dfin"] - tr_o["dfin"]) / np.maximum(tr_o["scale"], 1.0)
        assert err.max() <= 1e-12
    if precision == 32:
        ctx, s, c = make_pair(N=1300, precision=32, seed=12, stage_capacity=64)
        with ctx:
            q0, u0, L = ctx.get_state_fixed()
            tr, info = ctx.sweep_traced()
            _, _, st = replay_sweep(c, q0, u0, L, s.seed, 0, 0, tr, info, check_tol=1e-5, fp32=True)
            assert st["checked"] > 500


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp32_overflow_cells_untraced_match_traced(R):
    """Cells beyond the LDS capacity in untraced FP32 sequential sweeps (graph replay): three
    sweeps end bit for bit where three traced sweeps end (one wavefront per cell), and with two
    wavefronts per cell the running energy equals the FP64 recompute."""
    print(f"overflow untraced R={R}", flush=True)
    kw = dict(N=1300, precision=32, seed=12, R=R, stage_capacity=64, pass_schedule=0)
    a, _, _ = make_pair(waves_per_cell=1, **kw)
    b, _, _ = make_pair(waves_per_cell=1, **kw)
    with a, b:
        a.sweep(3)
        for _ in range(3):
            b.sweep_traced()
        qa, ua, _ = a.get_state_fixed()
        qb, ub, _ = b.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qa[d], qb[d]) and np.array_equal(ua[d], ub[d])
        assert a.stats()["accepted"] == b.stats()["accepted"] > 0
    c, _, _ = make_pair(waves_per_cell=2, **kw)
    with c:
        c.sweep(3)
        st = c.stats()
        assert st["attempted"] > 3000 and 0 < st["accepted"] < st["attempted"]
        tot = c.total_energy()
        assert st["cap_pairs"] == tot[1] and st["cap_colloid"] == tot[2]
        assert abs(st["energy_finite"] - tot[3]) <= 1e-6 * (abs(tot[3]) + 1300)


def test_waves_per_cell_do_not_change_fp64_results():
    # the wavefront split only reorders FP64 sums: identical decisions and states
    states = []
    for w in (1, 2, 3, 4, 8):
        q, u, L = layered_system(1500, 0.25, colloid_radius=3.0, seed=41)
        p = m.gb_params(1500, 0.25, 1.0, colloid_radius=3.0, precision=64, seed=6060,
                        energy_check_interval=0, waves_per_cell=w)
        with m.Context(p) as ctx:
            ctx.set_state_fixed(q, u, L)
            ctx.sweep(5)
            states.append(ctx.get_state_fixed())
    for qs, us, _ in states[1:]:
        for d in range(3):
            assert np.array_equal(qs[d], states[0][0][d])
            np.testing.assert_allclose(us[d], states[0][1][d], rtol=0, atol=1e-13)


@pytest.mark.parametrize("R,start", [(0.0, "layered"), (3.0, "layered"), (3.0, "lattice")])
def test_speculative_schedule_matches_oracle_fp64(R, start):
    # untraced sweeps run the speculative pass (moves of a cell evaluated in parallel against
    # the fixed part of the stencil, decided in order): same states as the sequential oracle
    ctx, s, c = make_pair(R=R, precision=64, seed=23, start=start)
    with ctx:
        ctx.sweep(4)
        for _ in range(4):
            s.sweep()
        q, u, _ = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q[d], s.q[d])
            np.testing.assert_allclose(u[d], s.u[d], rtol=0, atol=1e-12)
        st = ctx.stats()
        assert st["attempted"] == s.counters[0] and st["accepted"] == s.counters[1]
        assert st["rejected_left_cell"] == s.counters[2] and st["rejected_hard_core"] == s.counters[3]
        assert st["candidates"] == s.counters[4] and st["pairs"] >= s.counters[5]
        assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]


def test_pass_schedules_agree():
    # sequential, speculative and two-moves schedules: identical FP64 decisions, every
    # wavefront count (schedule 2 is the sequential one in FP64; schedule 3 runs two)
    ref = None
    for sched in (0, 1, 2, 3):
        for w in ((2,) if sched == 3 else (2, 3) if sched != 1 else (1, 2, 3, 4, 8)):
            q, u, L = layered_system(2500, 0.3, colloid_radius=3.0, seed=44)
            p = m.gb_params(2500, 0.3, 1.0, colloid_radius=3.0, precision=64, seed=7070,
                            energy_check_interval=0, waves_per_cell=w, pass_schedule=sched)
            with m.Context(p) as ctx:
                ctx.set_state_fixed(q, u, L)
                ctx.sweep(5)
                got = ctx.get_state_fixed()
                acc = ctx.stats()["accepted"]
            if ref is None:
                ref = (got, acc)
                continue
            assert acc == ref[1], (sched, w)
            for d in range(3):
                assert np.array_equal(got[0][d], ref[0][0][d])
                np.testing.assert_allclose(got[1][d], ref[0][1][d], rtol=0, atol=1e-13)


def test_autotune_keeps_results_and_picks_a_supported_width():
    ctx, s, c = make_pair(N=1300, precision=64, seed=17)
    with ctx:
        best, sched, ms = ctx.autotune(2)
        assert best in (1, 2, 3, 4, 8) and sched in (0, 1, 3)
        run = [k for k in ms if not k.startswith("seq4/") and (not k.startswith("duo/") or k == "duo/2")]
        assert all(ms[k] > 0 for k in run)
        assert all(ms[k] == 0 for k in ms if k not in run)    # seq4: FP32 only; duo: W = 2 only
        # 2 schedules x 5 widths, duo at W = 2; 2 sweeps each; with the sequential schedule kept,
        # three pacing settings of one warm-up and 2 timed sweeps each
        for _ in range((2 * 5 + 1) * 2 + (3 * (1 + 2) if sched == 0 else 0)):
            s.sweep()
        q, u, _ = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q[d], s.q[d])


def test_autotune_fp32_times_every_schedule_and_keeps_the_energy():
    # FP32: the four-wavefront build (schedule 2) is timed too; whichever schedule and width
    # win, the running energy after the tuning sweeps and four more equals the FP64 recompute
    ctx, s, c = make_pair(N=9000, R=5.0, precision=32, seed=19)
    with ctx:
        best, sched, ms = ctx.autotune(2)
        assert best in (1, 2, 3, 4, 8) and sched in (0, 1, 2, 3)
        assert all((v > 0) == (not k.startswith("duo/") or k == "duo/2") for k, v in ms.items())
        ctx.sweep(4)
        st = ctx.stats()
        tot = ctx.total_energy()
        assert st["cap_pairs"] == tot[1] and st["cap_colloid"] == tot[2]
        assert abs(st["energy_finite"] - tot[3]) <= 1e-6 * (abs(tot[3]) + 9000)
        assert st["attempted"] > 34 * 9000            # 30 tuning sweeps and 4 more


def test_graph_replay_matches_single_sweeps():
    # mcgb_sweep(6) replays a captured two-sweep HIP graph; six mcgb_sweep(1) calls launch
    # kernels one by one: identical FP64 trajectories
    q, u, L = layered_system(1500, 0.25, colloid_radius=3.0, seed=44)
    out = []
    for mode in ("graph", "single"):
        p = m.gb_params(1500, 0.25, 1.0, colloid_radius=3.0, precision=64, seed=123,
                        energy_check_interval=4)
        with m.Context(p) as ctx:
            ctx.set_state_fixed(q, u, L)
            if mode == "graph":
                ctx.sweep(6)
            else:
                for _ in range(6):
                    ctx.sweep(1)
            out.append((ctx.get_state_fixed(), ctx.stats()))
    (qa, ua, _), sa = out[0]
    (qb, ub, _), sb = out[1]
    for d in range(3):
        assert np.array_equal(qa[d], qb[d])
        assert np.array_equal(ua[d], ub[d])
    assert sa["accepted"] == sb["accepted"] and sa["energy"] == sb["energy"]
    assert sa["resummations"] == sb["resummations"] == 1
