"""Pass schedule 3 (two moves of a cell in flight, one per wavefront) against the oracle: the
second move of a step speculates on the first one's outcome through the single pair between the
two moved particles, so every decision, cap count and state must still equal the oracle's
sequential chain (FP64, move by move), FP32 moves must replay within 1e-5, and untraced sweeps
must land where traced ones do."""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.replay import replay_sweep
from tests.systems import layered_system
from tests.test_gpu_parity import cond_tol, make_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def check_traced(ctx, s, sweeps, tol=1e-12, cond=False, label=""):
    for sweep in range(sweeps):
        print(f"two-moves schedule {label}: sweep {sweep}", flush=True)
        tr_g, info_g = ctx.sweep_traced()
        tr_o, info_o, _ = s.sweep(trace=True)
        assert np.array_equal(info_g, info_o)
        assert np.array_equal(tr_g["pid"], tr_o["pid"])
        assert np.array_equal(tr_g["outcome"], tr_o["outcome"]), f"sweep {sweep}"
        assert np.array_equal(tr_g["dcap"], tr_o["dcap"]) and np.array_equal(tr_g["dccap"], tr_o["dccap"])
        if cond:
            assert np.all(np.abs(tr_g["dfin"] - tr_o["dfin"]) <= cond_tol(tol, tr_o["scale"]))
        else:
            err = np.abs(tr_g["dfin"] - tr_o["dfin"]) / np.maximum(tr_o["scale"], 1.0)
            assert err.max() <= tol
    q, u, _ = ctx.get_state_fixed()
    for d in range(3):
        assert np.array_equal(q[d], s.q[d])
        np.testing.assert_allclose(u[d], s.u[d], rtol=0, atol=1e-12)


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp64_two_moves_in_flight_match_oracle_move_by_move(R):
    ctx, s, c = make_pair(R=R, precision=64, pass_schedule=3)
    with ctx:
        check_traced(ctx, s, 4, label=f"R {R}")
        st = ctx.stats()
        assert st["attempted"] == s.counters[0] and st["accepted"] == s.counters[1]
        assert st["rejected_left_cell"] == s.counters[2] and st["rejected_hard_core"] == s.counters[3]
        assert st["candidates"] == s.counters[4] and st["pairs"] == s.counters[5]
        assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]
        assert st["energy_finite"] == pytest.approx(s.energy[2], rel=1e-10, abs=1e-6)


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp64_two_moves_in_flight_lattice_start_with_caps(R):
    # capped and near-cap pairs between the two moved particles of a step: the routed cap
    # counts and finite sums must give the oracle's decisions
    ctx, s, c = make_pair(R=R, rho=0.3, precision=64, start="lattice", pass_schedule=3)
    with ctx:
        check_traced(ctx, s, 3, cond=True, label=f"lattice R {R}")
        assert s.energy[0] > 0
        assert ctx.stats()["cap_pairs"] == s.energy[0]


def test_fp64_two_moves_in_flight_dense_cells_and_27_sublattices():
    # cutoff 6: ~90 particles per cell (colloid terms outside the LDS cache, stencils beyond the
    # staging capacity on the HBM pass); then the 27-sublattice grid
    ctx, s, c = make_pair(N=6000, rho=0.3, R=3.0, precision=64, seed=31, cutoff=6.0,
                          pass_schedule=3)
    with ctx:
        check_traced(ctx, s, 2, label="dense")
    ctx, s, c = make_pair(N=3000, rho=0.3, R=3.0, precision=64, seed=33, colors=3,
                          pass_schedule=3)
    with ctx:
        check_traced(ctx, s, 3, label="27 sublattices")


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp32_two_moves_in_flight_replayed_within_1e5(R):
    ctx, s, c = make_pair(R=R, precision=32, seed=8, pass_schedule=3)
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


@pytest.mark.parametrize("precision", [32, 64])
def test_two_moves_in_flight_untraced_equal_traced_and_sequential(precision):
    # untraced sweeps (graph replay) run the same kernel as traced ones: bit for bit; in FP64
    # the sequential schedule reaches the same state (decisions equal up to summation order)
    print(f"two-moves untraced vs traced, FP{precision}", flush=True)
    kw = dict(N=9000, R=5.0, precision=precision, seed=23)
    a, _, _ = make_pair(pass_schedule=3, **kw)
    b, _, _ = make_pair(pass_schedule=3, **kw)
    z, _, _ = make_pair(pass_schedule=0, waves_per_cell=2, **kw)
    with a, b, z:
        a.sweep(4)
        for _ in range(4):
            b.sweep_traced()
        z.sweep(4)
        qa, ua, _ = a.get_state_fixed()
        qb, ub, _ = b.get_state_fixed()
        qz, uz, _ = z.get_state_fixed()
        sa, sb, sz = a.stats(), b.stats(), z.stats()
        assert sa["accepted"] == sb["accepted"] > 0 and sa["energy"] == sb["energy"]
        for d in range(3):
            assert np.array_equal(qa[d], qb[d]) and np.array_equal(ua[d], ub[d])
        if precision == 64:
            assert sa["accepted"] == sz["accepted"]
            for d in range(3):
                assert np.array_equal(qa[d], qz[d])
                np.testing.assert_allclose(ua[d], uz[d], rtol=0, atol=1e-13)
        # the running energy against the FP64 recompute
        tot = a.total_energy()
        assert sa["cap_pairs"] == tot[1] and sa["cap_colloid"] == tot[2]
        tol = 1e-9 if precision == 64 else 1e-6
        assert abs(sa["energy_finite"] - tot[3]) <= tol * (abs(tot[3]) + 9000)


def test_two_moves_in_flight_replicas_match_the_oracle():
    # replicas of uneven size and temperature share the launches; each follows its own oracle
    # chain (key seed + replica)
    print("two-moves replicas", flush=True)
    seed, rho = 61, 0.28
    sizes = [(1500, 0.0, 1.0), (2600, 3.0, 1.1), (1300, 0.0, 1.2)]
    ps, sys_ = [], []
    for i, (N, R, T) in enumerate(sizes):
        q, u, L = layered_system(N, rho, colloid_radius=R, seed=70 + i)
        ps.append(m.gb_params(N, rho, T, colloid_radius=R, precision=64, seed=seed,
                              energy_check_interval=0, pass_schedule=3))
        c = oc.consts(temperature=T, colloid_radius=R, colloid_on=1 if R > 0 else 0)
        sys_.append((q, u, L, oc.System(c, q, u, L, seed=seed, replica=i, nthreads=4)))
    with m.Context(ps) as ctx:
        for i, (q, u, L, _) in enumerate(sys_):
            ctx.set_state_fixed(q, u, L, replica=i)
        ctx.sweep(3)
        for i, (_, _, _, s) in enumerate(sys_):
            for _ in range(3):
                s.sweep()
            qg, ug, _ = ctx.get_state_fixed(replica=i)
            for d in range(3):
                assert np.array_equal(qg[d], s.q[d]), f"replica {i}"
                np.testing.assert_allclose(ug[d], s.u[d], rtol=0, atol=1e-12)
            assert ctx.stats(replica=i)["accepted"] == s.counters[1]
