"""The 27-sublattice decomposition (sublattice_colors 3: cells >= r_c/2, partners within +-2
cells, 27 passes per sweep) on the GPU against the FP64 oracle running the same decomposition:
cell lists and orders bit-exact, FP64 move-by-move identity, FP32 replay within 1e-5, untraced
(graph) sweeps, the HBM path, energies, NPT, replicas, and the same equilibrium as 8 sublattices."""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.replay import replay_sweep
from tests.systems import cond_tol, layered_system
from tests.test_gpu_parity import make_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


@pytest.mark.parametrize("N,R", [(1300, 0.0), (9000, 5.0)])
def test_cell_list_and_order_bit_exact(N, R):
    print(f"colors3 cell list N={N} R={R}", flush=True)
    ctx, s, c = make_pair(N=N, R=R, colors=3)
    with ctx:
        for t in (0, 5, 777):
            cs, ids, shift, order, dims = ctx.cell_list(t)
            g, o, _ = oc.sweep_setup(s.seed, 0, t, 3)
            assert list(shift) == list(g) and list(order) == list(o) and len(order) == 27
            n = [oc.grid_dim(s.L[d], c.cutoff, 3) for d in range(3)]
            assert list(dims) == n and all(x % 3 == 0 for x in n)
            cs_o, so_o = oc.cell_list(s.q, np.array(n, dtype=np.int32), g)
            assert np.array_equal(cs, cs_o) and np.array_equal(ids, so_o)


@pytest.mark.parametrize("R,start", [(0.0, "layered"), (3.0, "layered"), (3.0, "lattice")])
def test_fp64_sweeps_match_oracle_move_by_move(R, start):
    print(f"colors3 fp64 traced R={R} {start}", flush=True)
    ctx, s, c = make_pair(R=R, precision=64, colors=3, start=start, rho=0.28)
    with ctx:
        for sweep in range(3):
            tr_g, info_g = ctx.sweep_traced()
            tr_o, info_o, _ = s.sweep(trace=True)
            assert np.array_equal(info_g, info_o) and info_g[5] == 27
            assert np.array_equal(tr_g["pid"], tr_o["pid"])
            assert np.array_equal(tr_g["outcome"], tr_o["outcome"]), f"sweep {sweep}"
            assert np.array_equal(tr_g["dcap"], tr_o["dcap"]) and np.array_equal(tr_g["dccap"], tr_o["dccap"])
            assert np.all(np.abs(tr_g["dfin"] - tr_o["dfin"]) <= cond_tol(1e-12, tr_o["scale"]))
            assert (tr_g["outcome"] == 4).sum() > 100
        q, u, L = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q[d], s.q[d])
            np.testing.assert_allclose(u[d], s.u[d], rtol=0, atol=1e-12)
        st = ctx.stats()
        assert st["attempted"] == s.counters[0] and st["accepted"] == s.counters[1]
        assert st["candidates"] == s.counters[4] and st["pairs"] == s.counters[5]
        assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]


@pytest.mark.parametrize("R", [0.0, 3.0])
def test_fp32_sweep_replayed_within_1e5(R):
    ctx, s, c = make_pair(R=R, precision=32, seed=8, colors=3)
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


@pytest.mark.parametrize("waves", [1, 2, 4])
def test_untraced_sweeps_match_oracle_fp64(waves):
    ctx, s, c = make_pair(R=3.0, precision=64, seed=23, colors=3, waves_per_cell=waves,
                          pass_schedule=1)
    with ctx:
        ctx.sweep(4)                       # graph replay; the speculative schedule is not used
        for _ in range(4):
            s.sweep()
        q, u, _ = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q[d], s.q[d]), f"W={waves}"
            np.testing.assert_allclose(u[d], s.u[d], rtol=0, atol=1e-12)


def test_stencil_beyond_lds_capacity_uses_global_path():
    ctx, s, c = make_pair(N=1300, precision=64, seed=12, stage_capacity=32, colors=3)
    with ctx:
        tr_g, _ = ctx.sweep_traced()
        tr_o, _, _ = s.sweep(trace=True)
        assert np.array_equal(tr_g["outcome"], tr_o["outcome"])
        err = np.abs(tr_g["dfin"] - tr_o["dfin"]) / np.maximum(tr_o["scale"], 1.0)
        assert err.max() <= 1e-12


@pytest.mark.parametrize("precision", [32, 64])
def test_energies_on_the_fine_grid(precision):
    # initial energy, running energy after sweeps and the periodic re-summation all use the
    # +-2-cell stencil of the 27-sublattice grid
    ctx, s, c = make_pair(N=9000, R=5.0, precision=precision, seed=3, colors=3, resum=3)
    with ctx:
        st0 = ctx.stats()
        assert st0["cap_pairs"] == s.energy[0]
        assert abs(st0["energy_finite"] - s.energy[2]) <= 1e-9 * (abs(s.energy[2]) + 9000)
        ctx.sweep(7)
        st = ctx.stats()
        assert st["resummations"] >= 2
        tot = ctx.total_energy()
        assert st["cap_pairs"] == tot[1] and st["cap_colloid"] == tot[2]
        tol = 1e-9 if precision == 64 else 1e-6
        assert abs(st["energy_finite"] - tot[3]) <= tol * (abs(tot[3]) + 9000)
        q, u, L = ctx.get_state_fixed()
        so = oc.System(c, q, u, L)
        assert so.energy[0] == tot[1] and abs(so.energy[2] - tot[3]) <= 1e-9 * (abs(tot[3]) + 9000)


def test_npt_matches_oracle_fp64():
    N, rho, P = 4000, 0.28, 2.0
    q, u, L = layered_system(N, rho, seed=19)
    p = m.gb_params(N, rho, 1.0, precision=64, seed=515, npt=True, pressure=P,
                    energy_check_interval=0, sublattice_colors=3)
    c = oc.consts(temperature=1.0, pressure=P, colors=3)
    s = oc.System(c, q, u, L, seed=515, npt=True, nthreads=4)
    with m.Context(p) as ctx:
        ctx.set_state_fixed(q, u, L)
        ctx.sweep(6)
        for _ in range(6):
            s.sweep()
        qg, ug, Lg = ctx.get_state_fixed()
        assert np.allclose(Lg, s.L, rtol=1e-14, atol=0)
        for d in range(3):
            assert np.array_equal(qg[d], s.q[d])
        st = ctx.stats()
        assert st["volume_attempted"] == s.counters[7] == 6
        assert st["volume_accepted"] == s.counters[8] and st["accepted"] == s.counters[1]


def test_replicas_match_oracle():
    seed = 4242
    specs = [(1300, 0.25, 0.9, 0.0), (2000, 0.28, 1.3, 3.0)]
    params, states, oracles = [], [], []
    for i, (N, rho, T, R) in enumerate(specs):
        q, u, L = layered_system(N, rho, colloid_radius=R, seed=60 + i)
        params.append(m.gb_params(N, rho, T, colloid_radius=R, precision=64, seed=seed,
                                  energy_check_interval=0, sublattice_colors=3))
        states.append((q, u, L))
        c = oc.consts(temperature=T, colloid_radius=R, colloid_on=1 if R > 0 else 0, colors=3)
        oracles.append(oc.System(c, q, u, L, seed=seed, replica=i, nthreads=4))
    with m.Context(params) as ctx:
        for i, (q, u, L) in enumerate(states):
            ctx.set_state_fixed(q, u, L, replica=i)
        ctx.sweep(3)
        for i, s in enumerate(oracles):
            for _ in range(3):
                s.sweep()
            qg, _, _ = ctx.get_state_fixed(replica=i)
            for d in range(3):
                assert np.array_equal(qg[d], s.q[d]), f"replica {i}"


def test_mixed_colours_in_one_batch_are_rejected():
    p2 = m.gb_params(1300, 0.25, 1.0)
    p3 = m.gb_params(1300, 0.25, 1.0, sublattice_colors=3)
    with pytest.raises(m.McgbError):
        m.Context([p2, p3])
    with pytest.raises(m.McgbError):
        m.Context(m.gb_params(1300, 0.25, 1.0, sublattice_colors=4))


def test_same_equilibrium_as_eight_sublattices():
    """FP32 NVT at rho* = 0.30, T* = 1.2, N = 8000: <E>/N and P2 of the two decompositions agree
    within 4 combined standard errors (tau_int-corrected)."""
    res = {}
    for col in (2, 3):
        q, u, L = layered_system(8000, 0.30, seed=71)
        p = m.gb_params(8000, 0.30, 1.2, precision=32, seed=999, sublattice_colors=col)
        with m.Context(p) as ctx:
            ctx.set_state_fixed(q, u, L)
            ctx.sweep(300)
            ctx.series_clear()
            ctx.sweep(1500)
            e, p2, _ = ctx.series()
        res[col] = []
        for x in (e, p2):
            tau, _ = m.tau_int(x)
            res[col].append((x.mean(), x.std() * np.sqrt(tau / len(x))))
    for k in range(2):
        (m2, s2), (m3, s3) = res[2][k], res[3][k]
        assert abs(m2 - m3) < 4.0 * np.hypot(s2, s3) + 1e-9, (k, res)
