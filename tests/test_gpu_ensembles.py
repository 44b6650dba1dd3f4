"""NPT ln V moves and batched replicas on the GPU against the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.systems import layered_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def test_npt_fp64_matches_oracle():
    N, rho, T, P = 1400, 0.25, 1.0, 2.0
    q, u, L = layered_system(N, rho, seed=21)
    p = m.gb_params(N, rho, T, precision=64, seed=4242, npt=True, pressure=P,
                    npt_max_dlnv=0.01, energy_check_interval=0)
    c = oc.consts(temperature=T, pressure=P)
    s = oc.System(c, q, u, L, seed=4242, npt=True, npt_max_dlnv=0.01, nthreads=4)
    with m.Context(p) as ctx:
        ctx.set_state_fixed(q, u, L)
        for sweep in range(8):
            ctx.sweep(1)
            s.sweep()
            st = ctx.stats()
            assert st["volume_attempted"] == s.counters[7]
            assert st["volume_accepted"] == s.counters[8], f"sweep {sweep}"
            np.testing.assert_allclose(st["box"], s.L, rtol=1e-15)
        assert s.counters[8] >= 1               # at least one volume move accepted
        qg, ug, Lg = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qg[d], s.q[d])
        assert st["cap_pairs"] == s.energy[0]
        assert st["energy_finite"] == pytest.approx(s.energy[2], rel=1e-10, abs=1e-6)
        e, p2, dens = ctx.series()
        assert len(dens) == 8
        assert dens[-1] == pytest.approx(N / np.prod(s.L), rel=1e-13)


def test_npt_ideal_gas_volume_gpu():
    # eps0 = 0 and tiny sigma0: U = 0, so <V> = (N+1)/(beta P) for ln V sampling
    N, T, target = 200, 1.0, 8000.0
    P = (N + 1) / target
    p = m.gb_params(N, N / target, T, precision=64, seed=77, npt=True, pressure=P,
                    npt_max_dlnv=0.08, cutoff=3.0, max_displacement=0.5, energy_check_interval=0)
    p.gb_epsilon0 = 0.0
    p.gb_sigma_sidetoside = 0.01
    rng = np.random.default_rng(3)
    L = np.array([20.0] * 3)
    pos = [rng.uniform(0, 20.0, size=N) for _ in range(3)]
    u = rng.normal(size=(3, N))
    with m.Context(p) as ctx:
        ctx.set_state(pos, list(u), L)
        ctx.sweep(6000)
        _, _, rho = ctx.series()
    V = N / rho[500:]
    tau, _ = m.tau_int(V)
    err = V.std() * np.sqrt(tau / len(V))
    assert abs(V.mean() - target) < 4 * err + 0.005 * target


def test_replicas_match_oracle_and_single_contexts():
    seed = 5150
    specs = [(1300, 0.25, 0.8, 0.0), (1500, 0.25, 1.2, 3.0), (1400, 0.25, 1.6, 0.0)]
    params, states, oracles = [], [], []
    for i, (N, rho, T, R) in enumerate(specs):
        q, u, L = layered_system(N, rho, colloid_radius=R, seed=30 + i)
        params.append(m.gb_params(N, rho, T, colloid_radius=R, precision=64, seed=seed,
                                  energy_check_interval=0))
        states.append((q, u, L))
        c = oc.consts(temperature=T, colloid_radius=R, colloid_on=1 if R > 0 else 0)
        oracles.append(oc.System(c, q, u, L, seed=seed, replica=i, nthreads=4))
    with m.Context(params) as ctx:
        for i, (q, u, L) in enumerate(states):
            ctx.set_state_fixed(q, u, L, replica=i)
        ctx.sweep(3)
        for i, s in enumerate(oracles):
            for _ in range(3):
                s.sweep()
            qg, ug, _ = ctx.get_state_fixed(replica=i)
            for d in range(3):
                assert np.array_equal(qg[d], s.q[d]), f"replica {i}"
                np.testing.assert_allclose(ug[d], s.u[d], rtol=0, atol=1e-12)
            st = ctx.stats(replica=i)
            assert st["accepted"] == s.counters[1] and st["attempted"] == s.counters[0]
    # replica i of a batch equals a single-replica context keyed seed + i * 2^32
    N, rho, T, R = specs[1]
    q, u, L = states[1]
    p1 = m.gb_params(N, rho, T, colloid_radius=R, precision=64, seed=seed + (1 << 32),
                     energy_check_interval=0)
    with m.Context(p1) as ctx1:
        ctx1.set_state_fixed(q, u, L)
        ctx1.sweep(3)
        qg, _, _ = ctx1.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qg[d], oracles[1].q[d])
