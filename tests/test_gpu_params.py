"""Sweeps under Gay-Berne parameters other than the benchmark's GB(3, 5, 2, 1), sigma0 = eps0 = 1
(the reference reads sigma_ss, eps0, kappa, kappa', mu, nu and the cutoff from its parameter
file, MCnptColloid.c:4092-4127): every constant the pass kernels use must follow them, in the
pair (FP64 and the FP32 eps exponent cases), the colloid term, the pre-filter radii and the
cell grid.

* GB(3, 5, 1, 2) with sigma0 = 1.2, eps0 = 0.7 and an R = 3 colloid;
* GB(4.4, 20, 1, 1) (long rods) with a cutoff of 5.

FP64: two traced sweeps move by move and two untraced sweeps against the oracle.  FP32: a
traced sweep replayed against the brute-force restatement within 1e-5, and the running energy
of the two-wavefront kernel equal to the FP64 recompute after sweeps."""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.replay import replay_sweep
from tests.systems import layered_system

pytestmark = pytest.mark.gpu

N, RHO, SEED = 1300, 0.15, 20261208
# name: (sigma0, eps0, kappa, kappa', mu, nu, cutoff, colloid R, layer spacing of the start)
SETS = {"gb3512_s12_e07": (1.2, 0.7, 3.0, 5.0, 1.0, 2.0, 4.0, 3.0, 3.9),
        "gb44_20_11_rc5": (1.0, 1.0, 4.4, 20.0, 1.0, 1.0, 5.0, 0.0, 4.8)}


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def pair(name, precision, **kw):
    s0, e0, k, kp, mu, nu, rc, R, layer = SETS[name]
    p = m.gb_params(N, RHO, 1.0, colloid_radius=R, precision=precision, seed=SEED, kappa=k,
                    kappa_prime=kp, mu=mu, nu=nu, cutoff=rc, energy_check_interval=0, **kw)
    p.gb_sigma_sidetoside, p.gb_epsilon0 = s0, e0
    c = oc.consts(sigma0=s0, epsilon0=e0, kappa=k, kappa_prime=kp, mu=mu, nu=nu, cutoff=rc,
                  temperature=1.0, colloid_radius=R, colloid_on=1 if R > 0 else 0)
    q, u, L = layered_system(N, RHO, colloid_radius=R, seed=SEED, layer=layer)
    ctx = m.Context(p)
    ctx.set_state_fixed(q, u, L)
    return ctx, c, (q, u, L)


@pytest.mark.parametrize("name", sorted(SETS))
def test_fp64_sweeps_match_oracle(name):
    print(f"parameters {name}: FP64", flush=True)
    ctx, c, (q, u, L) = pair(name, 64)
    s = oc.System(c, q, u, L, seed=SEED, nthreads=8)
    with ctx:
        tot = ctx.total_energy()
        e = s.total_energy()
        assert tot[1] == e[0] and tot[2] == e[1]
        assert tot[3] == pytest.approx(e[2], rel=1e-12, abs=1e-9)
        for sweep in range(2):
            tr_g, info_g = ctx.sweep_traced()
            tr_o, info_o, _ = s.sweep(trace=True)
            assert np.array_equal(info_g, info_o)
            assert np.array_equal(tr_g["outcome"], tr_o["outcome"]), f"sweep {sweep}"
            assert np.array_equal(tr_g["dcap"], tr_o["dcap"])
            assert np.array_equal(tr_g["dccap"], tr_o["dccap"])
            err = np.abs(tr_g["dfin"] - tr_o["dfin"]) / np.maximum(tr_o["scale"], 1.0)
            assert err.max() <= 1e-12
            assert (tr_g["outcome"] == 4).sum() > 100 and (tr_g["outcome"] == 3).sum() > 100
        ctx.sweep(2)
        for _ in range(2):
            s.sweep()
        qg, ug, _ = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qg[d], s.q[d])
            np.testing.assert_allclose(ug[d], s.u[d], rtol=0, atol=1e-12)
        st = ctx.stats()
        assert st["attempted"] == s.counters[0] and st["accepted"] == s.counters[1]
        assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]
        assert st["energy_finite"] == pytest.approx(s.energy[2], rel=1e-10, abs=1e-6)


@pytest.mark.parametrize("name", sorted(SETS))
def test_fp32_sweep_replayed_and_running_energy(name):
    print(f"parameters {name}: FP32", flush=True)
    ctx, c, _ = pair(name, 32, pass_schedule=0, waves_per_cell=2)
    with ctx:
        ctx.sweep(2)
        q0, u0, L = ctx.get_state_fixed()
        t = ctx.stats()["sweeps"]
        tr, info = ctx.sweep_traced()
        qf, uf, st = replay_sweep(c, q0, u0, L, SEED, 0, t, tr, info, check_tol=1e-5, fp32=True)
        assert st["checked"] > 500
        q1, u1, _ = ctx.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q1[d], qf[d])
            np.testing.assert_allclose(u1[d], uf[d], rtol=0, atol=1e-6)
        ctx.sweep(4)
        s = ctx.stats()
        tot = ctx.total_energy()
        assert s["cap_pairs"] == tot[1] and s["cap_colloid"] == tot[2]
        assert abs(s["energy_finite"] - tot[3]) <= 1e-6 * (abs(tot[3]) + N)
