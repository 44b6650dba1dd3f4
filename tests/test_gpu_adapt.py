"""Step-size adaptation on the GPU (the controller of MCnptColloid.c:987-1031, run on the device
after every sweep) against the oracle running the same controller: the step sizes after every
sweep equal the oracle's exactly in FP64, through both 5 % branches and both clamps, in single
contexts and in a replica batch; the state stays bit-exact."""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.systems import layered_system

pytestmark = pytest.mark.gpu

# (start d_max, start d_rot, T*, rho*): each case reaches a different branch within 12 sweeps
CASES = {"shrink": (0.8, 0.1, 1.0, 0.25), "grow": (0.02, 0.1, 1.0, 0.25),
         "upper_clamp": (1.45, 0.2, 50.0, 0.05), "lower_clamp": (1.04e-4, 0.6, 1.0, 0.25)}


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def _pair(case, precision=64, N=1300, seed=5, adapt_moves=1000):
    d0, r0, T, rho = CASES[case]
    q, u, L = layered_system(N, rho, seed=seed)
    p = m.gb_params(N, rho, T, precision=precision, seed=900 + seed, energy_check_interval=0,
                    max_displacement=d0, max_rotation=r0, adapt_moves=adapt_moves)
    c = oc.consts(temperature=T, max_displacement=d0, max_rotation=r0)
    s = oc.System(c, q, u, L, seed=900 + seed, nthreads=4, adapt_moves=adapt_moves)
    ctx = m.Context(p)
    ctx.set_state_fixed(q, u, L)
    return ctx, s


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_step_sizes_follow_oracle_every_sweep(case):
    print(f"adapt case {case}", flush=True)
    ctx, s = _pair(case)
    with ctx:
        start = ctx.step_sizes()
        assert start[0] == np.float32(CASES[case][0]) and start[2] == 0
        drots = set()
        for k in range(12):
            ctx.sweep(1)
            s.sweep()
            d, r, n = ctx.step_sizes()
            assert (d, r) == (s.dmax, s.drot), f"sweep {k}: gpu {(d, r)} oracle {(s.dmax, s.drot)}"
            assert n == s.adapts == k + 1
            drots.add(float(r))
        q, u, _ = ctx.get_state_fixed()
        for dim in range(3):
            assert np.array_equal(q[dim], s.q[dim])
            np.testing.assert_allclose(u[dim], s.u[dim], rtol=0, atol=1e-12)
        if case.endswith("clamp"):
            assert len(drots) == 1 and drots != {CASES[case][1]}   # the clamp moved d_rot once
        # graph replay of several sweeps runs the same controller
        ctx.sweep(4)
        for _ in range(4):
            s.sweep()
        d, r, n = ctx.step_sizes()
        assert (d, r, n) == (s.dmax, s.drot, s.adapts)


def test_adaptation_threshold_spans_sweeps():
    # 1300 moves a sweep, threshold 3000: the controller acts after sweeps 3, 6, 9
    ctx, s = _pair("shrink", adapt_moves=3000)
    with ctx:
        for k in range(9):
            ctx.sweep(1)
            s.sweep()
            d, r, n = ctx.step_sizes()
            assert (d, r, n) == (s.dmax, s.drot, s.adapts)
            assert n == (k + 1) // 3


def test_replica_batch_adapts_each_replica():
    seed = 6160
    params, oracles = [], []
    for i, case in enumerate(("shrink", "grow", "upper_clamp")):
        d0, r0, T, rho = CASES[case]
        q, u, L = layered_system(1300, rho, seed=40 + i)
        params.append(m.gb_params(1300, rho, T, precision=64, seed=seed, energy_check_interval=0,
                                  max_displacement=d0, max_rotation=r0, adapt_moves=1000))
        c = oc.consts(temperature=T, max_displacement=d0, max_rotation=r0)
        oracles.append((oc.System(c, q, u, L, seed=seed, replica=i, nthreads=4,
                                  adapt_moves=1000), (q, u, L)))
    with m.Context(params) as ctx:
        for i, (_, (q, u, L)) in enumerate(oracles):
            ctx.set_state_fixed(q, u, L, replica=i)
        ctx.sweep(6)
        for i, (s, _) in enumerate(oracles):
            for _ in range(6):
                s.sweep()
            assert ctx.step_sizes(replica=i) == (s.dmax, s.drot, s.adapts), f"replica {i}"
            qg, _, _ = ctx.get_state_fixed(replica=i)
            assert np.array_equal(qg[0], s.q[0])


def test_off_by_default_and_manual_step_sizes():
    ctx, s = _pair("shrink", adapt_moves=0)
    with ctx:
        ctx.sweep(3)
        for _ in range(3):
            s.sweep()
        assert ctx.step_sizes() == (np.float32(0.8), np.float32(0.1), 0)
        ctx.set_step_sizes(0.25, 0.05)
        assert ctx.step_sizes()[:2] == (np.float32(0.25), np.float32(0.05))
        s.c.max_displacement, s.c.max_rotation = float(np.float32(0.25)), float(np.float32(0.05))
        ctx.sweep(3)
        for _ in range(3):
            s.sweep()
        q, _, L = ctx.get_state_fixed()
        for dim in range(3):                       # the later sweeps used the new sizes
            assert np.array_equal(q[dim], s.q[dim])
        with pytest.raises(m.McgbError):
            ctx.set_step_sizes(0.6 * L[0], 0.1)
        with pytest.raises(m.McgbError):
            ctx.set_step_sizes(-1.0, 0.1)
        assert ctx.step_sizes()[:2] == (np.float32(0.25), np.float32(0.05))


def test_fp32_acceptance_settles_in_the_window():
    ctx, _ = _pair("shrink", precision=32, N=8000, seed=9)
    with ctx:
        accs = []
        for _ in range(40):
            a0 = ctx.stats()
            ctx.sweep(1)
            a1 = ctx.stats()
            accs.append((a1["accepted"] - a0["accepted"]) / (a1["attempted"] - a0["attempted"]))
        d, r, n = ctx.step_sizes()
        assert accs[0] < 0.3 and n == 40 and d < 0.3
        assert 0.3 <= np.mean(accs[-8:]) <= 0.55
