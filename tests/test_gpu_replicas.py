"""The 32-replica batch at its full size (the benchmark's replica workload: N log-uniform in
[2000, 32000] from a seed, T* = 0.8 .. 1.6, an R = 3 colloid on every fourth replica, 27
sublattices): every replica shares the pass launches, so each must still follow its own chain.
FP64: every replica's fixed-point state, orientations and counters equal its oracle chain after
two sweeps.  FP32: after 20 sweeps (graph replays) every replica's running energy equals the
FP64 recompute of its state (cap counts exactly)."""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.systems import layered_system

pytestmark = pytest.mark.gpu

SEED = 20261205


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def batch(precision):
    rng = np.random.default_rng(SEED)
    Ns = np.round(np.exp(rng.uniform(np.log(2000), np.log(32000), 32))).astype(int)
    params, states, specs = [], [], []
    for i in range(32):
        R = 3.0 if i % 4 == 0 else 0.0
        T = 0.8 + i * 0.8 / 31
        params.append(m.gb_params(int(Ns[i]), 0.30, T, colloid_radius=R, precision=precision,
                                  seed=SEED, sublattice_colors=3, energy_check_interval=0))
        states.append(layered_system(int(Ns[i]), 0.30, colloid_radius=R, seed=SEED + i))
        specs.append((int(Ns[i]), T, R))
    return params, states, specs


def test_fp64_full_replica_batch_matches_the_oracle_per_replica():
    params, states, specs = batch(64)
    assert sum(n for n, _, _ in specs) > 200000
    with m.Context(params) as ctx:
        for i, (q, u, L) in enumerate(states):
            ctx.set_state_fixed(q, u, L, replica=i)
        print("32 replicas: 2 FP64 sweeps on the GPU", flush=True)
        ctx.sweep(2)
        for i, ((q, u, L), (N, T, R)) in enumerate(zip(states, specs)):
            c = oc.consts(temperature=T, colloid_radius=R, colloid_on=1 if R > 0 else 0, colors=3)
            s = oc.System(c, q, u, L, seed=SEED, replica=i, nthreads=16)
            for _ in range(2):
                s.sweep()
            qg, ug, _ = ctx.get_state_fixed(replica=i)
            for d in range(3):
                assert np.array_equal(qg[d], s.q[d]), f"replica {i} (N {N}, T {T:.3f}, R {R})"
                np.testing.assert_allclose(ug[d], s.u[d], rtol=0, atol=1e-12)
            st = ctx.stats(replica=i)
            assert st["attempted"] == s.counters[0] and st["accepted"] == s.counters[1]
            assert st["cap_pairs"] == s.energy[0] and st["cap_colloid"] == s.energy[1]
            assert st["energy_finite"] == pytest.approx(s.energy[2], rel=1e-10, abs=1e-6)


def test_fp32_full_replica_batch_running_energies_equal_the_fp64_recompute():
    params, states, specs = batch(32)
    with m.Context(params) as ctx:
        for i, (q, u, L) in enumerate(states):
            ctx.set_state_fixed(q, u, L, replica=i)
        print("32 replicas: 20 FP32 sweeps on the GPU", flush=True)
        ctx.sweep(20)
        for i, (N, T, R) in enumerate(specs):
            st = ctx.stats(replica=i)
            # moves drawn in empty cells of the layered start are not attempts
            assert st["sweeps"] == 20 and st["attempted"] > 5 * N
            tot = ctx.total_energy(replica=i)
            assert st["cap_pairs"] == tot[1] and st["cap_colloid"] == tot[2]
            # FP32 pair terms summed in FP64 per move: each accepted move adds its rounding,
            # a few 1e-7 of its terms; bounded here by 1e-6 of (|E| + N)
            assert abs(st["energy_finite"] - tot[3]) <= 1e-6 * (abs(tot[3]) + N), f"replica {i}"
