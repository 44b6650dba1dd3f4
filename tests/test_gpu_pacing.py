"""Pacing of the sequential pass (mcgb_set_pace) changes only the issue priority of the cells'
wavefronts, never the chain: from one state, sweeps with pacing off, at the default and at a
setting far below the cells' mean move time (most cells behind, priorities changing all the
time) give the same states, counters and running energies bit for bit.  The box is the
benchmark's geometry (rho* = 0.32, colloid R = 5) at sizes where the pass holds more
wavefronts than the SIMDs can run side by side."""
import numpy as np
import pytest

from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.systems import layered_system

pytestmark = pytest.mark.gpu

SEED = 20261206


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def run(pace, precision, N, sweeps=6):
    p = m.gb_params(N, 0.32, 1.0, colloid_radius=5.0, precision=precision, seed=SEED,
                    waves_per_cell=2, pass_schedule=0, energy_check_interval=0)
    q, u, L = layered_system(N, 0.32, colloid_radius=5.0, seed=SEED)
    with m.Context(p) as ctx:
        ctx.set_state_fixed(q, u, L)
        ctx.set_pace(pace)
        assert ctx.pace() == pace
        ctx.sweep(sweeps)
        qg, ug, _ = ctx.get_state_fixed()
        return qg, ug, ctx.stats()


@pytest.mark.parametrize("precision,N", [(32, 256000), (64, 64000)])
def test_pacing_leaves_the_chain_unchanged(precision, N):
    print(f"pacing: FP{precision} N {N}, pace 0 / 109 / 40", flush=True)
    q0, u0, s0 = run(0, precision, N)
    assert s0["attempted"] > 3 * N
    for pace in (109, 40):
        q, u, s = run(pace, precision, N)
        for d in range(3):
            assert np.array_equal(q[d], q0[d]), f"pace {pace}"
            assert np.array_equal(u[d], u0[d]), f"pace {pace}"
        for k in ("attempted", "accepted", "cap_pairs", "cap_colloid", "energy_finite", "pairs"):
            assert s[k] == s0[k], f"pace {pace}: {k}"


def test_pace_setting_is_checked():
    p = m.gb_params(4000, 0.32, 1.0, precision=32, seed=SEED)
    with m.Context(p) as ctx:
        assert ctx.pace() == 109
        with pytest.raises(Exception):
            ctx.set_pace(-1)
        with pytest.raises(Exception):
            ctx.set_pace(1001)
        assert ctx.pace() == 109
