"""Statistical parity of the FP32 GPU sweep with the FP64 oracle: from the same equilibrated
state, both chains sample the same ensemble, so <E>/N, P2 and (NPT) the density agree within
autocorrelation-corrected error bars (Sokal tau_int, c = 5).  The FP32 and FP64 chains take
different decisions after a few moves (rounding), so this is the comparison that remains.
The z-scores are printed; the test bound is 2 combined standard errors (the 2 sigma of the
statistical parity criterion).  The chains are seeded, so each comparison is deterministic."""
import os

import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.systems import layered_system

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def _threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 4


def _mean_se(x):
    x = np.asarray(x, dtype=np.float64)
    tau, _ = m.tau_int(x)
    return x.mean(), x.std() * np.sqrt(max(tau, 1.0) / len(x))


def _z(a, b):
    (ma, sa), (mb, sb) = _mean_se(a), _mean_se(b)
    return abs(ma - mb) / max(np.hypot(sa, sb), 1e-300), ma, mb


def test_nvt_energy_and_order_match_oracle():
    N, rho, T, S = 2000, 0.30, 1.0, 1500
    q, u, L = layered_system(N, rho, seed=12)
    p = m.gb_params(N, rho, T, precision=32, seed=77)
    with m.Context(p) as ctx:
        ctx.set_state_fixed(q, u, L)
        ctx.sweep(300)
        q0, u0, L0 = ctx.get_state_fixed()
        ctx.series_clear()
        ctx.sweep(S)
        eg, pg, _ = ctx.series()
    s = oc.System(oc.consts(temperature=T), q0, u0, L0, seed=78, nthreads=_threads())
    eo, po = np.empty(S), np.empty(S)
    for k in range(S):
        s.sweep()
        eo[k] = s.energy_value() / N
        po[k] = s.order_parameter()[0]
    ze, mge, moe = _z(eg, eo)
    zp, mgp, mop = _z(pg, po)
    print(f"E/N gpu {mge:.5f} oracle {moe:.5f} z {ze:.2f}; P2 gpu {mgp:.4f} oracle {mop:.4f} z {zp:.2f}")
    assert ze < 2 and zp < 2


def test_npt_density_matches_oracle():
    N, rho0, T, P, S = 3000, 0.28, 1.0, 2.0, 1500
    q, u, L = layered_system(N, rho0, seed=13)
    p = m.gb_params(N, rho0, T, precision=32, seed=91, npt=True, pressure=P, npt_max_dlnv=0.01)
    with m.Context(p) as ctx:
        ctx.set_state_fixed(q, u, L)
        ctx.sweep(500)
        q0, u0, L0 = ctx.get_state_fixed()
        ctx.series_clear()
        ctx.sweep(S)
        _, pg, rg = ctx.series()
        st = ctx.stats()
    assert st["volume_accepted"] > 0.2 * st["volume_attempted"]
    s = oc.System(oc.consts(temperature=T, pressure=P), q0, u0, L0, seed=92, npt=True,
                  npt_max_dlnv=0.01, nthreads=_threads())
    ro, po = np.empty(S), np.empty(S)
    for k in range(S):
        s.sweep()
        ro[k] = N / np.prod(s.L)
        po[k] = s.order_parameter()[0]
    zr, mgr, mor = _z(rg, ro)
    zp, mgp, mop = _z(pg, po)
    print(f"rho gpu {mgr:.5f} oracle {mor:.5f} z {zr:.2f}; P2 gpu {mgp:.4f} oracle {mop:.4f} z {zp:.2f}")
    assert zr < 2 and zp < 2
