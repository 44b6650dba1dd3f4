"""The oracle's step-size controller (MCnptColloid.c:987-1031) against what its arithmetic fixes:
a 5 % change of a float step, the float quotient of the acceptance ratio at the window edges,
the two clamps and the rotation update, and the convergence of the acceptance into the window
on a small oracle run.  Retyping the rule is not used as the check."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from tests.systems import layered_system


def adapt(d, r, acc, att, mt=1.0, half_box=1e9):
    dd, rr = C.c_float(d), C.c_float(r)
    oc.lib().orc_adapt_steps(C.byref(dd), C.byref(rr), acc, att, mt, half_box)
    return np.float32(dd.value), np.float32(rr.value)


def ulps(a, b):
    return abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32)))


@pytest.mark.parametrize("d", [0.1, 0.37, 1.0, 0.0123])
def test_five_percent_steps(d):
    d32 = np.float32(d)
    lo, r = adapt(d32, 0.1, 1, 10)                  # ratio 0.1
    hi, r2 = adapt(d32, 0.1, 9, 10)                 # ratio 0.9
    assert ulps(lo, np.float64(d32) * 0.95) <= 1 and r == np.float32(0.1)
    assert ulps(hi, np.float64(d32) * 1.05) <= 1 and r2 == np.float32(0.1)


def test_window_edges_follow_the_float_quotient():
    d = np.float32(0.2)
    # (float)3/(float)10 = 0.30000001 is not below 0.3: no change
    assert float(np.float32(3) / np.float32(10)) > 0.3
    assert adapt(d, 0.1, 3, 10)[0] == d
    # (float)11/(float)20 = 0.55000001 is above 0.55: the step grows
    assert float(np.float32(11) / np.float32(20)) > 0.55
    assert adapt(d, 0.1, 11, 20)[0] > d
    # inside the window nothing changes
    for acc in (31, 40, 54):
        assert adapt(d, 0.1, acc, 100) == (d, np.float32(0.1))
    # 29/100 shrinks
    assert adapt(d, 0.1, 29, 100)[0] < d


def test_lower_clamp_and_rotation_halving():
    mt = np.float32(1.0)
    d = np.float32(1.0e-4 * 1.04)                   # 0.95 d falls below 1e-4 min_thick
    nd, nr = adapt(d, 0.2, 0, 100, mt)
    assert ulps(nd, 0.01) <= 1 and nr == np.float32(0.1)
    nd, nr = adapt(d, 1e-6, 0, 100, mt)             # rotation at its floor stays
    assert nr == np.float32(1e-6)
    nd, nr = adapt(np.float32(0.5), 0.2, 0, 100, mt)    # no clamp: rotation unchanged
    assert nr == np.float32(0.2)


def test_upper_clamp_and_rotation_growth():
    mt = np.float32(3.0)                            # min(sigma_ee, sigma_ss) of GB(3,5,2,1) is 1
    d = np.float32(4.4)                             # 1.05 d > 1.5 min_thick = 4.5
    nd, nr = adapt(d, 0.2, 90, 100, mt)
    assert ulps(nd, 2.25) <= 1 and ulps(nr, 0.3) <= 1
    nd, nr = adapt(d, 0.75, 90, 100, mt)            # rotation above its ceiling stays
    assert nr == np.float32(0.75)
    # 0.7f = 0.69999999 is below the double 0.7 of the comparison: it still grows
    nd, nr = adapt(d, 0.7, 90, 100, mt)
    assert ulps(nr, 1.05) <= 1
    nd, nr = adapt(np.float32(1.0), 0.2, 90, 100, mt)   # no clamp: rotation unchanged
    assert nr == np.float32(0.2)


def test_increase_bounded_by_half_box():
    d = np.float32(1.0)
    assert adapt(d, 0.2, 90, 100, 10.0, half_box=1.04) == (d, np.float32(0.2))
    assert adapt(d, 0.2, 90, 100, 10.0, half_box=1.06)[0] > d


def test_oracle_run_brings_acceptance_into_the_window():
    q, u, L = layered_system(1300, 0.25, seed=3)
    c = oc.consts(max_displacement=1.5, max_rotation=0.1)
    s = oc.System(c, q, u, L, seed=77, adapt_moves=1000, nthreads=4)
    accs = []
    for _ in range(70):
        before = s.counters.copy()
        s.sweep()
        accs.append((s.counters[1] - before[1]) / (s.counters[0] - before[0]))
    assert accs[0] < 0.1 and s.adapts == 70
    # one sweep of 1300 moves per action: the last 15 sweeps sit in the window (with noise)
    assert 0.3 <= np.mean(accs[-15:]) <= 0.55 and s.dmax < 0.3
    assert s.drot == np.float32(0.1)                       # no clamp was reached
    assert c.max_displacement == pytest.approx(1.5)      # the caller's constants are untouched
