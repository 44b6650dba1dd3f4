"""The oracle's sweep samples the exact Boltzmann distribution.

Every other oracle test replays the sweep's own formulas; these tests pin what those formulas
must produce, independently of them:

* the move draws: the rotation vector is uniform on the unit sphere (the distribution of the
  reference's Marsaglia table, pto_ransphr/Marsaglia4MC, MCnptColloid.c:3036-3078, used at
  :2526-2550): mean 0, second moments I/3, KS tests on z and the azimuth; the fixed-point step is
  uniform and symmetric (:2502-2509); the Metropolis threshold -log(U) T is exponential
  (:2566-2574);
* the chain: one rod near the colloid against the exact (distance, angle) density by quadrature,
  and two Gay-Berne rods against importance sampling of e^{-beta U}, for both sublattice
  decompositions;
* that these checks have teeth: each of six plausible mistakes planted in the oracle (a biased
  azimuth or polar draw, beta for T in the threshold, a dropped trial colloid term, an asymmetric
  step, a dropped old-state pair term) makes them fail.
"""
import numpy as np
import pytest
from scipy import stats

from oracle import oracle_ctypes as oc
from tests import stationary as st

SEED = 20261204


@pytest.fixture(autouse=True)
def _no_mutation():
    oc.set_mutation(0)
    yield
    oc.set_mutation(0)


def draws(n_per=50000, T=0.7, dmax=3.0, L=(16.5, 17.0, 18.0)):
    out = [oc.move_draws(SEED, r, t, p, cell, 0, n_per, np.array(L), dmax, T)
           for r, t, p, cell in ((0, 0, 0, 0), (1, 5, 3, 17), (2, 9, 7, 63), (0, 123, 5, 40))]
    j = np.concatenate([o[0] for o in out])
    dq = np.concatenate([o[1] for o in out])
    v = np.concatenate([o[2] for o in out])
    thr = np.concatenate([o[3] for o in out])
    K = np.array([round(np.float32(dmax) * 2.0 ** 32 / x) for x in L], dtype=np.float64)
    return j, dq, v, thr, K


def sphere_checks(v):
    """Failures of the unit-sphere checks on rotation vectors v (n, 3): [] when uniform."""
    n = len(v)
    bad = []
    if np.max(np.abs(np.sum(v * v, axis=1) - 1.0)) > 1e-14:
        bad.append("not unit vectors")
    m = v.mean(axis=0)
    if np.any(np.abs(m) > 4.0 * np.sqrt(1.0 / 3.0 / n)):
        bad.append(f"mean {m}")
    sq = (v ** 2).mean(axis=0) - 1.0 / 3.0
    if np.any(np.abs(sq) > 4.0 * np.sqrt((1.0 / 5.0 - 1.0 / 9.0) / n)):
        bad.append(f"<v_i^2> - 1/3 = {sq}")
    off = np.array([(v[:, 0] * v[:, 1]).mean(), (v[:, 0] * v[:, 2]).mean(),
                    (v[:, 1] * v[:, 2]).mean()])
    if np.any(np.abs(off) > 4.0 * np.sqrt(1.0 / 15.0 / n)):
        bad.append(f"<v_i v_j> = {off}")
    pz = stats.kstest(v[:, 2], "uniform", args=(-1.0, 2.0)).pvalue
    pphi = stats.kstest(np.arctan2(v[:, 1], v[:, 0]), "uniform", args=(-np.pi, 2 * np.pi)).pvalue
    if pz < 1e-3:
        bad.append(f"KS z p = {pz:.2e}")
    if pphi < 1e-3:
        bad.append(f"KS azimuth p = {pphi:.2e}")
    return bad


def test_rotation_draws_are_uniform_on_the_sphere():
    _, _, v, _, _ = draws()
    assert sphere_checks(v) == []


def test_step_selector_and_threshold_distributions():
    j, dq, v, thr, K = draws(T=0.7)
    n = len(j)
    # step: dq/K uniform on (-1/2, 1/2), symmetric (the proposal must be symmetric)
    for a in range(3):
        x = dq[:, a] / K[a]
        assert np.all(np.abs(dq[:, a]) <= K[a] / 2)
        assert stats.kstest(x, "uniform", args=(-0.5, 1.0)).pvalue > 1e-3
        assert abs(x.mean()) < 4.0 * np.sqrt(1.0 / 12.0 / n)
    # threshold -log(U) T is exponential with mean T
    assert stats.kstest(thr / 0.7, "expon").pvalue > 1e-3
    # particle selector uniform on [0, 2^32)
    assert stats.kstest(j / 2.0 ** 32, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("mutation", [oc.MUT_PHI_HALF, oc.MUT_Z_HALF])
def test_planted_rotation_bias_fails_the_sphere_checks(mutation):
    oc.set_mutation(mutation)
    _, _, v, _, _ = draws(n_per=20000)
    assert sphere_checks(v) != []


# ------------------------------------------------------------------ chains

def rod_consts(T=0.6):
    return oc.consts(temperature=T, max_displacement=3.0, max_rotation=1.0,
                     colloid_radius=st.ROD_R, colloid_on=1)


def run_rod(c, nsweeps, nch=8, stride=20, seed=SEED):
    rng = np.random.default_rng(seed)
    q = np.zeros((nch, 3, 1), dtype=np.uint32)           # box corner: far from the colloid
    u = st.sphere_points(rng, nch).reshape(nch, 3, 1)
    L = np.full((nch, 3), st.ROD_L)
    oq, ou, _, _ = oc.run_chains(c, q, u, L, seed, nsweeps, stride)
    return st.rod_observables(oq[..., 0], ou[..., 0], st.ROD_L)


def rod_verdict(c, d, cth, un):
    P = st.colloid_rod_reference(c, st.ROD_L, st.ROD_D_EDGES, st.ROD_C_EDGES).ravel()
    z, m, s = st.zscores(st.rod_indicators(d, cth), P)
    ok1, msg1 = st.check(st.ROD_NAMES, z, m, P, s, "(d, |cos theta|) bins")
    zm, mm, sm = st.zscores(st.orientation_moments(un), np.zeros(9))
    ok2, msg2 = st.check(st.MOMENT_NAMES, zm, mm, np.zeros(9), sm, "orientation moments")
    return ok1, ok2, msg1 + "\n" + msg2, z, zm


def test_one_rod_near_the_colloid_matches_quadrature():
    c = rod_consts()
    ok1, ok2, msg, _, _ = rod_verdict(c, *run_rod(c, 400000))
    assert ok1 and ok2, msg


@pytest.mark.parametrize("mutation", [oc.MUT_PHI_HALF, oc.MUT_Z_HALF, oc.MUT_BETA,
                                      oc.MUT_NO_TRIAL_COLLOID, oc.MUT_STEP_DRIFT])
def test_planted_mistakes_fail_the_one_rod_check(mutation):
    c = rod_consts()
    oc.set_mutation(mutation)
    obs = run_rod(c, 200000)
    oc.set_mutation(0)
    ok1, ok2, msg, z, zm = rod_verdict(c, *obs)
    assert not (ok1 and ok2), msg
    # a clear failure, not a marginal one
    assert max(np.abs(z).max(), np.abs(zm).max()) > 2 * st.MAX_Z, msg


def two_rod_consts(T=0.6, colors=2):
    return oc.consts(temperature=T, max_displacement=3.0 if colors == 2 else 1.8,
                     max_rotation=1.0, colors=colors)


def run_two(c, nsweeps, nch=8, stride=10, seed=SEED):
    rng = np.random.default_rng(seed)
    q = np.zeros((nch, 3, 2), dtype=np.uint32)
    q[:, 0, 1] = 2 ** 31                                 # half a box apart
    u = np.moveaxis(st.sphere_points(rng, 2 * nch).reshape(nch, 2, 3), 1, 2)
    L = np.full((nch, 3), st.ROD_L)
    oq, ou, _, _ = oc.run_chains(c, q, u, L, seed, nsweeps, stride)
    return two_rod_samples(c, oq, ou)


def two_rod_samples(c, q, u):
    """Indicators of TWO_ROD_OBS from states q (..., 3, 2), u (..., 3, 2); the pair energy is
    recomputed with the FP64 oracle."""
    dr, r, cg, u0, u1 = st.pair_of(q, u, st.ROD_L)
    U = np.zeros(r.shape)
    ins = r < st.TWO_ROD_OBS.rc
    Uv, cap = oc.gb_pair_batch(c, u0[ins], u1[ins], dr[ins])
    U[ins] = np.where(cap != 0, 1e9, Uv)
    return st.TWO_ROD_OBS.indicators(r, cg, U)


def two_rod_verdict(c, X):
    P, se, _ = st.two_rod_reference(c, st.ROD_L, st.TWO_ROD_OBS)
    z, m, s = st.zscores(X, P, se)
    ok, msg = st.check(st.TWO_ROD_OBS.names, z, m, P, s, "two rods")
    return ok, msg, z


@pytest.mark.parametrize("colors", [2, 3])
def test_two_rods_match_importance_sampling(colors):
    c = two_rod_consts(colors=colors)
    ok, msg, _ = two_rod_verdict(c, run_two(c, 400000 if colors == 2 else 200000))
    assert ok, msg


@pytest.mark.parametrize("mutation", [oc.MUT_BETA, oc.MUT_DROP_OLD])
def test_planted_mistakes_fail_the_two_rod_check(mutation):
    c = two_rod_consts()
    oc.set_mutation(mutation)
    X = run_two(c, 100000)
    oc.set_mutation(0)
    ok, msg, z = two_rod_verdict(c, X)
    assert not ok and np.abs(z).max() > 2 * st.MAX_Z, msg
