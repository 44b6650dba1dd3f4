"""The FP64 oracle against vectors of the reference's OWN compiled physics functions
(tests/golden/ref_physics.npz, ref_loader.json, written by tests/golden/make_golden.py from
oracle/_ref/libmcgb_ref.so), and — where the reference tree was built here — against the
live reference functions on fresh random inputs."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import oracle_ctypes as oc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PHYS = np.load(os.path.join(GOLD, "ref_physics.npz"))
with open(os.path.join(GOLD, "ref_loader.json")) as f:
    LOADER = json.load(f)

SETS = {name: v["params"] for name, v in LOADER["constants"].items()}


def consts_for(name, R=0.0):
    s0, e0, k, kp, mu, nu = SETS[name]
    return oc.consts(sigma0=s0, epsilon0=e0, kappa=k, kappa_prime=kp, mu=mu, nu=nu,
                     colloid_radius=R, colloid_on=1 if R > 0 else 0)


@pytest.mark.parametrize("name", sorted(SETS))
def test_constants_match_reference_second_call(name):
    # get_system_const (MCnptColloid.c:2752-2783) assigns kappa^2 after using it; its second
    # call gives the intended constants, which the oracle computes in dependency order.
    c = consts_for(name, R=5.0)
    ref = LOADER["constants"][name]["second_call"]
    assert c.chi == pytest.approx(ref["gb_chi"], rel=1e-15, abs=1e-15)
    assert c.chi2 == pytest.approx(ref["gb_chi_sq"], rel=1e-15, abs=1e-15)
    assert c.chi_prime == pytest.approx(ref["gb_chi_prime"], rel=1e-15, abs=1e-15)
    assert c.k_imu == pytest.approx(ref["gb_k_imu"], rel=1e-15)
    assert c.four_eps0 == pytest.approx(ref["gb_4eps0"], rel=1e-15)
    assert c.rc2 == pytest.approx(ref["cutoff_radius_sq"], rel=1e-15)
    assert c.colloid_cut == pytest.approx(ref["colloid_rod_cutoff"], rel=1e-15)
    assert c.colloid_R2 == pytest.approx(ref["colloid_radius_sq"], rel=1e-15)


def test_reference_first_call_has_chi_minus_one():
    # documents the reference defect the oracle does not reproduce (chi uses kappa^2 = 0)
    for name in SETS:
        assert LOADER["constants"][name]["first_call"]["gb_chi"] == -1.0


@pytest.mark.parametrize("name", sorted(SETS))
def test_gb_pair_matches_reference_vectors(name):
    c = consts_for(name)
    ei, ej, dr = PHYS[f"{name}/pair_ei"], PHYS[f"{name}/pair_ej"], PHYS[f"{name}/pair_dr"]
    ref = PHYS[f"{name}/pair_u"]
    u, cap = oc.gb_pair(c, ei, ej, dr)
    assert np.array_equal(cap == 1, ref == 1e9)
    np.testing.assert_allclose(u, ref, rtol=1e-13, atol=1e-13)
    assert (cap == 0).sum() > 100
    if name != "lj_limit":                 # sigma = sigma0 there: never inside the cap
        assert (cap == 1).sum() > 0        # both branches exercised


@pytest.mark.parametrize("name", sorted(SETS))
@pytest.mark.parametrize("R", [5, 3])
def test_colloid_matches_reference_vectors(name, R):
    c = consts_for(name, R=float(R))
    rel, e = PHYS[f"{name}/coll_R{R}_rel"], PHYS[f"{name}/coll_R{R}_e"]
    ref = PHYS[f"{name}/coll_R{R}_u"]
    rel = (50.0 + rel) - 50.0          # the reference saw the rod at colloid centre (50,50,50) + rel
    u, cap = oc.colloid(c, rel, e)
    assert np.array_equal(cap == 1, ref == 1e8)
    np.testing.assert_allclose(u, ref, rtol=1e-13, atol=1e-13)
    assert (ref == 0.0).sum() > 0 and (np.abs(ref) > 0).sum() > 50


def test_fixed_point_minimum_image_matches_reference():
    # the build stores positions as 32-bit fractions of the box; the wrap-around int32
    # difference must reproduce compute_distance_sq's minimum image (MCnptColloid.c:2690-2721)
    L = PHYS["minimg/box"]
    p1, p2, ref = PHYS["minimg/p1"], PHYS["minimg/p2"], PHYS["minimg/dr"]
    for d in range(3):
        q1 = oc.to_fixed(p1[:, d], L[d])
        q2 = oc.to_fixed(p2[:, d], L[d])
        dq = (q2.astype(np.int64) - q1.astype(np.int64)).astype(np.uint32).view(np.int32)
        dr = dq.astype(np.float64) * (L[d] / oc.TWO32)
        np.testing.assert_allclose(dr, ref[:, d], rtol=0, atol=2.0 * L[d] / oc.TWO32)


def test_order_parameter_matches_reference_vectors():
    # the reference returns 1.5 * eigenvalue[1] when eigenvalue[0] is strictly largest
    # (MCnptColloid.c:1890-1891); the oracle returns 1.5 * the true largest eigenvalue.
    U, ref = PHYS["p2/u"], PHYS["p2/p2"]
    for k in range(U.shape[0]):
        u = U[k]
        sys = [np.ascontiguousarray(u[:, d]) for d in range(3)]
        p2 = C.c_double()
        dvec = np.empty(3)
        oc.lib().orc_order_parameter(u.shape[0], *(oc._d(a) for a in sys), C.byref(p2), oc._d(dvec))
        Q = u.T @ u / u.shape[0] - np.eye(3) / 3
        lam = np.linalg.eigvalsh(Q)
        assert p2.value == pytest.approx(1.5 * lam[-1], abs=1e-12)
        # the reference's float result is one of the 1.5*eigenvalues (the largest unless D12 fires)
        assert np.min(np.abs(1.5 * lam - ref[k])) < 1e-4
        if abs(ref[k] - 1.5 * lam[-1]) < 1e-4:
            assert abs(abs(np.dot(dvec, PHYS["p2/director"][k])) - 1.0) < 1e-3


def _ref_or_skip():
    L = oc.ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return L


def test_gb_pair_matches_live_reference_random():
    L = _ref_or_skip()
    from tests.golden.make_golden import set_params, pair_inputs, rand_dirs
    rng = np.random.default_rng(7)
    for name in SETS:
        s0, e0, k, kp, mu, nu = SETS[name]
        set_params(L, s0, e0, k, kp, mu, nu)
        L.get_system_const()
        L.get_system_const()
        c = consts_for(name)
        n = 300
        ei = rand_dirs(rng, n) * rng.uniform(0.5, 1.5, size=(n, 1))
        ej = rand_dirs(rng, n)
        dr = rand_dirs(rng, n) * rng.uniform(0.3, 4.0, size=(n, 1))
        u, _ = oc.gb_pair(c, ei, ej, dr)
        dp = C.POINTER(C.c_double)
        for m in range(n):
            a, b, d = (np.ascontiguousarray(v[m]) for v in (ei, ej, dr))
            rij, irij, iei, iej, eiej = pair_inputs(a, b, d)
            ref = L.pij_GB(a.ctypes.data_as(dp), b.ctypes.data_as(dp), rij, irij, iei, iej, eiej,
                           d.ctypes.data_as(dp))
            assert u[m] == pytest.approx(ref, rel=1e-13, abs=1e-13)


@pytest.mark.parametrize("fname", ["params_n8000.txt", "params_n256000_colloid.txt", "params_unknown.txt"])
def test_reference_loader_parses_fixture_files(fname):
    # what the reference's own loader read from our fixture files (same key format)
    rec = LOADER["loader"][fname]
    assert rec["exit_code"] == 0
    assert rec["values"]["gb_kappa"] == 3.0


def test_reference_loader_error_texts():
    rec = LOADER["loader"]["params_missing.txt"]
    assert rec["exit_code"] == 1
    assert rec["stdout"][0] == "Error: Not all parameters were read correctly. Expected 23, got 21"
    rec = LOADER["loader"]["params_malformed.txt"]
    assert rec["exit_code"] == 1 and rec["stderr"] == ["Error reading parameter: num_particles"]
