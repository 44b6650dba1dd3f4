"""Closed forms, invariants, known-answer vectors and brute force that pin the FP64 oracle
beyond the reference's vectors (each chosen so that a dropped term, a wrong sign or index, or
a transposed operand fails)."""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc

X, Y, Z = np.eye(3)
C3521 = oc.consts()            # GB(kappa=3, kappa'=5, mu=2, nu=1), sigma0 = eps0 = 1


def U(c, ei, ej, dr):
    u, cap = oc.gb_pair(c, ei, ej, dr)
    return u[0], cap[0]


def test_constants_closed_forms():
    assert C3521.chi == pytest.approx(0.8, abs=1e-15)
    assert C3521.chi_prime == pytest.approx((np.sqrt(5) - 1) / (np.sqrt(5) + 1), rel=1e-15)
    assert C3521.sigma_ee == 3.0 and C3521.colloid_cut == 12.0


@pytest.mark.parametrize("r", [1.0, 2 ** (1 / 6), 1.3, 2.0, 3.5])
def test_side_by_side(r):
    # c = 1, a = b = 0: sigma = sigma0, eps = (1 - chi^2)^(-nu/2) = 5/3
    u, _ = U(C3521, X, X, r * Y)
    rho = 1.0 / r
    assert u == pytest.approx(4 * (5 / 3) * (rho ** 12 - rho ** 6), rel=1e-13)


def test_side_by_side_well_depth():
    u, _ = U(C3521, X, X, 2 ** (1 / 6) * Y)
    assert u == pytest.approx(-5 / 3, rel=1e-13)


def test_end_to_end_contact_and_depth():
    # c = a = b = 1: sigma = kappa*sigma0 = 3, eps = eps_ss / kappa' = 1/3
    s = 3.0
    for dist in (0.95, 1.0, 2 ** (1 / 6), 1.7):
        u, _ = U(C3521, X, X, (s - 1 + dist) * X)
        rho = 1 / dist
        assert u == pytest.approx(4 * (1 / 3) * (rho ** 12 - rho ** 6), rel=1e-12)
    u, _ = U(C3521, X, X, (s - 1 + 2 ** (1 / 6)) * X)
    assert u == pytest.approx(-1 / 3, rel=1e-12)


def test_t_shape():
    # e_i || r, e_j perpendicular: a = 1, b = 0, c = 0: g(w) = 1 - w
    chi, chip = C3521.chi, C3521.chi_prime
    sigma = 1 / np.sqrt(1 - chi)
    eps = (1 - chip) ** 2
    for r in (sigma + 0.05, sigma + 0.3, sigma + 1.0):
        u, _ = U(C3521, X, Y, r * X)
        rho = 1 / (r - sigma + 1)
        assert u == pytest.approx(4 * eps * (rho ** 12 - rho ** 6), rel=1e-12)


def test_x_shape():
    for r in (1.0, 1.2, 2.0):
        u, _ = U(C3521, X, Y, r * Z)
        assert u == pytest.approx(4 * (r ** -12 - r ** -6), rel=1e-13)


def test_lennard_jones_limit():
    c = oc.consts(kappa=1.0, kappa_prime=1.0, sigma0=1.3, epsilon0=0.7)
    rng = np.random.default_rng(0)
    for _ in range(50):
        ei, ej = rng.normal(size=3), rng.normal(size=3)
        dr = rng.normal(size=3)
        dr *= rng.uniform(1.1, 3.5) / np.linalg.norm(dr)
        u, _ = U(c, ei, ej, dr)
        r = np.linalg.norm(dr)
        assert u == pytest.approx(4 * 0.7 * ((1.3 / r) ** 12 - (1.3 / r) ** 6), rel=1e-12)


def test_symmetries():
    rng = np.random.default_rng(1)
    for _ in range(100):
        ei, ej = rng.normal(size=3), rng.normal(size=3)
        dr = rng.normal(size=3)
        dr *= rng.uniform(1.0, 4.0) / np.linalg.norm(dr)
        u0, _ = U(C3521, ei, ej, dr)
        assert U(C3521, ej, ei, -dr)[0] == pytest.approx(u0, rel=1e-12, abs=1e-14)
        assert U(C3521, -ei, ej, dr)[0] == pytest.approx(u0, rel=1e-12, abs=1e-14)
        assert U(C3521, ei, -ej, -dr)[0] == pytest.approx(u0, rel=1e-12, abs=1e-14)
        assert U(C3521, 2.5 * ei, 0.4 * ej, dr)[0] == pytest.approx(u0, rel=1e-12, abs=1e-14)
        R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        assert U(C3521, R @ ei, R @ ej, R @ dr)[0] == pytest.approx(u0, rel=1e-11, abs=1e-13)


def test_cap_boundary():
    # U = 1e9 iff r - sigma + sigma0 < 0.001 (MCnptColloid.c:2670); side by side sigma = 1
    # just outside the cap the finite energy is ~4 eps (1e3)^12: the cap LOWERS the energy
    u, cap = U(C3521, X, X, (0.001 + 1e-9) * Y)
    assert cap == 0 and u == pytest.approx(4 * (5 / 3) * (1 / (0.001 + 1e-9)) ** 12, rel=1e-6)
    u, cap = U(C3521, X, X, (0.001 - 1e-9) * Y)
    assert cap == 1 and u == 1e9


def test_colloid_closed_forms():
    R = 5.0
    c = oc.consts(colloid_radius=R, colloid_on=1)
    chi, chip = c.chi, c.chi_prime

    def ref(d, sin2):
        scr = np.sqrt((1 - chi * sin2) / (1 - chi))
        ecr = 5.0 * ((1 - chip) / (1 - chip * sin2)) ** 2
        x = d - R - scr + 1
        return ecr * (1 / (45 * x ** 9) - 1 / (6 * x ** 3) - 1 / (40 * d * x ** 8) + 1 / (4 * d * x ** 2))

    for d in (6.5, 7.0, 9.0, 16.9):
        u, _ = oc.colloid(c, [0, 0, d], [0, 0, 1])          # radial: theta = 0
        assert u[0] == pytest.approx(ref(d, 0.0), rel=1e-12)
        u, _ = oc.colloid(c, [0, d, 0], [1, 0, 0])          # tangential: theta = 90
        assert u[0] == pytest.approx(ref(d, 1.0), rel=1e-12)
    assert oc.colloid(c, [0, 0, R + 12.0001], [0, 0, 1])[0][0] == 0.0   # cutoff 4 sigma_ee
    u, cap = oc.colloid(c, [0, 0, R + 1.0], [0, 0, 1])                 # x <= 0 -> cap
    assert cap[0] == 1 and u[0] == 1e8


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(oc.philox(ctr, key)) == want


def test_cell_dim_rule():
    assert oc.cell_dim(29.876, 4.0) == 6
    assert oc.cell_dim(92.852, 4.0) == 22
    assert oc.cell_dim(16.0, 4.0) == 2 * int(16.0 / (8.0 * (1 + 1e-6))) or oc.cell_dim(16.0, 4.0) == 0
    assert oc.cell_dim(15.9, 4.0) == 0      # fewer than 4 cells: no 8-sublattice grid


def test_cell_list_brute_force():
    rng = np.random.default_rng(3)
    N = 3000
    q = [rng.integers(0, 2 ** 32, size=N, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    n = np.array([6, 8, 4], dtype=np.int32)
    g = rng.integers(0, 2 ** 32, size=3, dtype=np.uint64).astype(np.uint32)
    cs, so = oc.cell_list(q, n, g)
    assert cs[0] == 0 and cs[-1] == N and np.all(np.diff(cs) >= 0)
    assert sorted(so.tolist()) == list(range(N))
    for c in range(int(np.prod(n))):
        members = so[cs[c]:cs[c + 1]]
        assert np.all(np.diff(members) > 0)          # ordered by particle index
        cz, rem = divmod(c, n[0] * n[1])
        cy, cx = divmod(rem, n[0])
        for d, cd in zip(range(3), (cx, cy, cz)):
            f = ((q[d][members] - g[d]).astype(np.uint64) * int(n[d])) >> np.uint64(32)
            assert np.all(f == cd)


def test_sweep_setup_order_is_permutation():
    seen = set()
    for t in range(200):
        g, order, vol = oc.sweep_setup(42, 0, t)
        assert sorted(order.tolist()) == list(range(8)) and len(order) == 8
        seen.add(tuple(order.tolist()))
    assert len(seen) > 150        # the shuffle really shuffles
