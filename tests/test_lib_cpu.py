"""libmcgb host side without a GPU: exported symbols, parameter loader against the reference's
own loader, derived constants, start configurations, tau_int, eigen solver, fortSphere.1."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")
with open(os.path.join(GOLD, "ref_loader.json")) as f:
    LOADER = json.load(f)


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(REPO, "include", "mcgb.h")).read()
    names = sorted(set(re.findall(r"\b(mcgb_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 25
    L = m.lib()
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("fname", ["params_n8000.txt", "params_n256000_colloid.txt", "params_unknown.txt"])
def test_loader_matches_reference_loader(fname, capfd):
    ref = LOADER["loader"][fname]["values"]
    p = m.load_params(os.path.join(GOLD, fname))
    for k, v in ref.items():
        if k == "gb_sigma_endtoend":
            assert p.gb_sigma_sidetoside * p.gb_kappa == pytest.approx(v, rel=1e-15)
        else:
            assert getattr(p, k) == pytest.approx(v, rel=1e-7 if k.startswith("max_") else 0), k
    assert p.keys_read == (1 << 23) - 1


def test_loader_optional_and_unknown_keys(capfd):
    p = m.load_params(os.path.join(GOLD, "params_unknown.txt"))
    out = capfd.readouterr().out
    assert p.seed == 12345                       # optional key recognised
    assert "Warning: Unknown parameter 'foo_bar' in configuration file" in out
    assert "seed" not in out                     # the reference warned about it; we know it


def test_loader_missing_keys_message():
    with pytest.raises(m.McgbError) as e:
        m.load_params(os.path.join(GOLD, "params_missing.txt"))
    msg = str(e.value)
    ref = LOADER["loader"]["params_missing.txt"]["stdout"]
    assert e.value.code == -3
    for line in ref:
        assert line.strip() in msg


def test_loader_malformed_value():
    with pytest.raises(m.McgbError) as e:
        m.load_params(os.path.join(GOLD, "params_malformed.txt"))
    assert "Error reading parameter: num_particles" in str(e.value)


def test_loader_missing_file():
    with pytest.raises(m.McgbError) as e:
        m.load_params("/nonexistent/parametersFile.txt")
    assert e.value.code == -2


def test_derived_constants_match_reference_second_call():
    ref = LOADER["constants"]["gb3521"]["second_call"]
    p = m.gb_params(1000, 0.3, 1.0, colloid_radius=5.0)
    d = m.derived(p)
    names = ["gb_beta", "gb_4eps0", "cutoff_radius_sq", "gb_kappa_sq", "gb_chi", "gb_chi_sq",
             "gb_k_imu", "gb_chi_prime", None, "colloid_rod_cutoff", "colloid_radius_sq"]
    for v, n in zip(d, names):
        if n:
            assert v == pytest.approx(ref[n], rel=1e-15, abs=1e-15), n
    assert d[8] == 3.0


def test_init_lattice_geometry():
    (x, y, z), (ux, uy, uz), box = m.init_lattice(8000, 0.30, seed=4)
    L = (8000 / 0.3) ** (1 / 3)
    assert box[0] == pytest.approx(L, rel=1e-14)
    a = L / 20                                  # 20^3 simple cubic lattice
    for c in (x, y, z):
        k = c / a - 0.5
        assert np.allclose(k, np.round(k), atol=1e-9)
    assert np.allclose(ux ** 2 + uy ** 2 + uz ** 2, 1.0)
    assert abs(uz.mean()) < 0.03 and abs((uz ** 2).mean() - 1 / 3) < 0.02
    pts = np.round(np.stack([x, y, z], 1) / a - 0.5).astype(int)
    assert len({tuple(p) for p in pts}) == 8000


def test_init_lattice_colloid_exclusion():
    N, R = 20000, 5.0
    (x, y, z), _, box = m.init_lattice(N, 0.32, colloid_radius=R, exclusion_radius=5.5, jitter=0.05, seed=2)
    c = box[0] / 2
    d = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    assert d.min() >= 5.5 - 0.05 * np.sqrt(3) - 1e-9
    assert box[0] ** 3 == pytest.approx(N / 0.32 + 4 / 3 * np.pi * R ** 3, rel=1e-12)


def test_tau_int_ar1():
    rng = np.random.default_rng(0)
    for phi in (0.0, 0.5, 0.9):
        n = 200000
        e = rng.normal(size=n)
        x = np.empty(n)
        x[0] = e[0]
        for i in range(1, n):
            x[i] = phi * x[i - 1] + e[i]
        tau, w = m.tau_int(x)
        want = (1 + phi) / (1 - phi)
        assert tau == pytest.approx(want, rel=0.08), (phi, tau, want)


def test_eigen_sym3_against_numpy():
    rng = np.random.default_rng(1)
    for _ in range(200):
        A = rng.normal(size=(3, 3))
        A = A + A.T
        if _ % 7 == 0:
            A = np.diag(rng.normal(size=3))
        if _ % 11 == 0:
            v = rng.normal(size=3)
            A = np.outer(v, v) - np.eye(3) * 0.2         # repeated eigenvalue
        q = [A[0, 0], A[1, 1], A[2, 2], A[0, 1], A[0, 2], A[1, 2]]
        ev, V = m.eigen_sym3(q)
        assert np.allclose(ev, np.linalg.eigvalsh(A), atol=1e-10)
        for k in range(3):
            assert np.allclose(A @ V[:, k], ev[k] * V[:, k], atol=1e-8)


def test_fort1_round_trip(tmp_path):
    (x, y, z), (ux, uy, uz), box = m.init_lattice(500, 0.2, seed=9)
    path = str(tmp_path / "fortSphere.1")
    L = m.lib()
    m.check(L.mcgb_write_fort1(path.encode(), 500, *(m._d(a) for a in (x, y, z, ux, uy, uz)),
                               m._d(box), 5.0))
    lines = open(path).read().splitlines()
    assert len(lines) == 1 + 4 * 500 and lines[2].split() == ["0.00000000"] * 3
    arrs = [np.empty(500) for _ in range(6)]
    b2 = np.empty(3)
    R = np.empty(1)
    m.check(L.mcgb_read_fort1(path.encode(), 500, *(m._d(a) for a in arrs), m._d(b2), m._d(R)))
    assert np.allclose(arrs[0], x, atol=1e-8) and np.allclose(arrs[5], uz, atol=1e-8)
    assert R[0] == 5.0 and np.allclose(b2, box, atol=1e-8)


def test_resume_file_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(3)
    n = 257
    q = [rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    q[0][:2] = (0, 2**32 - 1)
    u = [rng.standard_normal(n) for _ in range(3)]
    u[0][:3] = (-0.0, 5e-324, np.nextafter(1.0, 2.0))
    box = np.array([31.25, np.nextafter(40.0, 0.0), 1e-3 / 3])
    rec = m.Resume(seed=2**64 - 1, sweep_index=2**32 - 1, max_displacement=np.float32(0.1),
                   max_rotation=np.float32(1e-7), adapt_attempted=12345678901,
                   adapt_accepted=7, adaptations=2**40, capped_pairs=3,
                   capped_colloid=2**40 + 1, energy_finite=-np.nextafter(-1234.5, 0.0))
    path = str(tmp_path / "fortSphere.1.mcgb")
    m.write_resume(path, q, u, box, rec)
    q2, u2, b2, r2 = m.read_resume(path, n)
    for a, b in zip(q + u + [box], q2 + u2 + [b2]):
        assert a.tobytes() == b.tobytes()
    assert r2.as_dict() == rec.as_dict()
    with pytest.raises(m.McgbError, match="holds 257 particles, 256 requested"):
        m.read_resume(path, 256)
    bad = str(tmp_path / "bad")
    with open(bad, "w") as f:
        f.write(open(path).read().replace("mcgb-resume 1", "mcgb-resume 2", 1))
    with pytest.raises(m.McgbError, match="error reading"):
        m.read_resume(bad, n)
    with open(bad, "w") as f:
        f.write("\n".join(open(path).read().splitlines()[:-1]))
    with pytest.raises(m.McgbError, match="error reading"):
        m.read_resume(bad, n)


def test_create_without_gpu_fails_loudly():
    if m.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(m.McgbError) as e:
        m.Context(m.gb_params(8000, 0.3, 1.2))
    assert e.value.code == -4


def test_driver_reports_missing_parameter_file(tmp_path):
    drv = os.path.join(REPO, "parellel_mc_mpi_for_liquidcrystalsystems_amd", "mcgb")
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    p = subprocess.run([drv, str(tmp_path / "nope.txt")], capture_output=True, text=True)
    assert p.returncode == 1 and "Error opening the parameters file" in p.stderr


def _driver_params(tmp_path, **over):
    txt = open(os.path.join(GOLD, "params_n8000.txt")).read()
    for k, v in over.items():
        txt = re.sub(r"(?m)^%s\s+\S+" % k, "%s %s" % (k, v), txt)
    path = tmp_path / "parametersFile.txt"
    path.write_text(txt)
    return path


def test_driver_cubic_start_writes_trajectory_before_the_gpu(tmp_path):
    """initial_configuration_type 2 in the driver: the first Configuraciones.dat frame is the
    cubic start of mcgb_init_cubic (written before the GPU context exists, as the reference
    writes it before its first sweep, MCnptColloid.c:877)."""
    drv = os.path.join(REPO, "parellel_mc_mpi_for_liquidcrystalsystems_amd", "mcgb")
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    if m.device_count() > 0:
        pytest.skip("a GPU is visible (the GPU suite runs the driver end to end)")
    path = _driver_params(tmp_path, num_particles=1000, system_density=0.02,
                          initial_configuration_type=2, total_mc_cycles=0)
    p = subprocess.run([drv, str(path)], capture_output=True, text=True, cwd=tmp_path)
    assert p.returncode == 1 and "cannot create the GPU context" in p.stderr
    pos, u, box = m.init_cubic(1000, 0.02, 0.0, seed=m.load_params(str(path)).seed)
    want = tmp_path / "want.dat"
    m.write_trajectory(want, pos, u, box)
    assert (tmp_path / "Configuraciones.dat").read_bytes() == want.read_bytes()


def test_driver_random_start_reports_jammed_density(tmp_path):
    """Type 1 at rho* = 0.3 with sigma_ee = 3 (packing fraction 4.2) cannot be built by random
    sequential addition; the reference loops forever there, the driver stops with an error."""
    drv = os.path.join(REPO, "parellel_mc_mpi_for_liquidcrystalsystems_amd", "mcgb")
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    path = _driver_params(tmp_path, num_particles=2000, initial_configuration_type=1)
    p = subprocess.run([drv, str(path)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert p.returncode == 1 and "no room for particle" in p.stderr


@pytest.mark.parametrize("field", ["gb_epsilon0", "system_temperature", "max_rotation",
                                   "colloid_radius"])
def test_create_refuses_non_finite_parameters(field):
    # checked before any device call: the same answer with or without a GPU
    p = m.gb_params(8000, 0.3, 1.2)
    setattr(p, field, float("nan"))
    with pytest.raises(m.McgbError) as e:
        m.Context(p)
    assert e.value.code == m.ERR_ARG and "not finite" in str(e.value)
