"""Writes tests/golden/ref_physics.npz and ref_loader.json: input/output vectors of the
reference's OWN compiled physics functions (oracle/_ref/libmcgb_ref.so, built from
MCnptColloid.c by oracle/build_ref.sh) on seeded inputs.

Run in the container that holds the reference tree:
    make -C oracle ref && python tests/golden/make_golden.py
The fixtures are data only (inputs and the reference's outputs); no reference text is stored.
"""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
from oracle import oracle_ctypes as oc  # noqa: E402

PARAM_SETS = {
    # name: sigma0, eps0, kappa, kappa', mu, nu
    "gb3521": (1.0, 1.0, 3.0, 5.0, 2.0, 1.0),
    "gb44_20_1_1": (1.0, 1.0, 4.4, 20.0, 1.0, 1.0),
    "lj_limit": (1.0, 1.0, 1.0, 1.0, 2.0, 1.0),
    "gb3512_s12": (1.2, 0.7, 3.0, 5.0, 1.0, 2.0),
}


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_params(L, s0, e0, k, kp, mu, nu, T=1.0, rc=4.0, R=5.0):
    for name, val in (("gb_sigma_sidetoside", s0), ("gb_epsilon0", e0), ("gb_kappa", k),
                      ("gb_kappa_prime", kp), ("gb_mu_exponent", mu), ("gb_nu_exponent", nu),
                      ("system_temperature", T), ("cutoff_radius", rc), ("verlet_radius", rc + 1),
                      ("colloid_radius", R)):
        oc.ref_set(L, name, val)
    oc.ref_set(L, "gb_kappa_sq", 0.0)
    oc.ref_set(L, "gb_sigma_endtoend", s0 * k)   # load_initial_conditions, MCnptColloid.c:4130


CONST_NAMES = ["gb_beta", "gb_4eps0", "cutoff_radius_sq", "gb_kappa_sq", "gb_chi", "gb_chi_sq",
               "gb_k_imu", "gb_chi_prime", "colloid_rod_cutoff", "colloid_radius_sq"]


def rand_dirs(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def pair_inputs(ei, ej, dr):
    """rij, irij, iei, iej, eiej formed exactly as GBij (MCnptColloid.c:1524-1535)."""
    r2 = dot(dr, dr)
    rij = np.sqrt(r2)
    irij = 1.0 / rij
    iei = 1.0 / np.sqrt(dot(ei, ei))
    iej = 1.0 / np.sqrt(dot(ej, ej))
    eiej = iei * iej * dot(ei, ej)
    return rij, irij, iei, iej, eiej


def special_geometries(s0, k):
    """side-by-side, end-to-end, T, X at distances around contact."""
    x, y, z = np.eye(3)
    out = []
    for r in (0.9 * s0, 1.0 * s0, 2 ** (1 / 6) * s0, 1.5 * s0, 2.5 * s0):
        out.append((x, x, r * y))        # side by side
        out.append((x, y, r * z))        # X
    for r in (k * s0 * 0.999, k * s0, k * s0 + 0.12 * s0, k * s0 + 0.5 * s0):
        out.append((x, x, r * x))        # end to end
    for r in (0.5 * (1 + k) * s0, 0.5 * (1 + k) * s0 + 0.1 * s0, 0.5 * (1 + k) * s0 + 1.0 * s0):
        out.append((x, y, r * x))        # T
    return out


def main():
    L = oc.ref_lib()
    if L is None:
        sys.exit("oracle/_ref/libmcgb_ref.so is missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20261203)
    data = {}
    consts = {}
    for name, (s0, e0, k, kp, mu, nu) in PARAM_SETS.items():
        set_params(L, s0, e0, k, kp, mu, nu)
        L.get_system_const()
        first = {n: oc.ref_get(L, n) for n in CONST_NAMES}
        L.get_system_const()                     # second call: kappa^2 is now set
        second = {n: oc.ref_get(L, n) for n in CONST_NAMES}
        consts[name] = {"params": [s0, e0, k, kp, mu, nu], "first_call": first,
                        "second_call": second}

        # ---- pij_GB on random and special pairs (non-unit orientation vectors included)
        n = 200
        ei = rand_dirs(rng, n) * rng.uniform(0.6, 1.4, size=(n, 1))
        ej = rand_dirs(rng, n) * rng.uniform(0.6, 1.4, size=(n, 1))
        ei[: n // 2] /= np.linalg.norm(ei[: n // 2], axis=1, keepdims=True)
        ej[: n // 2] /= np.linalg.norm(ej[: n // 2], axis=1, keepdims=True)
        r = rng.uniform(0.35 * s0, 4.0 * s0, size=n)
        dr = rand_dirs(rng, n) * r[:, None]
        sp = special_geometries(s0, k)
        ei = np.vstack([ei, [g[0] for g in sp]])
        ej = np.vstack([ej, [g[1] for g in sp]])
        dr = np.vstack([dr, [g[2] for g in sp]])
        u = np.empty(len(dr))
        for m in range(len(dr)):
            a, b, d = (np.ascontiguousarray(v[m]) for v in (ei, ej, dr))
            rij, irij, iei, iej, eiej = pair_inputs(a, b, d)
            u[m] = L.pij_GB(dp(a), dp(b), rij, irij, iei, iej, eiej, dp(d))
        data[f"{name}/pair_ei"] = ei
        data[f"{name}/pair_ej"] = ej
        data[f"{name}/pair_dr"] = dr
        data[f"{name}/pair_u"] = u

        # ---- GBCollRod around colloids of radius 5 and 3
        for R in (5.0, 3.0):
            oc.ref_set(L, "colloid_radius", R)
            L.get_system_const()
            box = (C.c_double * 3).in_dll(L, "box_size")
            for dd in range(3):
                box[dd] = 100.0
            m = 150
            d = rng.uniform(R - 0.5, R + 4 * s0 * k + 2.0, size=m)
            rel = rand_dirs(rng, m) * d[:, None]
            e = rand_dirs(rng, m)
            e[m // 2:] *= rng.uniform(0.8, 1.2, size=(m - m // 2, 1))
            # exact closed-form geometries: radial (theta = 0) and tangential (theta = 90 deg)
            extra_rel, extra_e = [], []
            for dist in (R + 1.5, R + 2.0, R + 3.0, R + 5.0):
                extra_rel.append([0.0, 0.0, dist]); extra_e.append([0.0, 0.0, 1.0])
                extra_rel.append([0.0, 0.0, dist]); extra_e.append([1.0, 0.0, 0.0])
            rel = np.vstack([rel, extra_rel])
            e = np.vstack([e, extra_e])
            cpos = np.array([50.0, 50.0, 50.0])
            uc = np.empty(len(rel))
            for q in range(len(rel)):
                ri = np.ascontiguousarray(cpos + rel[q])
                eq = np.ascontiguousarray(e[q])
                uc[q] = L.GBCollRod(0, dp(ri), dp(eq), dp(cpos))
            data[f"{name}/coll_R{int(R)}_rel"] = rel
            data[f"{name}/coll_R{int(R)}_e"] = e
            data[f"{name}/coll_R{int(R)}_u"] = uc

    # ---- compute_distance_sq minimum image in a cubic box
    Lb = 29.876
    box = (C.c_double * 3).in_dll(L, "box_size")
    for dd in range(3):
        box[dd] = Lb
    p1 = rng.uniform(0, Lb, size=(300, 3))
    p2 = rng.uniform(0, Lb, size=(300, 3))
    d_out = np.empty((300, 3))
    r2_out = np.empty(300)
    for m in range(300):
        dd = np.zeros(3)
        r2_out[m] = L.ref_compute_distance_sq(*p1[m], *p2[m], dp(dd))
        d_out[m] = dd
    data["minimg/box"] = np.array([Lb, Lb, Lb])
    data["minimg/p1"] = p1
    data["minimg/p2"] = p2
    data["minimg/dr"] = d_out
    data["minimg/r2"] = r2_out

    # ---- order parameter (float Q sums, Jacobi; its P2/director outputs)
    p2s, dirs, us = [], [], []
    for s in range(12):
        n = 400
        if s < 4:
            u = rand_dirs(rng, n)
        else:
            director = rand_dirs(rng, 1)[0]
            kappa = 2.0 * s
            u = director + rng.normal(size=(n, 3)) / np.sqrt(kappa)
            u /= np.linalg.norm(u, axis=1, keepdims=True)
            u *= np.sign(rng.uniform(-1, 1, size=(n, 1)))
        ux, uy, uz = (np.ascontiguousarray(u[:, d]) for d in range(3))
        fx, fy, fz, fp2 = C.c_float(), C.c_float(), C.c_float(), C.c_float()
        L.compute_liquid_crystal_order_parameters(n, dp(ux), dp(uy), dp(uz), C.byref(fx),
                                                  C.byref(fy), C.byref(fz), C.byref(fp2))
        us.append(u)
        p2s.append(fp2.value)
        dirs.append([fx.value, fy.value, fz.value])
    data["p2/u"] = np.stack(us)
    data["p2/p2"] = np.array(p2s)
    data["p2/director"] = np.array(dirs)

    np.savez_compressed(os.path.join(HERE, "ref_physics.npz"), **data)

    # ---- the reference loader on our parameter files (exit/print behaviour via a subprocess)
    loader = {}
    for fname in sorted(os.listdir(HERE)):
        if not fname.startswith("params_") or not fname.endswith(".txt"):
            continue
        code = (
            "import ctypes as C, sys, json\n"
            f"sys.path.insert(0, {REPO!r})\n"
            "from oracle import oracle_ctypes as oc\n"
            "L = oc.ref_lib()\n"
            f"L.load_initial_conditions({os.path.join(HERE, fname)!r}.encode())\n"
            "ints = ['mc_ensemble_type','num_particles','initial_configuration_type',"
            "'total_mc_cycles','configuration_save_interval','statistics_print_interval',"
            "'colloid_wall_option']\n"
            "flts = ['max_displacement','max_rotation']\n"
            "dbls = ['system_density','system_temperature','system_pressure','gb_sigma_sidetoside',"
            "'gb_epsilon0','gb_kappa','gb_kappa_prime','gb_mu_exponent','gb_nu_exponent',"
            "'cutoff_radius','verlet_radius','colloid_interaction_epsilon',"
            "'colloid_interaction_amplitude','colloid_radius','gb_sigma_endtoend']\n"
            "out = {n: oc.ref_get(L, n, C.c_int) for n in ints}\n"
            "out.update({n: oc.ref_get(L, n, C.c_float) for n in flts})\n"
            "out.update({n: oc.ref_get(L, n) for n in dbls})\n"
            "sys.stdout.flush()\n"
            "print('JSON' + json.dumps(out))\n"
        )
        p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
        values = None
        lines = []
        for line in p.stdout.splitlines():
            if line.startswith("JSON"):
                values = json.loads(line[4:])
            else:
                lines.append(line)
        loader[fname] = {"exit_code": p.returncode, "stdout": lines,
                         "stderr": p.stderr.splitlines(), "values": values}

    with open(os.path.join(HERE, "ref_loader.json"), "w") as f:
        json.dump({"constants": consts, "loader": loader}, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "ref_physics.npz"), "and ref_loader.json")


if __name__ == "__main__":
    main()
