"""Independent Python restatement of one checkerboard sweep (test helper).

It re-derives every trial move from the Philox draws and recomputes each dE by brute force
over ALL particles with the reference's minimum image (compute_distance_sq,
MCnptColloid.c:2690-2721) and a numpy restatement of pij_GB/GBCollRod.  The oracle and the
GPU must both agree with it."""
import numpy as np

from oracle import oracle_ctypes as oc

TWO32 = oc.TWO32


def gb_numpy(c, ei, ej, dr):
    """pij_GB (MCnptColloid.c:2618-2681) for one e_i against rows ej, dr."""
    r2 = np.einsum("ij,ij->i", dr, dr)
    rij = np.sqrt(r2)
    irij = 1.0 / rij
    iei = 1.0 / np.sqrt(ei @ ei)
    iej = 1.0 / np.sqrt(np.einsum("ij,ij->i", ej, ej))
    eiej = iei * iej * (ej @ ei)
    a = irij * iei * (dr @ ei)
    b = irij * iej * np.einsum("ij,ij->i", dr, ej)
    ap2, am2 = (a + b) ** 2, (a - b) ** 2

    def g(w):
        return 1.0 - 0.5 * w * (ap2 / (1.0 + w * eiej) + am2 / (1.0 - w * eiej))

    eps = (1.0 / np.sqrt(1.0 - c.chi2 * eiej * eiej)) ** c.nu * g(c.chi_prime) ** c.mu
    sigma = c.sigma0 / np.sqrt(g(c.chi))
    dist = rij - sigma + c.sigma0
    cap = dist < 0.001
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        rho6 = (c.sigma0 / dist) ** 6
        u = c.four_eps0 * eps * rho6 * (rho6 - 1.0)
    return np.where(cap, 0.0, u), cap


def colloid_numpy(c, rel, e):
    u, cap = oc.colloid(c, rel, e)
    return (0.0 if cap[0] else u[0]), int(cap[0])


def min_image(d, L):
    return np.where(d > 0.5 * L, d - L, np.where(d < -0.5 * L, d + L, d))


def particle_energy(c, x, u, i, xi, ui, L):
    """finite part and cap count of particle i at (xi, ui) against all others."""
    dr = np.stack([min_image(x[d] - xi[d], L[d]) for d in range(3)], axis=1)
    mask = np.ones(x[0].shape[0], dtype=bool)
    mask[i] = False
    r2 = np.einsum("ij,ij->i", dr, dr)
    mask &= r2 < c.rc2
    if not mask.any():
        fin, cap = 0.0, 0
    else:
        uj = np.stack([u[d][mask] for d in range(3)], axis=1)
        e, cp = gb_numpy(c, np.asarray(ui, dtype=np.float64), uj, dr[mask])
        fin, cap = float(e.sum()), int(cp.sum())
    return fin, cap


def step_q(r, K):
    m = 2 * int(r) - 4294967295
    p = m * K
    return (p >> 33) if p >= 0 else -((-p) >> 33)


def i32(v):
    v = int(v) & 0xFFFFFFFF
    return v - (1 << 32) if v >= (1 << 31) else v


def u01(r):
    return (float(r) + 0.5) / TWO32


def replay_sweep(c, q, u, L, seed, replica, t, trace, info, check_tol=1e-9, fp32=False,
                 max_scale=np.inf, collect=None):
    """Replays one sweep on copies of (q, u); checks every trace record; returns final state.

    fp32=True mirrors the FP32 kernel's storage: the rotation vector and the stored orientation
    are rounded to float32; dE is checked to check_tol relative to the sum of |pair terms| and
    an accept/reject within that tolerance of the threshold follows the record.  Moves whose
    sum of |pair terms| exceeds max_scale (near-cap energies) are not checked for dE.
    collect (a list, optional) receives (record dfin, replayed FP64 dfin, record scale, dcap)
    for every move that reached the energy evaluation."""
    q = [a.astype(np.uint32).copy() for a in q]
    u = [a.astype(np.float64).copy() for a in u]
    N = q[0].shape[0]
    col = 3 if c.colors == 3 else 2
    P = col ** 3
    n = [oc.grid_dim(L[d], c.cutoff, col) for d in range(3)]
    nc = n[0] * n[1] * n[2]
    k = -(-N // nc)
    g, order, _ = oc.sweep_setup(seed, replica, t, col)
    assert list(info[:3]) == list(g) and list(info[6:6 + P]) == list(order)
    assert info[3] == k and info[4] == nc and info[5] == P
    cs, so = oc.cell_list(q, np.array(n, dtype=np.int32), g)
    key = [seed & 0xFFFFFFFF, ((seed >> 32) + replica) & 0xFFFFFFFF]
    K = [int(np.floor(c.max_displacement * TWO32 / L[d] + 0.5)) for d in range(3)]
    s = [L[d] / TWO32 for d in range(3)]
    hx, hy, hz = n[0] // col, n[1] // col, n[2] // col
    nact = hx * hy * hz
    stats = {"checked": 0, "accepted": 0}
    for p in range(P):
        sig = int(order[p])
        sx, sy, sz = sig % col, (sig // col) % col, sig // (col * col)
        for a in range(nact):
            ax, ay, az = a % hx, (a // hx) % hy, a // (hx * hy)
            cx, cy, cz = col * ax + sx, col * ay + sy, col * az + sz
            cell = cx + n[0] * (cy + n[1] * cz)
            members = so[cs[cell]:cs[cell + 1]]
            for m in range(k):
                rec = trace[p * nact * k + a * k + m]
                if len(members) == 0:
                    assert rec["pid"] == -1 and rec["outcome"] == 0
                    continue
                A = oc.philox([m, cell, t, p << 8], key)
                B = oc.philox([m, cell, t, (p << 8) | 1], key)
                i = int(members[(int(A[0]) * len(members)) >> 32])
                assert rec["pid"] == i
                qn = [(int(q[d][i]) + step_q(A[1 + d], K[d])) & 0xFFFFFFFF for d in range(3)]
                cn = [((((qn[d] - int(g[d])) & 0xFFFFFFFF) * n[d]) >> 32) for d in range(3)]
                if cn != [cx, cy, cz]:
                    assert rec["outcome"] == 1
                    continue
                rel_n = np.array([i32(qn[d] - 0x80000000) * s[d] for d in range(3)])
                if c.colloid_on and rel_n @ rel_n <= c.colloid_R2:
                    assert rec["outcome"] == 2
                    continue
                zz = 1.0 - 2.0 * u01(B[0])
                phi = 2.0 * np.pi * u01(B[1])
                sxy = np.sqrt(max(0.0, 1.0 - zz * zz))
                ui = np.array([u[d][i] for d in range(3)])
                v = np.array([sxy * np.cos(phi), sxy * np.sin(phi), zz])
                if fp32:
                    v = v.astype(np.float32).astype(np.float64)
                un = ui + c.max_rotation * v
                nn = un @ un
                un = ui.copy() if nn <= 1e-12 else un / np.sqrt(nn)
                if fp32:
                    un = un.astype(np.float32).astype(np.float64)
                x = [q[d].astype(np.float64) * s[d] for d in range(3)]
                xi = np.array([x[d][i] for d in range(3)])
                xn = np.array([qn[d] * s[d] for d in range(3)])
                fo, co = particle_energy(c, x, u, i, xi, ui, L)
                fn, cn_ = particle_energy(c, x, u, i, xn, un, L)
                dfin = fn - fo
                dccap = 0
                if c.colloid_on:
                    rel_o = np.array([i32(int(q[d][i]) - 0x80000000) * s[d] for d in range(3)])
                    eo, cco = colloid_numpy(c, rel_o, ui)
                    en, ccn = colloid_numpy(c, rel_n, un)
                    dfin += en - eo
                    dccap = ccn - cco
                assert rec["dcap"] == cn_ - co
                assert rec["dccap"] == dccap
                if collect is not None:
                    collect.append((float(rec["dfin"]), dfin, float(rec["scale"]), cn_ - co))
                if rec["scale"] <= max_scale:
                    assert abs(rec["dfin"] - dfin) <= check_tol * max(rec["scale"], 1.0), \
                        (rec["dfin"], dfin, rec["scale"])
                    stats["checked"] += 1
                dE = (cn_ - co) * 1e9 + dccap * 1e8 + dfin
                thr = -np.log(u01(B[2])) * c.temperature
                accept = dE <= thr
                if abs(dE - thr) > check_tol * max(abs(thr), rec["scale"], 1.0):
                    assert rec["outcome"] == (4 if accept else 3)
                else:
                    accept = rec["outcome"] == 4
                if accept:
                    for d in range(3):
                        q[d][i] = np.uint32(qn[d])
                        u[d][i] = un[d]
                    stats["accepted"] += 1
    return q, u, stats
