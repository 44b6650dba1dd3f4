"""Seeded synthetic configurations for the tests (lattice starts, as in the benchmark configs)."""
import numpy as np

from oracle import oracle_ctypes as oc


def lattice_system(N, rho, colloid_radius=0.0, exclusion=None, jitter=0.05, seed=0):
    """Cubic box V = N/rho + 4/3 pi R^3; n_s^3 jittered sites (minus those within `exclusion`
    of the centre); N of them chosen at random; orientations uniform on the sphere.
    Returns (q (3,N) uint32, u (3,N) float64, L (3,))."""
    rng = np.random.default_rng(seed)
    V = N / rho + (4.0 / 3.0) * np.pi * colloid_radius ** 3
    L = V ** (1.0 / 3.0)
    excl = (colloid_radius + 0.5) if exclusion is None else exclusion
    ns = int(np.ceil(N ** (1.0 / 3.0)))
    while True:
        a = L / ns
        g = (np.arange(ns) + 0.5) * a
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        sites = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        if colloid_radius > 0:
            d = np.linalg.norm(sites - 0.5 * L, axis=1)
            sites = sites[d >= excl]
        if sites.shape[0] >= N:
            break
        ns += 1
    sites = sites[rng.permutation(sites.shape[0])[:N]]
    sites = sites + rng.uniform(-jitter, jitter, size=sites.shape)
    u = rng.normal(size=(N, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    Lv = np.array([L, L, L])
    q = [oc.to_fixed(sites[:, d], L) for d in range(3)]
    return q, [np.ascontiguousarray(u[:, d]) for d in range(3)], Lv


def layered_system(N, rho, colloid_radius=0.0, seed=0, tilt=0.1, layer=3.25):
    """Near-nematic layered start with moderate energies (no capped or near-capped pairs):
    rods along z with a random tilt (std `tilt` rad, random head/tail) on square layers spaced
    `layer` apart in z; sites within colloid_radius + 2 of the centre removed.
    Returns (q (3,N) uint32, u (3,N) float64, L (3,))."""
    rng = np.random.default_rng(seed)
    V = N / rho + (4.0 / 3.0) * np.pi * colloid_radius ** 3
    L = V ** (1.0 / 3.0)
    nz = max(1, int(L // layer))
    excl = colloid_radius + 2.0 if colloid_radius > 0 else 0.0
    nx = int(np.ceil(np.sqrt(N / nz)))
    while True:
        a, c = L / nx, L / nz
        gx = (np.arange(nx) + 0.5) * a
        gz = (np.arange(nz) + 0.5) * c
        X, Y, Zg = np.meshgrid(gx, gx, gz, indexing="ij")
        sites = np.stack([X.ravel(), Y.ravel(), Zg.ravel()], axis=1)
        if colloid_radius > 0:
            sites = sites[np.linalg.norm(sites - 0.5 * L, axis=1) >= excl]
        if sites.shape[0] >= N:
            break
        nx += 1
    sites = sites[np.sort(rng.permutation(sites.shape[0])[:N])]
    sites[:, :2] += rng.uniform(-0.03, 0.03, size=(N, 2))
    u = np.zeros((N, 3))
    u[:, 2] = 1.0
    u[:, :2] += rng.normal(scale=tilt, size=(N, 2))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u *= np.where(rng.uniform(size=(N, 1)) < 0.5, -1.0, 1.0)
    Lv = np.array([L, L, L])
    q = [oc.to_fixed(sites[:, d], L) for d in range(3)]
    return q, [np.ascontiguousarray(u[:, d]) for d in range(3)], Lv


def box_system(N, rho, aspect, colloid_radius=0.0, seed=0, tilt=0.1, layer=3.25):
    """The layered start in a rectangular box: volume N/rho + 4/3 pi R^3, edge lengths in the
    ratios `aspect` (so every axis has its own cell count and fixed-point scale); square-ish
    layers spaced at least `layer` apart in z, sites within R + 2 of the centre removed.
    Returns (q (3,N) uint32, u (3,N) float64, L (3,))."""
    rng = np.random.default_rng(seed)
    V = N / rho + (4.0 / 3.0) * np.pi * colloid_radius ** 3
    a = np.asarray(aspect, dtype=np.float64)
    Lv = a * (V / np.prod(a)) ** (1.0 / 3.0)
    nz = max(1, int(Lv[2] // layer))
    excl = colloid_radius + 2.0 if colloid_radius > 0 else 0.0
    per_layer = N / nz
    nx = int(np.ceil(np.sqrt(per_layer * Lv[0] / Lv[1])))
    ny = int(np.ceil(np.sqrt(per_layer * Lv[1] / Lv[0])))
    while True:
        gx = (np.arange(nx) + 0.5) * (Lv[0] / nx)
        gy = (np.arange(ny) + 0.5) * (Lv[1] / ny)
        gz = (np.arange(nz) + 0.5) * (Lv[2] / nz)
        X, Y, Zg = np.meshgrid(gx, gy, gz, indexing="ij")
        sites = np.stack([X.ravel(), Y.ravel(), Zg.ravel()], axis=1)
        if colloid_radius > 0:
            sites = sites[np.linalg.norm(sites - 0.5 * Lv, axis=1) >= excl]
        if sites.shape[0] >= N:
            break
        nx += 1
        ny += 1
    sites = sites[np.sort(rng.permutation(sites.shape[0])[:N])]
    sites[:, :2] += rng.uniform(-0.03, 0.03, size=(N, 2))
    u = np.zeros((N, 3))
    u[:, 2] = 1.0
    u[:, :2] += rng.normal(scale=tilt, size=(N, 2))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u *= np.where(rng.uniform(size=(N, 1)) < 0.5, -1.0, 1.0)
    q = [oc.to_fixed(sites[:, d], Lv[d]) for d in range(3)]
    return q, [np.ascontiguousarray(u[:, d]) for d in range(3)], Lv


def cond_tol(base, scale):
    """Tolerance of a summed energy difference whose largest term may sit near the overlap cap:
    a pair at r - sigma + sigma0 = sigma0/rho carries a relative condition number of about
    12 (r + sigma) rho / sigma0 <= 96 rho (r < 4, sigma < 4), and rho is bounded through
    4 eps rho^12 <= sum |U| = scale (eps >= 0.38 for GB(3,5,2,1))."""
    rho_max = np.maximum(1.0, (np.asarray(scale) / 1.5) ** (1.0 / 12.0))
    return base * np.maximum(np.asarray(scale), 1.0) * (1.0 + 96.0 * rho_max)
