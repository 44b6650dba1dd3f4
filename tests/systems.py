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
