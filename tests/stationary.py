"""Exact stationary distributions of tiny systems, and the batch-means statistics that compare a
Monte Carlo chain with them.

The sweep is a Metropolis chain whose target is the Boltzmann distribution of the configuration
(positions uniform in the periodic box, orientations uniform on the sphere, weight e^{-beta U}).
For one rod near the fixed colloid, and for two Gay-Berne rods, that distribution reduces to
integrals that do not need the chain at all:

* one rod, colloid of radius R at the centre of a cubic box of half-width h: the energy depends
  on the rod's distance d from the centre and on c = |cos theta| (theta between the rod axis and
  the radial direction), and c is uniform on [0, 1] for isotropic orientations.  So
  p(d, c) = S(d) e^{-beta U(d, c)} / Z, where S(d) = d^2 Omega(h/d) is the area of the sphere of
  radius d inside the cube; Omega follows from one integral over the azimuth (below).  This is
  computed by midpoint quadrature on a fine (d, c) grid.
* two rods in a cubic box: the relative vector is uniform on the torus, so
  P(observable) = [V_out 1_out + V_ball <w 1_obs>_ball] / [V_out + V_ball <w>_ball], with
  w = e^{-beta U} and <.>_ball an average over the pair vector uniform in the cutoff ball and
  both orientations uniform on the sphere: plain importance sampling with independent draws.

The pair and colloid energies come from the oracle's FP64 functions, which are pinned against
the reference's own compiled pij_GB/GBCollRod (MCnptColloid.c:2607-2681, :1758-1854) elsewhere;
what these integrals pin is the sampling: the proposal (MCnptColloid.c:2502-2550), the
Metropolis rule (:2566-2574) and the checkerboard schedule.
"""
import numpy as np

from oracle import oracle_ctypes as oc

TWO32 = 4294967296.0


# ------------------------------------------------------------------ one rod near the colloid

def cube_solid_angle(a, nphi=20000):
    """Solid angle of the directions n with max_i |n_i| <= a (a = h/d).  Per octant, with
    n = (sin t cos f, sin t sin f, cos t) and m(f) = max(cos f, sin f), the allowed cos t lie in
    [sqrt(max(0, 1 - a^2/m^2)), min(a, 1)]."""
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    f = (np.arange(nphi) + 0.5) * (0.5 * np.pi / nphi)
    m = np.maximum(np.cos(f), np.sin(f))[None, :]
    hi = np.minimum(a, 1.0)[:, None]
    lo = np.sqrt(np.maximum(0.0, 1.0 - (a[:, None] / m) ** 2))
    return 8.0 * np.clip(hi - lo, 0.0, None).sum(axis=1) * (0.5 * np.pi / nphi)


def colloid_rod_reference(c, L, d_edges, c_edges, dd=0.002, nc=400):
    """Exact probabilities of the (d, |cos theta|) bins for one rod in a cubic box of side L
    with the colloid at its centre: array (len(d_edges)-1, len(c_edges)-1).  d_edges must start
    at or below the colloid radius and end at or beyond L sqrt(3)/2."""
    h = 0.5 * L
    d0, d1 = max(float(c.colloid_radius), 1e-9), h * np.sqrt(3.0)
    nd = int(np.ceil((d1 - d0) / dd))
    d = d0 + (np.arange(nd) + 0.5) * ((d1 - d0) / nd)
    cc = (np.arange(nc) + 0.5) / nc
    S = d ** 2 * cube_solid_angle(h / d)
    D, CC = np.meshgrid(d, cc, indexing="ij")
    rel = np.stack([np.zeros(D.size), np.zeros(D.size), D.ravel()], axis=1)
    e = np.stack([np.sqrt(1.0 - CC.ravel() ** 2), np.zeros(D.size), CC.ravel()], axis=1)
    U, cap = oc.colloid_batch(c, rel, e)
    w = np.where(cap != 0, 0.0, np.exp(-np.minimum(c.beta * U, 700.0))).reshape(D.shape)
    w = w * S[:, None]
    H = np.zeros((len(d_edges) - 1, len(c_edges) - 1))
    di = np.clip(np.searchsorted(d_edges, d, side="right") - 1, 0, len(d_edges) - 2)
    ci = np.clip(np.searchsorted(c_edges, cc, side="right") - 1, 0, len(c_edges) - 2)
    np.add.at(H, (di[:, None].repeat(nc, 1), ci[None, :].repeat(nd, 0)), w)
    return H / H.sum()


def rod_observables(q, u, L):
    """(d, |cos theta|, u) of rods from fixed-point positions q (..., 3) and orientations u
    (..., 3): d = distance of the rod centre from the box centre (minimum image), theta the
    angle between the rod axis and the radial direction."""
    rel = (q.astype(np.int64) - 2 ** 31).astype(np.float64) * (L / TWO32)
    d = np.linalg.norm(rel, axis=-1)
    un = u / np.linalg.norm(u, axis=-1, keepdims=True)
    cth = np.abs(np.sum(rel * un, axis=-1)) / np.maximum(d, 1e-300)
    return d, np.minimum(cth, 1.0), un


# ------------------------------------------------------------------ two Gay-Berne rods

def sphere_points(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


class TwoRodObservables:
    """Indicator observables of two rods: pair distance bins (the last one is r >= r_c), and,
    for pairs inside the cutoff, joint bins of (r < r_split) x |u1.u2| and bins of the pair
    energy."""

    def __init__(self, rc, r_edges, r_split, c_edges, u_edges):
        self.rc = rc
        self.r_edges = np.asarray(r_edges, dtype=np.float64)
        self.r_split = r_split
        self.c_edges = np.asarray(c_edges, dtype=np.float64)
        self.u_edges = np.asarray(u_edges, dtype=np.float64)
        self.names = ([f"r in [{a:g},{b:g})" for a, b in zip(self.r_edges[:-1], self.r_edges[1:])]
                      + [f"r >= {rc:g}"]
                      + [f"{'r<' if k == 0 else 'r>='}{r_split:g}, |c| in [{a:g},{b:g})"
                         for k in (0, 1) for a, b in zip(self.c_edges[:-1], self.c_edges[1:])]
                      + [f"U in [{a:g},{b:g})" for a, b in zip(self.u_edges[:-1], self.u_edges[1:])])

    def indicators(self, r, cosg, U):
        """Indicator matrix (n, n_obs) for samples with pair distance r, |u1.u2| cosg and pair
        energy U (U is ignored where r >= r_c)."""
        r = np.asarray(r)
        inside = r < self.rc
        nr = len(self.r_edges) - 1
        cols = []
        for k in range(nr):
            cols.append(inside & (r >= self.r_edges[k]) & (r < self.r_edges[k + 1]))
        cols.append(~inside)
        ac = np.abs(cosg)
        for near in (True, False):
            sel = inside & ((r < self.r_split) if near else (r >= self.r_split))
            for k in range(len(self.c_edges) - 1):
                hi = self.c_edges[k + 1] if k + 2 < len(self.c_edges) else np.inf
                cols.append(sel & (ac >= self.c_edges[k]) & (ac < hi))
        for k in range(len(self.u_edges) - 1):
            cols.append(inside & (U >= self.u_edges[k]) & (U < self.u_edges[k + 1]))
        return np.stack(cols, axis=-1).astype(np.float64)


def two_rod_reference(c, L, obs, M=4_000_000, seed=12345, groups=20):
    """Exact probabilities of obs's indicators for two rods in a cubic box of side L at the
    temperature of c, by importance sampling (pair vector uniform in the cutoff ball,
    orientations uniform), with their standard errors from `groups` independent groups."""
    rng = np.random.default_rng(seed)
    rc = obs.rc
    v_ball = 4.0 / 3.0 * np.pi * rc ** 3
    v_out = L ** 3 - v_ball
    n_obs = len(obs.names)
    est = []
    per = M // groups
    for _ in range(groups):
        r = rc * rng.uniform(size=per) ** (1.0 / 3.0)
        dr = sphere_points(rng, per) * r[:, None]
        e1, e2 = sphere_points(rng, per), sphere_points(rng, per)
        U, cap = oc.gb_pair_batch(c, e1, e2, dr)
        w = np.where(cap != 0, 0.0, np.exp(-np.minimum(c.beta * U, 700.0)))
        ind = obs.indicators(r, np.sum(e1 * e2, axis=1), U)
        z = v_out + v_ball * w.mean()
        p = v_ball * (w[:, None] * ind).mean(axis=0) / z
        p[len(obs.r_edges) - 1] = v_out / z            # the r >= r_c bin
        est.append(p)
    est = np.array(est)
    return est.mean(axis=0), est.std(axis=0, ddof=1) / np.sqrt(groups), n_obs


def pair_of(q, u, L):
    """Pair vector (minimum image), distance and |u1.u2| of two rods: q (..., 3, 2) uint32,
    u (..., 3, 2)."""
    dq = (q[..., 1].astype(np.int64) - q[..., 0].astype(np.int64))
    dq = ((dq + 2 ** 31) % 2 ** 32) - 2 ** 31                 # wrap-around int32 difference
    dr = dq.astype(np.float64) * (L / TWO32)
    u0 = u[..., 0] / np.linalg.norm(u[..., 0], axis=-1, keepdims=True)
    u1 = u[..., 1] / np.linalg.norm(u[..., 1], axis=-1, keepdims=True)
    return dr, np.linalg.norm(dr, axis=-1), np.sum(u0 * u1, axis=-1), u0, u1


# ------------------------------------------------------------------ chain statistics

def batch_means(x, nbatch):
    """x (n_chains, n_records, k): mean over everything and its standard error from the batch
    means of nbatch consecutive batches per chain (batches much longer than the correlation time
    are nearly independent)."""
    x = np.asarray(x, dtype=np.float64)
    nch, nrec = x.shape[:2]
    bl = nrec // nbatch
    xb = x[:, :bl * nbatch].reshape(nch, nbatch, bl, *x.shape[2:]).mean(axis=2)
    xb = xb.reshape(nch * nbatch, *x.shape[2:])
    return xb.mean(axis=0), xb.std(axis=0, ddof=1) / np.sqrt(xb.shape[0])


def zscores(x, p_ref, se_ref=None, nbatch=20):
    """z-scores of the chain means of indicators x (n_chains, n_records, k) against exact
    probabilities p_ref (k,), combining the chain's batch-means error with the reference's.
    An indicator's error is at least the independent-sample bound sqrt(p (1-p) / n), which
    autocorrelation can only raise: a rare bin that drew no sample has no batch-means spread."""
    m, se = batch_means(x, nbatch)
    p = np.asarray(p_ref, dtype=np.float64)
    n = x.shape[0] * x.shape[1]
    floor = np.where((p > 0.0) & (p < 1.0), p * (1.0 - p) / n, 0.0)
    s2 = np.maximum(se ** 2, floor) + (0.0 if se_ref is None else np.asarray(se_ref) ** 2)
    return (m - p_ref) / np.sqrt(np.maximum(s2, 1e-300)), m, np.sqrt(s2)


def describe(names, z, m, p, s):
    return "\n".join(f"  {n:>32s}: chain {a:.5f} exact {b:.5f} +- {e:.5f}  z = {zz:+.2f}"
                     for n, zz, a, b, e in zip(names, z, m, p, s))


# ------------------------------------------------------------------ the standard test systems

ROD_L = 16.5                       # 4 cells of the cutoff per axis (8 sublattices) or 6 (27)
ROD_R = 3.0
ROD_D_EDGES = np.array([3.0, 4.5, 5.0, 5.5, 6.0, 7.0, 8.25, 10.0, 11.67, 15.0])
ROD_C_EDGES = np.array([0.0, 0.25, 0.5, 0.75, 1.0001])
TWO_ROD_OBS = TwoRodObservables(4.0, [0.0, 2.0, 2.5, 3.0, 3.5, 4.0], 2.5,
                                [0.0, 0.5, 0.8, 0.95, 1.0],
                                [-np.inf, -1.5, -0.8, -0.3, -0.05, np.inf])
MAX_Z = 4.0                        # per observable; about 0.2% false alarms over ~36 of them


ROD_NAMES = [f"d [{a:g},{b:g}) |c| [{e:g},{f:g})" for a, b in zip(ROD_D_EDGES[:-1], ROD_D_EDGES[1:])
             for e, f in zip(ROD_C_EDGES[:-1], ROD_C_EDGES[1:])]
MOMENT_NAMES = ["<ux>", "<uy>", "<uz>", "<ux uy>", "<ux uz>", "<uy uz>", "<ux^2>-1/3",
                "<uy^2>-1/3", "<uz^2>-1/3"]


def rod_indicators(d, cth):
    """One-hot (d, |cos theta|) bin of every sample: (..., n_d_bins * n_c_bins)."""
    nc = len(ROD_C_EDGES) - 1
    di = np.clip(np.searchsorted(ROD_D_EDGES, d, side="right") - 1, 0, len(ROD_D_EDGES) - 2)
    ci = np.clip(np.searchsorted(ROD_C_EDGES, cth, side="right") - 1, 0, nc - 1)
    k = di * nc + ci
    X = np.zeros(np.shape(d) + ((len(ROD_D_EDGES) - 1) * nc,))
    np.put_along_axis(X, k[..., None], 1.0, axis=-1)
    return X


def orientation_moments(un):
    """<u_i>, <u_i u_j> (i < j), <u_i^2> - 1/3 per sample: all nine vanish for an isotropic
    orientation distribution (and for one rod near a spherical colloid in a cubic box, whose
    distribution is invariant under the cube's symmetries and u -> -u)."""
    return np.concatenate([un, un[..., [0, 0, 1]] * un[..., [1, 2, 2]], un ** 2 - 1.0 / 3.0],
                          axis=-1)


def chi2_limit(k, p=0.999):
    from scipy.stats import chi2
    return float(chi2.ppf(p, k))


def check(names, z, m, p, s, label):
    """Assertion message and verdict: every |z| <= MAX_Z and sum z^2 below the 0.999 quantile of
    chi^2 with len(z) degrees of freedom."""
    c2 = float(np.sum(z ** 2))
    ok = bool(np.all(np.abs(z) <= MAX_Z) and c2 <= chi2_limit(len(z)))
    msg = f"{label}: chi2 {c2:.1f} (limit {chi2_limit(len(z)):.1f}), max |z| {np.abs(z).max():.2f}\n"
    return ok, msg + describe(names, z, m, p, s)
