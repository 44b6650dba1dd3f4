"""ctypes binding of the FP64 CPU oracle (liboracle.so) and of the reference's own compiled
physics functions (_ref/libmcgb_ref.so, built only where the reference tree exists).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmcgb_ref.so")

TWO32 = 4294967296.0


class OrcConsts(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "sigma0", "epsilon0", "kappa", "kappa_prime", "mu", "nu",
        "temperature", "pressure", "cutoff", "max_displacement", "max_rotation",
        "colloid_radius")] + [("colloid_on", C.c_int), ("colors", C.c_int)] + [
        (n, C.c_double) for n in (
            "beta", "four_eps0", "rc2", "kappa2", "chi", "chi2", "k_imu", "chi_prime",
            "sigma_ee", "colloid_cut", "colloid_R2")]


class MoveTrace(C.Structure):
    _fields_ = [("pid", C.c_int32), ("outcome", C.c_int32), ("dcap", C.c_int32),
                ("dccap", C.c_int32), ("dfin", C.c_double), ("scale", C.c_double)]


TRACE_DTYPE = np.dtype([("pid", "<i4"), ("outcome", "<i4"), ("dcap", "<i4"), ("dccap", "<i4"),
                        ("dfin", "<f8"), ("scale", "<f8")])

_lib = None


def lib():
    """Loads liboracle.so (building it with make if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
        L = C.CDLL(ORACLE_SO)
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        ip = C.POINTER(C.c_int)
        L.orc_consts_init.argtypes = [C.POINTER(OrcConsts)]
        L.orc_gb_pair.argtypes = [C.POINTER(OrcConsts), dp, dp, dp, ip]
        L.orc_gb_pair.restype = C.c_double
        L.orc_colloid.argtypes = [C.POINTER(OrcConsts), dp, dp, ip]
        L.orc_colloid.restype = C.c_double
        L.orc_philox.argtypes = [up, up, up]
        L.orc_cell_dim.argtypes = [C.c_double, C.c_double]
        L.orc_cell_dim.restype = C.c_int
        L.orc_grid_dim.argtypes = [C.c_double, C.c_double, C.c_int]
        L.orc_grid_dim.restype = C.c_int
        L.orc_cell_list.argtypes = [C.c_int, up, up, up, ip, up, ip, ip]
        L.orc_sweep_setup.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, up, up, up]
        L.orc_sweep.argtypes = [C.POINTER(OrcConsts), C.c_int, up, up, up, dp, dp, dp, dp,
                                C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_double,
                                dp, C.POINTER(C.c_int64), C.POINTER(MoveTrace), up, C.c_int,
                                C.c_int]
        L.orc_sweep.restype = C.c_int
        L.orc_adapt_steps.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                                      C.c_int64, C.c_float, C.c_double]
        L.orc_adapt_steps.restype = None
        L.orc_total_energy.argtypes = [C.POINTER(OrcConsts), C.c_int, up, up, up, dp, dp, dp, dp,
                                       dp, C.c_int]
        L.orc_total_energy.restype = C.c_int
        L.orc_total_energy_brute.argtypes = [C.POINTER(OrcConsts), C.c_int, dp, dp, dp, dp, dp,
                                             dp, dp, dp]
        L.orc_order_parameter.argtypes = [C.c_int, dp, dp, dp, dp, dp]
        i32p = C.POINTER(C.c_int32)
        L.orc_move_draws.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32,
                                     C.c_int, C.c_int, dp, C.c_double, C.c_double, up, i32p, dp,
                                     dp]
        L.orc_move_draws.restype = None
        L.orc_gb_pair_batch.argtypes = [C.POINTER(OrcConsts), C.c_int, dp, dp, dp, dp, ip]
        L.orc_gb_pair_batch.restype = None
        L.orc_colloid_batch.argtypes = [C.POINTER(OrcConsts), C.c_int, dp, dp, dp, ip]
        L.orc_colloid_batch.restype = None
        L.orc_run_chains.argtypes = [C.POINTER(OrcConsts), C.c_int, C.c_int, up, dp, dp,
                                     C.c_uint64, C.c_uint32, C.c_int, C.c_int, up, dp, dp, C.c_int]
        L.orc_run_chains.restype = C.c_int
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_set_mutation.restype = None
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def consts(sigma0=1.0, epsilon0=1.0, kappa=3.0, kappa_prime=5.0, mu=2.0, nu=1.0,
           temperature=1.0, pressure=0.0, cutoff=4.0, max_displacement=0.1, max_rotation=0.1,
           colloid_radius=0.0, colloid_on=0, colors=2):
    c = OrcConsts(sigma0=sigma0, epsilon0=epsilon0, kappa=kappa, kappa_prime=kappa_prime, mu=mu,
                  nu=nu, temperature=temperature, pressure=pressure, cutoff=cutoff,
                  max_displacement=max_displacement, max_rotation=max_rotation,
                  colloid_radius=colloid_radius, colloid_on=colloid_on, colors=colors)
    lib().orc_consts_init(C.byref(c))
    return c


def gb_pair(c, ei, ej, dr):
    """Vectorised over rows: returns (u, capped)."""
    ei = np.ascontiguousarray(np.atleast_2d(ei), dtype=np.float64)
    ej = np.ascontiguousarray(np.atleast_2d(ej), dtype=np.float64)
    dr = np.ascontiguousarray(np.atleast_2d(dr), dtype=np.float64)
    n = dr.shape[0]
    u = np.empty(n)
    cap = np.zeros(n, dtype=np.int32)
    L = lib()
    cp = C.c_int()
    for k in range(n):
        u[k] = L.orc_gb_pair(C.byref(c), _d(ei[k]), _d(ej[k]), _d(dr[k]), C.byref(cp))
        cap[k] = cp.value
    return u, cap


def colloid(c, rel, e):
    rel = np.ascontiguousarray(np.atleast_2d(rel), dtype=np.float64)
    e = np.ascontiguousarray(np.atleast_2d(e), dtype=np.float64)
    n = rel.shape[0]
    u = np.empty(n)
    cap = np.zeros(n, dtype=np.int32)
    L = lib()
    cp = C.c_int()
    for k in range(n):
        u[k] = L.orc_colloid(C.byref(c), _d(rel[k]), _d(e[k]), C.byref(cp))
        cap[k] = cp.value
    return u, cap


def gb_pair_batch(c, ei, ej, dr):
    """orc_gb_pair over rows (one C call): returns (u, capped)."""
    ei, ej, dr = (np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64) for a in (ei, ej, dr))
    n = dr.shape[0]
    u = np.empty(n)
    cap = np.empty(n, dtype=np.int32)
    lib().orc_gb_pair_batch(C.byref(c), n, _d(ei), _d(ej), _d(dr), _d(u), _i(cap))
    return u, cap


def colloid_batch(c, rel, e):
    """orc_colloid over rows (one C call): returns (u, capped)."""
    rel, e = (np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64) for a in (rel, e))
    n = rel.shape[0]
    u = np.empty(n)
    cap = np.empty(n, dtype=np.int32)
    lib().orc_colloid_batch(C.byref(c), n, _d(rel), _d(e), _d(u), _i(cap))
    return u, cap


def move_draws(seed, replica, t, pass_, cell, m0, n, L, max_displacement, temperature):
    """Draws of moves m0..m0+n-1 of one active cell, by the sweep's own code:
    (j (n,) uint32, dq (n,3) int32, v (n,3), thr (n,))."""
    j = np.empty(n, dtype=np.uint32)
    dq = np.empty((n, 3), dtype=np.int32)
    v = np.empty((n, 3))
    thr = np.empty(n)
    Lv = np.ascontiguousarray(L, dtype=np.float64)
    lib().orc_move_draws(seed, replica, t, pass_, cell, m0, n, _d(Lv), max_displacement,
                         temperature, _u(j), dq.ctypes.data_as(C.POINTER(C.c_int32)), _d(v),
                         _d(thr))
    return j, dq, v, thr


def run_chains(c, q, u, L, seed, nsweeps, stride, t0=0, nthreads=8):
    """Independent chains of a small system (chain ch = replica ch): q (nch, 3, N) uint32,
    u (nch, 3, N), L (nch, 3).  Returns recorded (q, u, E) every `stride` sweeps:
    q (nch, nrec, 3, N), u (nch, nrec, 3, N), E (nch, nrec), and the final (q, u, L)."""
    q = np.ascontiguousarray(q, dtype=np.uint32).copy()
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    L = np.ascontiguousarray(L, dtype=np.float64).copy()
    nch, _, N = q.shape
    nrec = nsweeps // stride
    oq = np.empty((nch, nrec, 3, N), dtype=np.uint32)
    ou = np.empty((nch, nrec, 3, N))
    oe = np.empty((nch, nrec))
    rc = lib().orc_run_chains(C.byref(c), N, nch, _u(q), _d(u), _d(L), seed, t0, nsweeps, stride,
                              _u(oq), _d(ou), _d(oe), nthreads)
    if rc != 0:
        raise ValueError("orc_run_chains failed (box too small for the cell grid?)")
    return oq, ou, oe, (q, u, L)


def set_mutation(m):
    """Plant one of the oracle's ORC_MUT_* mistakes (0 removes it).  Tests only."""
    lib().orc_set_mutation(int(m))


MUT_PHI_HALF, MUT_Z_HALF, MUT_BETA, MUT_NO_TRIAL_COLLOID, MUT_STEP_DRIFT, MUT_DROP_OLD = range(1, 7)


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32)
    key = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    lib().orc_philox(_u(ctr), _u(key), _u(out))
    return out


def cell_dim(L, rc):
    return lib().orc_cell_dim(float(L), float(rc))


def grid_dim(L, rc, colors=2):
    """Cells per axis of the sweep grid for 2 or 3 sublattice colours per axis."""
    return lib().orc_grid_dim(float(L), float(rc), int(colors))


def cell_list(q, n, g):
    q = [np.ascontiguousarray(a, dtype=np.uint32) for a in q]
    N = q[0].shape[0]
    n = np.ascontiguousarray(n, dtype=np.int32)
    g = np.ascontiguousarray(g, dtype=np.uint32)
    cs = np.empty(int(np.prod(n)) + 1, dtype=np.int32)
    so = np.empty(max(N, 1), dtype=np.int32)
    lib().orc_cell_list(N, _u(q[0]), _u(q[1]), _u(q[2]), _i(n), _u(g), _i(cs), _i(so))
    return cs, so[:N]


def sweep_setup(seed, replica, t, colors=2):
    g = np.empty(3, dtype=np.uint32)
    order = np.empty(27, dtype=np.uint32)
    vol = np.empty(2, dtype=np.uint32)
    lib().orc_sweep_setup(seed, replica, t, int(colors), _u(g), _u(order), _u(vol))
    return g, order[:colors ** 3], vol


class System:
    """One replica held by the oracle: fixed-point positions q (3, N) uint32, orientations u
    (3, N) float64, box L (3,), running energy {cap_pairs, cap_colloid, finite}."""

    def __init__(self, c, q, u, L, seed=1, replica=0, npt=False, npt_max_dlnv=0.01, nthreads=1,
                 resum_interval=0, adapt_moves=0):
        self.c = c
        self.adapt_moves = adapt_moves
        if adapt_moves:                    # the controller changes the move sizes: own copy
            self.c = OrcConsts.from_buffer_copy(c)
        self.dmax = np.float32(self.c.max_displacement)
        self.drot = np.float32(self.c.max_rotation)
        self.adapts = 0
        self._ad = [0, 0]
        self.q = [np.ascontiguousarray(a, dtype=np.uint32).copy() for a in q]
        self.u = [np.ascontiguousarray(a, dtype=np.float64).copy() for a in u]
        self.L = np.ascontiguousarray(L, dtype=np.float64).copy()
        self.N = self.q[0].shape[0]
        self.seed, self.replica = seed, replica
        self.npt, self.npt_max_dlnv = npt, npt_max_dlnv
        self.nthreads = nthreads
        self.resum_interval = resum_interval
        self.t = 0
        self.counters = np.zeros(9, dtype=np.int64)
        e = self.total_energy()
        self.energy = np.array(e[:3], dtype=np.float64)

    def total_energy(self):
        out = np.empty(4)
        rc = lib().orc_total_energy(C.byref(self.c), self.N, _u(self.q[0]), _u(self.q[1]),
                                    _u(self.q[2]), _d(self.u[0]), _d(self.u[1]), _d(self.u[2]),
                                    _d(self.L), _d(out), self.nthreads)
        if rc != 0:
            raise ValueError("box too small for the cell grid")
        return out

    def energy_value(self, e=None):
        e = self.energy if e is None else e
        return e[0] * 1e9 + e[1] * 1e8 + e[2]

    def positions(self):
        return [self.q[d].astype(np.float64) * (self.L[d] / TWO32) for d in range(3)]

    def sweep(self, trace=False):
        col = 3 if self.c.colors == 3 else 2
        nc = int(np.prod([grid_dim(self.L[d], self.c.cutoff, col) for d in range(3)]))
        if nc == 0:
            raise ValueError("box too small for the cell grid")
        k = -(-self.N // nc)
        tr = np.zeros(nc * k, dtype=TRACE_DTYPE) if trace else None
        info = np.zeros(36, dtype=np.uint32)
        cnt = np.zeros(9, dtype=np.int64)
        rc = lib().orc_sweep(
            C.byref(self.c), self.N, _u(self.q[0]), _u(self.q[1]), _u(self.q[2]),
            _d(self.u[0]), _d(self.u[1]), _d(self.u[2]), _d(self.L),
            self.seed, self.replica, self.t, 1 if self.npt else 0, self.npt_max_dlnv,
            _d(self.energy), cnt.ctypes.data_as(C.POINTER(C.c_int64)),
            tr.ctypes.data_as(C.POINTER(MoveTrace)) if trace else None, _u(info),
            1 if (self.resum_interval and (self.t + 1) % self.resum_interval == 0) else 0,
            self.nthreads)
        if rc != 0:
            raise ValueError("box too small for the cell grid")
        self.t += 1
        self.counters += cnt
        if self.adapt_moves:
            self._ad[0] += int(cnt[0])
            self._ad[1] += int(cnt[1])
            if self._ad[0] >= self.adapt_moves:
                d, r = C.c_float(self.dmax), C.c_float(self.drot)
                s_ee = self.c.sigma_ee
                min_thick = np.float32(min(s_ee, self.c.sigma0))
                lib().orc_adapt_steps(C.byref(d), C.byref(r), self._ad[1], self._ad[0],
                                      float(min_thick), 0.5 * float(self.L.min()))
                self.dmax, self.drot = np.float32(d.value), np.float32(r.value)
                self.c.max_displacement = float(self.dmax)
                self.c.max_rotation = float(self.drot)
                self._ad = [0, 0]
                self.adapts += 1
        return (tr, info, cnt)

    def order_parameter(self):
        p2 = C.c_double()
        d = np.empty(3)
        lib().orc_order_parameter(self.N, _d(self.u[0]), _d(self.u[1]), _d(self.u[2]),
                                  C.byref(p2), _d(d))
        return p2.value, d


def to_fixed(x, L):
    """Positions (any real) -> 32-bit fixed-point fractions of the box."""
    f = np.asarray(x, dtype=np.float64) / L
    f = f - np.floor(f)
    q = np.floor(f * TWO32)
    q = np.where(q >= TWO32, 0.0, q)
    return q.astype(np.uint64).astype(np.uint32)


def ref_lib():
    """The reference's own physics functions (None when the reference tree was not built)."""
    if not os.path.exists(REF_SO):
        return None
    L = C.CDLL(REF_SO)
    dp = C.POINTER(C.c_double)
    L.pij_GB.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, dp]
    L.pij_GB.restype = C.c_double
    L.GBCollRod.argtypes = [C.c_int, dp, dp, dp]
    L.GBCollRod.restype = C.c_double
    L.get_system_const.argtypes = []
    L.ref_compute_distance_sq.argtypes = [C.c_double] * 6 + [dp]
    L.ref_compute_distance_sq.restype = C.c_double
    L.ref_g_omega.argtypes = [C.c_double] * 4
    L.ref_g_omega.restype = C.c_double
    fp = C.POINTER(C.c_float)
    L.compute_liquid_crystal_order_parameters.argtypes = [C.c_int, dp, dp, dp, fp, fp, fp, fp]
    L.load_initial_conditions.argtypes = [C.c_char_p]
    return L


def ref_set(L, name, value, ctype=C.c_double):
    ctype.in_dll(L, name).value = value


def ref_get(L, name, ctype=C.c_double):
    return ctype.in_dll(L, name).value
