"""ctypes binding of libmcgb.so (include/mcgb.h) for the tests and the benchmark.

The compute path is the HIP library; this module only marshals arrays.  Loading fails loudly
when libmcgb.so has not been built (python -m parellel_mc_mpi_for_liquidcrystalsystems_amd.build).
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmcgb.so")

ERRORS = {0: "OK", -1: "ARG", -2: "IO", -3: "PARAM", -4: "HIP", -5: "NOMEM", -6: "GEOMETRY",
          -7: "STATE", -8: "CAPACITY"}
globals().update({"ERR_" + v: k for k, v in ERRORS.items() if k < 0})
REQUIRED_KEYS = [
    "mc_ensemble_type", "num_particles", "system_density", "system_temperature",
    "system_pressure", "gb_sigma_sidetoside", "gb_epsilon0", "gb_kappa",
    "gb_kappa_prime", "gb_mu_exponent", "gb_nu_exponent", "max_displacement",
    "max_rotation", "initial_configuration_type", "cutoff_radius",
    "verlet_radius", "total_mc_cycles", "configuration_save_interval",
    "statistics_print_interval", "colloid_wall_option", "colloid_interaction_epsilon",
    "colloid_interaction_amplitude", "colloid_radius"]


class McgbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mcgb error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("mc_ensemble_type", C.c_int), ("num_particles", C.c_int),
        ("system_density", C.c_double), ("system_temperature", C.c_double),
        ("system_pressure", C.c_double), ("gb_sigma_sidetoside", C.c_double),
        ("gb_epsilon0", C.c_double), ("gb_kappa", C.c_double), ("gb_kappa_prime", C.c_double),
        ("gb_mu_exponent", C.c_double), ("gb_nu_exponent", C.c_double),
        ("max_displacement", C.c_float), ("max_rotation", C.c_float),
        ("initial_configuration_type", C.c_int), ("cutoff_radius", C.c_double),
        ("verlet_radius", C.c_double), ("total_mc_cycles", C.c_int),
        ("configuration_save_interval", C.c_int), ("statistics_print_interval", C.c_int),
        ("colloid_wall_option", C.c_int), ("colloid_interaction_epsilon", C.c_double),
        ("colloid_interaction_amplitude", C.c_double), ("colloid_radius", C.c_double),
        ("seed", C.c_uint64), ("precision", C.c_int), ("npt_max_dlnv", C.c_double),
        ("energy_check_interval", C.c_int), ("lattice_jitter", C.c_double),
        ("colloid_exclusion_radius", C.c_double), ("stage_capacity", C.c_int),
        ("waves_per_cell", C.c_int), ("pass_schedule", C.c_int), ("adapt_moves", C.c_int),
        ("sublattice_colors", C.c_int), ("keys_read", C.c_uint32)]


class MoveTrace(C.Structure):
    _fields_ = [("pid", C.c_int32), ("outcome", C.c_int32), ("dcap", C.c_int32),
                ("dccap", C.c_int32), ("dfin", C.c_double), ("scale", C.c_double)]


TRACE_DTYPE = np.dtype([("pid", "<i4"), ("outcome", "<i4"), ("dcap", "<i4"), ("dccap", "<i4"),
                        ("dfin", "<f8"), ("scale", "<f8")])


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "sweeps", "attempted", "accepted", "rejected_left_cell", "rejected_hard_core",
        "candidates", "pairs", "colloid_evals", "volume_attempted", "volume_accepted")] + [
        ("cap_pairs", C.c_int64), ("cap_colloid", C.c_int64), ("energy_finite", C.c_double),
        ("energy", C.c_double), ("box", C.c_double * 3), ("volume", C.c_double),
        ("density", C.c_double), ("max_drift", C.c_double), ("resummations", C.c_uint64),
        ("da_stage1_rejected", C.c_uint64), ("da_stage2_rejected", C.c_uint64),
        ("max_drift_forced", C.c_double), ("resummations_forced", C.c_uint64)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["box"] = list(self.box)
        return d


class Resume(C.Structure):
    """Everything beyond the configuration that a replica's chain depends on: the Philox key
    (seed), the Philox counter (sweeps done), the step sizes, the adaptation counters and the
    running energy (capped-term counts and finite part)."""
    _fields_ = [("seed", C.c_uint64), ("sweep_index", C.c_uint64),
                ("max_displacement", C.c_float), ("max_rotation", C.c_float),
                ("adapt_attempted", C.c_uint64), ("adapt_accepted", C.c_uint64),
                ("adaptations", C.c_uint64), ("capped_pairs", C.c_int64),
                ("capped_colloid", C.c_int64), ("energy_finite", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None
dp = C.POINTER(C.c_double)
up = C.POINTER(C.c_uint32)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)
u64p = C.POINTER(C.c_uint64)
llp = C.POINTER(C.c_longlong)
SUR_NPARAM = 4545
DA_BINS = 66


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run "
                          "`python -m parellel_mc_mpi_for_liquidcrystalsystems_amd.build`")
    L = C.CDLL(LIB_PATH)
    ctx = C.c_void_p
    sig = {
        "mcgb_params_default": (None, [C.POINTER(Params)]),
        "mcgb_params_load": (C.c_int, [C.c_char_p, C.POINTER(Params)]),
        "mcgb_params_derived": (C.c_int, [C.POINTER(Params), dp]),
        "mcgb_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(ctx)]),
        "mcgb_destroy": (None, [ctx]),
        "mcgb_last_error": (C.c_char_p, []),
        "mcgb_device_count": (C.c_int, []),
        "mcgb_set_state": (C.c_int, [ctx, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_set_state_fixed": (C.c_int, [ctx, C.c_int, C.c_int, up, up, up, dp, dp, dp, dp]),
        "mcgb_get_state": (C.c_int, [ctx, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_get_state_fixed": (C.c_int, [ctx, C.c_int, C.c_int, up, up, up, dp, dp, dp, dp]),
        "mcgb_sweep": (C.c_int, [ctx, C.c_int]),
        "mcgb_sweep_traced": (C.c_int, [ctx, C.POINTER(MoveTrace), C.c_int, up]),
        "mcgb_get_stats": (C.c_int, [ctx, C.c_int, C.POINTER(Stats)]),
        "mcgb_reset_counters": (C.c_int, [ctx, C.c_int]),
        "mcgb_total_energy": (C.c_int, [ctx, C.c_int, dp]),
        "mcgb_resum_energy": (C.c_int, [ctx, C.c_int, dp]),
        "mcgb_order_parameter": (C.c_int, [ctx, C.c_int, dp, dp]),
        "mcgb_series": (C.c_int, [ctx, C.c_int, C.c_int, dp, dp, dp]),
        "mcgb_series_clear": (C.c_int, [ctx]),
        "mcgb_kernel_times": (C.c_int, [ctx, dp, dp, dp]),
        "mcgb_autotune": (C.c_int, [ctx, C.c_int, ip, ip, dp]),
        "mcgb_set_pace": (C.c_int, [ctx, C.c_int]),
        "mcgb_get_pace": (C.c_int, [ctx, ip]),
        "mcgb_cell_list": (C.c_int, [ctx, C.c_int, C.c_uint32, ip, C.c_int, ip, C.c_int, up, up, ip]),
        "mcgb_pair_energy_batch": (C.c_int, [C.POINTER(Params), C.c_int, dp, dp, dp, C.c_int, dp, ip]),
        "mcgb_colloid_energy_batch": (C.c_int, [C.POINTER(Params), C.c_int, dp, dp, C.c_int, dp, ip]),
        "mcgb_pair_energy_batch_noalloc": (C.c_int, [C.POINTER(Params), C.c_int, dp, dp, dp, C.c_int,
                                                     dp, ip]),
        "mcgb_philox_batch": (C.c_int, [C.c_int, up, up, up]),
        "mcgb_sweep_draws": (C.c_int, [ctx, C.c_int, C.c_int, up, C.POINTER(C.c_int32), dp, dp,
                                       up]),
        "mcgb_init_lattice": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64, dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_tau_int": (C.c_int, [dp, C.c_int, C.c_double, dp, ip]),
        "mcgb_get_step_sizes": (C.c_int, [ctx, C.c_int, fp, fp, u64p]),
        "mcgb_sweep_async": (C.c_int, [ctx, C.c_int]),
        "mcgb_set_kernel_timing": (C.c_int, [ctx, C.c_int]),
        "mcgb_sync": (C.c_int, [ctx]),
        "mcgb_snapshot_begin": (C.c_int, [ctx, C.c_int]),
        "mcgb_snapshot_wait": (C.c_int, [ctx, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_set_step_sizes": (C.c_int, [ctx, C.c_int, C.c_float, C.c_float]),
        "mcgb_init_random": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                       C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_init_cubic": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_uint64,
                                      dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_write_trajectory": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_write_qmga": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_write_xyz": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, dp, dp, dp]),
        "mcgb_eigen_sym3": (C.c_int, [dp, dp, dp]),
        "mcgb_write_fort1": (C.c_int, [C.c_char_p, C.c_int, dp, dp, dp, dp, dp, dp, dp, C.c_double]),
        "mcgb_read_fort1": (C.c_int, [C.c_char_p, C.c_int, dp, dp, dp, dp, dp, dp, dp, dp]),
        "mcgb_get_resume": (C.c_int, [ctx, C.c_int, C.POINTER(Resume)]),
        "mcgb_resume_state": (C.c_int, [ctx, C.c_int, C.c_int, up, up, up, dp, dp, dp, dp,
                                        C.POINTER(Resume)]),
        "mcgb_write_resume": (C.c_int, [C.c_char_p, C.c_int, up, up, up, dp, dp, dp, dp,
                                        C.POINTER(Resume)]),
        "mcgb_read_resume": (C.c_int, [C.c_char_p, C.c_int, up, up, up, dp, dp, dp, dp,
                                       C.POINTER(Resume)]),
        "mcgb_surrogate_set": (C.c_int, [ctx, fp]),
        "mcgb_surrogate_get": (C.c_int, [ctx, fp]),
        "mcgb_surrogate_init": (C.c_int, [ctx, C.c_uint32]),
        "mcgb_surrogate_eval": (C.c_int, [ctx, C.c_int, fp, C.c_int, fp]),
        "mcgb_surrogate_grad": (C.c_int, [ctx, fp, C.c_int, fp, fp, C.c_float, fp, dp]),
        "mcgb_surrogate_train": (C.c_int, [ctx, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                           C.c_uint32, fp]),
        "mcgb_pairs_reserve": (C.c_int, [ctx, C.c_longlong]),
        "mcgb_pairs_clear": (C.c_int, [ctx]),
        "mcgb_pairs_count": (C.c_int, [ctx, llp]),
        "mcgb_pairs_get": (C.c_int, [ctx, C.c_longlong, C.c_int, fp]),
        "mcgb_pairs_put": (C.c_int, [ctx, C.c_int, fp]),
        "mcgb_emit_pairs": (C.c_int, [ctx, C.c_int, C.c_int, C.c_int, llp]),
        "mcgb_da_enable": (C.c_int, [ctx, C.c_int, C.c_int]),
        "mcgb_da_histogram": (C.c_int, [ctx, C.c_int, u64p, C.c_int]),
        "mcgb_da_guard": (C.c_int, [ctx, C.c_int, C.c_double, C.c_double]),
        "mcgb_da_status": (C.c_int, [ctx, C.POINTER(C.c_int), u64p, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc is not None and rc < 0:
        raise McgbError(rc, lib().mcgb_last_error().decode(errors="replace"))
    return rc


def _d(a):
    return a.ctypes.data_as(dp)


def _u(a):
    return a.ctypes.data_as(up)


def _i(a):
    return a.ctypes.data_as(ip)


def _f(a):
    return a.ctypes.data_as(fp)


def default_params(**kw):
    p = Params()
    lib().mcgb_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def gb_params(N, density, temperature, colloid_radius=0.0, precision=32, seed=1, npt=False,
              pressure=0.0, kappa=3.0, kappa_prime=5.0, mu=2.0, nu=1.0, cutoff=4.0,
              max_displacement=0.1, max_rotation=0.1, **kw):
    """Parameter block of the benchmark family GB(kappa, kappa', mu, nu), sigma0 = eps0 = 1."""
    return default_params(
        mc_ensemble_type=2 if npt else 1, num_particles=N, system_density=density,
        system_temperature=temperature, system_pressure=pressure, gb_sigma_sidetoside=1.0,
        gb_epsilon0=1.0, gb_kappa=kappa, gb_kappa_prime=kappa_prime, gb_mu_exponent=mu,
        gb_nu_exponent=nu, max_displacement=max_displacement, max_rotation=max_rotation,
        initial_configuration_type=1, cutoff_radius=cutoff, verlet_radius=cutoff + 1.0,
        total_mc_cycles=0, configuration_save_interval=0, statistics_print_interval=0,
        colloid_wall_option=1 if colloid_radius > 0 else 0, colloid_interaction_epsilon=0.0,
        colloid_interaction_amplitude=0.0, colloid_radius=colloid_radius, seed=seed,
        precision=precision, keys_read=(1 << 23) - 1, **kw)


def load_params(path):
    p = Params()
    check(lib().mcgb_params_load(path.encode(), C.byref(p)))
    return p


def derived(p):
    out = np.empty(11)
    check(lib().mcgb_params_derived(C.byref(p), _d(out)))
    return out


def device_count():
    return lib().mcgb_device_count()


def init_lattice(N, density, colloid_radius=0.0, exclusion_radius=-1.0, jitter=0.0, seed=1):
    arrs = [np.empty(N) for _ in range(6)]
    box = np.empty(3)
    check(lib().mcgb_init_lattice(N, density, colloid_radius, exclusion_radius, jitter, seed,
                                  *(_d(a) for a in arrs), _d(box)))
    return arrs[:3], arrs[3:], box


def init_random(N, density, colloid_radius=0.0, sigma_ee=3.0, seed=1, max_tries=1000000):
    arrs = [np.empty(N) for _ in range(6)]
    box = np.empty(3)
    check(lib().mcgb_init_random(N, density, colloid_radius, sigma_ee, seed, max_tries,
                                 *(_d(a) for a in arrs), _d(box)))
    return arrs[:3], arrs[3:], box


def init_cubic(N, density, colloid_radius=0.0, seed=1):
    arrs = [np.empty(N) for _ in range(6)]
    box = np.empty(3)
    check(lib().mcgb_init_cubic(N, density, colloid_radius, seed, *(_d(a) for a in arrs), _d(box)))
    return arrs[:3], arrs[3:], box


def write_trajectory(path, pos, u, box, append=False):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in list(pos) + list(u)]
    b = np.ascontiguousarray(box, dtype=np.float64)
    check(lib().mcgb_write_trajectory(os.fsencode(path), 1 if append else 0, a[0].shape[0],
                                      *(_d(v) for v in a), _d(b)))


def write_qmga(path, pos, u, box, append=False):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in list(pos) + list(u)]
    b = np.ascontiguousarray(box, dtype=np.float64)
    check(lib().mcgb_write_qmga(os.fsencode(path), 1 if append else 0, a[0].shape[0],
                                *(_d(v) for v in a), _d(b)))


def write_xyz(path, pos, box, append=False):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in pos]
    b = np.ascontiguousarray(box, dtype=np.float64)
    check(lib().mcgb_write_xyz(os.fsencode(path), 1 if append else 0, a[0].shape[0],
                               *(_d(v) for v in a), _d(b)))


def write_resume(path, q, u, box, record):
    """Write a resume file: fixed-point positions, orientations and box as exact values, and
    the record (a Resume)."""
    q = [np.ascontiguousarray(a, dtype=np.uint32) for a in q]
    u = [np.ascontiguousarray(a, dtype=np.float64) for a in u]
    box = np.ascontiguousarray(box, dtype=np.float64)
    check(lib().mcgb_write_resume(os.fsencode(path), q[0].shape[0], *(_u(a) for a in q),
                                  *(_d(a) for a in u), _d(box), C.byref(record)))


def read_resume(path, n):
    """Read a resume file of n particles: (q, u, box, record)."""
    q = [np.empty(n, dtype=np.uint32) for _ in range(3)]
    u = [np.empty(n) for _ in range(3)]
    box = np.empty(3)
    rec = Resume()
    check(lib().mcgb_read_resume(os.fsencode(path), n, *(_u(a) for a in q), *(_d(a) for a in u),
                                 _d(box), C.byref(rec)))
    return q, u, box, rec


def tau_int(x, c=5.0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    tau = C.c_double()
    w = C.c_int()
    check(lib().mcgb_tau_int(_d(x), x.shape[0], c, C.byref(tau), C.byref(w)))
    return tau.value, w.value


def eigen_sym3(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    ev = np.empty(3)
    V = np.empty(9)
    check(lib().mcgb_eigen_sym3(_d(q), _d(ev), _d(V)))
    return ev, V.reshape(3, 3)


def pair_energy_batch(p, ei, ej, dr, precision):
    ei, ej, dr = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3) for a in (ei, ej, dr))
    n = ei.shape[0]
    u = np.empty(n)
    cap = np.empty(n, dtype=np.int32)
    check(lib().mcgb_pair_energy_batch(C.byref(p), n, _d(ei), _d(ej), _d(dr), precision, _d(u), _i(cap)))
    return u, cap


NOALLOC_MAX_PAIRS = 65536


def pair_energy_batch_noalloc(p, ei, ej, dr, precision):
    """mcgb_pair_energy_batch_noalloc: the pair batch with its device staging in the library's
    static array (at most NOALLOC_MAX_PAIRS pairs per call)."""
    ei, ej, dr = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3) for a in (ei, ej, dr))
    n = ei.shape[0]
    u = np.empty(n)
    cap = np.empty(n, dtype=np.int32)
    check(lib().mcgb_pair_energy_batch_noalloc(C.byref(p), n, _d(ei), _d(ej), _d(dr), precision,
                                               _d(u), _i(cap)))
    return u, cap


def colloid_energy_batch(p, rel, e, precision=64):
    rel, e = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3) for a in (rel, e))
    n = rel.shape[0]
    u = np.empty(n)
    cap = np.empty(n, dtype=np.int32)
    check(lib().mcgb_colloid_energy_batch(C.byref(p), n, _d(rel), _d(e), precision, _d(u), _i(cap)))
    return u, cap


def philox_batch(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.empty_like(ctr)
    check(lib().mcgb_philox_batch(ctr.shape[0], _u(ctr), _u(key), _u(out)))
    return out


class Context:
    """One GPU context holding one or more replicas (mcgb_create/mcgb_destroy)."""

    def __init__(self, params):
        if isinstance(params, Params):
            params = [params]
        arr = (Params * len(params))(*params)
        self.params = list(params)
        self.nrep = len(params)
        h = C.c_void_p()
        check(lib().mcgb_create(arr, self.nrep, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().mcgb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, pos, u, box, replica=0):
        pos = [np.ascontiguousarray(a, dtype=np.float64) for a in pos]
        u = [np.ascontiguousarray(a, dtype=np.float64) for a in u]
        box = np.ascontiguousarray(box, dtype=np.float64)
        check(lib().mcgb_set_state(self.h, replica, pos[0].shape[0], *(_d(a) for a in pos),
                                   *(_d(a) for a in u), _d(box)))

    def set_state_fixed(self, q, u, box, replica=0):
        q = [np.ascontiguousarray(a, dtype=np.uint32) for a in q]
        u = [np.ascontiguousarray(a, dtype=np.float64) for a in u]
        box = np.ascontiguousarray(box, dtype=np.float64)
        check(lib().mcgb_set_state_fixed(self.h, replica, q[0].shape[0], *(_u(a) for a in q),
                                         *(_d(a) for a in u), _d(box)))

    def get_state(self, replica=0):
        n = self.params[replica].num_particles
        pos = [np.empty(n) for _ in range(3)]
        u = [np.empty(n) for _ in range(3)]
        box = np.empty(3)
        check(lib().mcgb_get_state(self.h, replica, n, *(_d(a) for a in pos), *(_d(a) for a in u), _d(box)))
        return pos, u, box

    def get_state_fixed(self, replica=0):
        n = self.params[replica].num_particles
        q = [np.empty(n, dtype=np.uint32) for _ in range(3)]
        u = [np.empty(n) for _ in range(3)]
        box = np.empty(3)
        check(lib().mcgb_get_state_fixed(self.h, replica, n, *(_u(a) for a in q), *(_d(a) for a in u),
                                         _d(box)))
        return q, u, box

    def sweep(self, n=1):
        check(lib().mcgb_sweep(self.h, n))

    def sweep_traced(self):
        p = self.params[0]
        cap = p.num_particles + 4 * p.num_particles + 100000
        buf = (MoveTrace * cap)()
        info = np.zeros(36, dtype=np.uint32)
        check(lib().mcgb_sweep_traced(self.h, buf, cap, _u(info)))
        nrec = int(info[3]) * int(info[4])
        arr = np.frombuffer(buf, dtype=TRACE_DTYPE, count=nrec).copy()
        return arr, info

    def sweep_draws(self, replica=0):
        """Draws of the last sweep: (j, dq (n,3) int32, v (n,3), thr, info {t, k, nact, P})."""
        p = self.params[replica]
        cap = 4 * p.num_particles + 200000
        j = np.empty(cap, dtype=np.uint32)
        dq = np.empty((cap, 3), dtype=np.int32)
        v = np.empty((cap, 3))
        thr = np.empty(cap)
        info = np.zeros(4, dtype=np.uint32)
        n = check(lib().mcgb_sweep_draws(self.h, replica, cap, _u(j),
                                         dq.ctypes.data_as(C.POINTER(C.c_int32)), _d(v), _d(thr),
                                         _u(info)))
        return j[:n].copy(), dq[:n].copy(), v[:n].copy(), thr[:n].copy(), info

    def stats(self, replica=0):
        s = Stats()
        check(lib().mcgb_get_stats(self.h, replica, C.byref(s)))
        return s.as_dict()

    def reset_counters(self, replica=0):
        check(lib().mcgb_reset_counters(self.h, replica))

    def total_energy(self, replica=0):
        out = np.empty(4)
        check(lib().mcgb_total_energy(self.h, replica, _d(out)))
        return out

    def resum_energy(self, replica=0):
        d = C.c_double()
        check(lib().mcgb_resum_energy(self.h, replica, C.byref(d)))
        return d.value

    def order_parameter(self, replica=0):
        p2 = C.c_double()
        d = np.empty(3)
        check(lib().mcgb_order_parameter(self.h, replica, C.byref(p2), _d(d)))
        return p2.value, d

    def series(self, replica=0, max_n=1 << 16):
        e, p2, rho = np.empty(max_n), np.empty(max_n), np.empty(max_n)
        n = check(lib().mcgb_series(self.h, replica, max_n, _d(e), _d(p2), _d(rho)))
        return e[:n], p2[:n], rho[:n]

    def series_clear(self):
        check(lib().mcgb_series_clear(self.h))

    def autotune(self, sweeps=4):
        """Returns (waves_per_cell, schedule, {"seq/W", "spec/W", "seq4/W" or "duo/W": ms per
        pass}; 0 for a combination not run).  With delayed acceptance on, every schedule runs
        the delayed-acceptance kernel, so it tunes W for that kernel."""
        best, sched = C.c_int(), C.c_int()
        ms = np.empty(20)
        check(lib().mcgb_autotune(self.h, sweeps, C.byref(best), C.byref(sched), _d(ms)))
        keys = [f"{s}/{w}" for s in ("seq", "spec", "seq4", "duo") for w in (1, 2, 3, 4, 8)]
        return best.value, sched.value, dict(zip(keys, ms.tolist()))

    def set_pace(self, percent):
        """Sequential-pass pacing in percent of the last sweep's mean move time (0: off)."""
        check(lib().mcgb_set_pace(self.h, int(percent)))

    def pace(self):
        p = C.c_int()
        check(lib().mcgb_get_pace(self.h, C.byref(p)))
        return p.value

    def set_kernel_timing(self, on=True):
        check(lib().mcgb_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(lib().mcgb_kernel_times(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"ms_pass": a.value, "ms_rebuild": b.value, "ms_sweep": c.value}

    def cell_list(self, t, replica=0):
        p = self.params[replica]
        n = p.num_particles
        cs = np.empty(2_000_000 if n > 100000 else 200000, dtype=np.int32)
        ids = np.empty(n, dtype=np.int32)
        shift = np.empty(3, dtype=np.uint32)
        order = np.empty(27, dtype=np.uint32)
        dims = np.empty(3, dtype=np.int32)
        check(lib().mcgb_cell_list(self.h, replica, t, _i(cs), cs.shape[0], _i(ids), n, _u(shift),
                                   _u(order), _i(dims)))
        nc = int(np.prod(dims))
        npass = 27 if p.sublattice_colors == 3 else 8
        return cs[:nc + 1].copy(), ids, shift, order[:npass], dims

    # ------------------------------------------------------------ learned first stage

    def surrogate_set(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.shape == (SUR_NPARAM,)
        check(lib().mcgb_surrogate_set(self.h, _f(p)))

    def surrogate_get(self):
        p = np.empty(SUR_NPARAM, dtype=np.float32)
        check(lib().mcgb_surrogate_get(self.h, _f(p)))
        return p

    def surrogate_init(self, seed):
        check(lib().mcgb_surrogate_init(self.h, seed))

    def surrogate_eval(self, x, mfma=True):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 4)
        y = np.empty(x.shape[0], dtype=np.float32)
        check(lib().mcgb_surrogate_eval(self.h, x.shape[0], _f(x), 1 if mfma else 0, _f(y)))
        return y

    def surrogate_grad(self, x, t, clip=10.0, params=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(SUR_NPARAM, dtype=np.float32)
        loss = C.c_double()
        pp = None if params is None else _f(np.ascontiguousarray(params, dtype=np.float32))
        check(lib().mcgb_surrogate_grad(self.h, pp, x.shape[0], _f(x), _f(t), clip, _f(g),
                                        C.byref(loss)))
        return g, loss.value

    def surrogate_train(self, steps, batch=8192, lr=1e-3, clip=10.0, seed=1, wfloor=0.0):
        losses = np.empty(steps, dtype=np.float32)
        check(lib().mcgb_surrogate_train(self.h, steps, batch, lr, clip, wfloor, seed, _f(losses)))
        return losses

    def pairs_reserve(self, capacity):
        check(lib().mcgb_pairs_reserve(self.h, capacity))

    def pairs_clear(self):
        check(lib().mcgb_pairs_clear(self.h))

    def pairs_count(self):
        n = C.c_longlong()
        check(lib().mcgb_pairs_count(self.h, C.byref(n)))
        return n.value

    def pairs_get(self, first=0, n=None):
        if n is None:
            n = self.pairs_count() - first
        out = np.empty((n, 5), dtype=np.float32)
        check(lib().mcgb_pairs_get(self.h, first, n, _f(out)))
        return out

    def pairs_put(self, rec):
        rec = np.ascontiguousarray(rec, dtype=np.float32).reshape(-1, 5)
        check(lib().mcgb_pairs_put(self.h, rec.shape[0], _f(rec)))

    def emit_pairs(self, stride=1, offset=0, replica=0):
        n = C.c_longlong()
        check(lib().mcgb_emit_pairs(self.h, replica, stride, offset, C.byref(n)))
        return n.value

    def da_enable(self, on=True, judge=False):
        check(lib().mcgb_da_enable(self.h, 1 if on else 0, 1 if judge else 0))

    def da_guard(self, window_sweeps=20, max_stage2_rejection=0.5, min_stage1_acceptance=0.02):
        check(lib().mcgb_da_guard(self.h, window_sweeps, max_stage2_rejection,
                                  min_stage1_acceptance))

    def da_status(self):
        """{"on": bool, "fallbacks": int, "stage2_rejection": r, "stage1_acceptance": a} (the
        last guarded window)."""
        on, fb = C.c_int(), C.c_uint64()
        rates = np.zeros(2)
        check(lib().mcgb_da_status(self.h, C.byref(on), C.byref(fb), _d(rates)))
        return {"on": bool(on.value), "fallbacks": fb.value, "stage2_rejection": rates[0],
                "stage1_acceptance": rates[1]}

    def da_histogram(self, replica=0, clear=False):
        h = np.zeros(DA_BINS, dtype=np.uint64)
        check(lib().mcgb_da_histogram(self.h, replica, h.ctypes.data_as(u64p), 1 if clear else 0))
        return h

    def sweep_async(self, n=1):
        check(lib().mcgb_sweep_async(self.h, n))

    def sync(self):
        check(lib().mcgb_sync(self.h))

    def snapshot_begin(self, replica=0):
        check(lib().mcgb_snapshot_begin(self.h, replica))

    def snapshot_wait(self, replica=0):
        """(pos, u, box) of the snapshot begun last on this replica, FP64, particle order."""
        N = self.params[replica].num_particles
        arrs = [np.empty(N) for _ in range(6)]
        box = np.empty(3)
        check(lib().mcgb_snapshot_wait(self.h, replica, N, *(_d(a) for a in arrs), _d(box)))
        return arrs[:3], arrs[3:], box

    def step_sizes(self, replica=0):
        """(max_displacement, max_rotation, adaptations) of a replica; float32 values."""
        d, r, n = C.c_float(), C.c_float(), C.c_uint64()
        check(lib().mcgb_get_step_sizes(self.h, replica, C.byref(d), C.byref(r), C.byref(n)))
        return np.float32(d.value), np.float32(r.value), n.value

    def get_resume(self, replica=0):
        """The replica's resume record (a Resume): seed, sweeps done, step sizes, adaptation
        counters.  Together with get_state_fixed it fixes every later sweep."""
        rec = Resume()
        check(lib().mcgb_get_resume(self.h, replica, C.byref(rec)))
        return rec

    def resume_state(self, q, u, box, record, replica=0):
        """Upload a state verbatim with its resume record; the chain then continues exactly as
        the one the record was taken from.  The record's seed must equal the replica's."""
        q = [np.ascontiguousarray(a, dtype=np.uint32) for a in q]
        u = [np.ascontiguousarray(a, dtype=np.float64) for a in u]
        box = np.ascontiguousarray(box, dtype=np.float64)
        check(lib().mcgb_resume_state(self.h, replica, q[0].shape[0], *(_u(a) for a in q),
                                      *(_d(a) for a in u), _d(box), C.byref(record)))

    def set_step_sizes(self, max_displacement, max_rotation, replica=0):
        check(lib().mcgb_set_step_sizes(self.h, replica, float(max_displacement),
                                        float(max_rotation)))
