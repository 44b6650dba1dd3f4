"""Benchmark of the checkerboard sweep on one MI355X: plain FP32 vs learned delayed acceptance.

Plain workload (the `value` of the line): N = 256000 Gay-Berne ellipsoids GB(3, 5, 2, 1) at
rho* = 0.32, T* = 1.0, one colloid of radius 5 sigma0 fixed at the box centre (GBCollRod form),
seeded jittered 64^3 lattice (jitter 0.05, sites within 5.5 of the colloid removed), moves
0.1 sigma0 / 0.1 rad, FP32 pair math with FP64 totals.  A step is one sweep (8 sublattice
passes, ~N trial moves).  The lattice start's overlaps sit inside the 1e9 cap of the pair
energy, behind a ~1e36 barrier that 0.1 moves cannot cross; 300 untimed sweeps at 0.6 / 0.6
remove them, then the timed sweeps run at 0.1 / 0.1.

Effective samples (the `ess` object): E/N and P2 series of the same geometry from a nematic
start (rods along z on layers), equilibrated, then measured: tau_int (Sokal window) with its
error, n, n/tau, the means of both halves and their z, a stationarity flag (|z| < 2) and whether
n >= 50 tau.  Delayed acceptance (the `da` object): the same geometry at T* = 0.8; the per-pair
MLP 4-64-64-1 (SiLU, BF16 matrix-core inference) is trained on the GPU on exact pair energies
emitted from GPU trajectories at both temperatures; plain and delayed acceptance then run from
the same equilibrated state, with the plain kernel as fallback (mcgb_da_guard).

Also: the NPT workload (density), the 32-replica batch, the 8000-particle CPU configuration
(`n8000`: N = 8000 without a colloid, the FP64 oracle on the host cores beside the GPU in FP32
and FP64), and `cpu_baseline`,
the FP64 oracle timed on a bounded sample of the plain workload.  One JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m  # noqa: E402

METRIC = "trial moves/s and ESS/s of <E>,P2 at N=256k: plain vs surrogate-DA"
PEAK_FP32_TFLOPS = 157.3
# algorithmic flop counts: candidate test 17, GB pair 68, colloid-rod term 45
F_CAND, F_PAIR, F_COLL = 17.0, 68.0, 45.0
# HBM bytes per launch of the W=2 pass kernel on this workload: rocprofv3 FETCH_SIZE x 2.000 +
# WRITE_SIZE x 1.000 (the factors a plain 256 MiB float32 copy reads), from the two counter
# passes summarised in profiles/sweep_fp32_N256000_colloid_hbm_traffic.md
TRAFFIC_PASS_BYTES = 29.74e6
# pass schedules (mcgb_params.pass_schedule)
SCHEDULES = ("sequential", "speculative", "sequential_4waves", "two_moves_in_flight")
# step sizes that move the lattice start's capped overlaps out of the cap (untimed)
RELAX_STEP, RELAX_ROT = 0.6, 0.6


def workload(N=256000, rho=0.32, T=1.0, R=5.0, seed=20261203, colors=2):
    p = m.gb_params(N, rho, T, colloid_radius=R, precision=32, seed=seed,
                    lattice_jitter=0.05, colloid_exclusion_radius=5.5, sublattice_colors=colors)
    pos, u, box = m.init_lattice(N, rho, colloid_radius=R, exclusion_radius=5.5, jitter=0.05, seed=seed)
    return p, pos, u, box


def nematic_start(N, rho, R, seed, tilt=0.1, layer=3.25):
    """Rods along z (random tilt of std `tilt` rad, random head/tail) on square layers `layer`
    apart, sites within R + 2 of the colloid removed, box V = N/rho + 4/3 pi R^3: a start with
    no capped or near-capped pair, close to the nematic states of this density."""
    rng = np.random.default_rng(seed)
    L = (N / rho + 4.0 / 3.0 * np.pi * R ** 3) ** (1.0 / 3.0)
    nz = max(1, int(L // layer))
    nx = int(np.ceil(np.sqrt(N / nz)))
    while True:
        gx = (np.arange(nx) + 0.5) * (L / nx)
        gz = (np.arange(nz) + 0.5) * (L / nz)
        X, Y, Z = np.meshgrid(gx, gx, gz, indexing="ij")
        sites = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        if R > 0:
            sites = sites[np.linalg.norm(sites - 0.5 * L, axis=1) >= R + 2.0]
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
    return [sites[:, d].copy() for d in range(3)], [u[:, d].copy() for d in range(3)], np.full(3, L)


def host_cpu():
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(1, min(16, ncpu)), model


def cpu_baseline(pos, u, box, N, rho, T, R, seed, sweeps=1):
    """The FP64 oracle (project C restatement of the reference) on the same start state."""
    from oracle import oracle_ctypes as oc
    threads, model = host_cpu()
    c = oc.consts(temperature=T, colloid_radius=R, colloid_on=1)
    q = [oc.to_fixed(pos[d], box[d]) for d in range(3)]
    s = oc.System(c, q, u, box, seed=seed, nthreads=threads)
    t0 = time.perf_counter()
    for _ in range(sweeps):
        s.sweep()
    dt = time.perf_counter() - t0
    return {"value": float(s.counters[0] / dt), "unit": "trial moves/s", "cores": threads,
            "kind": "port", "cpu_model": model,
            "sample": f"{sweeps} FP64 sweep(s) ({int(s.counters[0])} trial moves) of the same "
                      f"N={N} colloid workload from the same start state, OpenMP over the "
                      f"independent cells of each pass, {dt:.2f} s wall"}


def series_report(x, seconds):
    """Effective samples of a per-sweep series: tau_int (Sokal window, c = 5) and its error
    tau sqrt(2 (2M + 1) / n), n, ESS = n / tau, ESS/s, the means of both halves with their
    tau-corrected errors and z, `stationary` (|z| < 2) and `converged` (stationary and
    n >= 50 tau)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    if n < 16:
        return None
    tau, w = m.tau_int(x)
    tau = max(tau, 1.0)

    def half(y):
        t, _ = m.tau_int(y)
        return float(y.mean()), float(y.std() * np.sqrt(max(t, 1.0) / len(y)))

    h = n // 2
    (m1, e1), (m2, e2) = half(x[:h]), half(x[h:])
    z = (m1 - m2) / max(np.hypot(e1, e2), 1e-300)
    return {"mean": float(x.mean()), "se": float(x.std() * np.sqrt(tau / n)),
            "tau_int": float(tau), "tau_err": float(tau * np.sqrt(2.0 * (2 * w + 1) / n)),
            "window": int(w), "n": n, "n_over_tau": float(n / tau), "ess": float(n / tau),
            "ess_per_s": float(n / tau / seconds) if seconds > 0 else None,
            "halves": [m1, m2], "halves_z": float(z), "stationary": bool(abs(z) < 2.0),
            "converged": bool(abs(z) < 2.0 and n >= 50 * tau)}


def progress(msg):
    """One progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def advance(ctx, sweeps, label, chunk=20000):
    """Untimed sweeps in chunks, with a progress line per chunk."""
    done = 0
    while done < sweeps:
        n = min(chunk, sweeps - done)
        ctx.sweep(n)
        done += n
        progress(f"{label}: {done}/{sweeps} sweeps")


def timed(ctx, sweeps, chunk=20000, label="timed"):
    """Runs `sweeps` sweeps (mcgb_sweep returns after hipStreamSynchronize) and returns the wall
    seconds spent in them, the counters, and the per-sweep E/N, P2 and density series, read
    after every `chunk` sweeps (the device ring holds 65536 entries; the reads are not timed)."""
    ctx.reset_counters()
    ctx.series_clear()
    dt = 0.0
    parts = []
    left = sweeps
    while left > 0:
        n = min(chunk, left)
        t0 = time.perf_counter()
        ctx.sweep(n)
        dt += time.perf_counter() - t0
        parts.append(ctx.series(max_n=n + 16))
        ctx.series_clear()
        left -= n
        progress(f"{label}: {sweeps - left}/{sweeps} sweeps, {dt:.1f} s")
    st = ctx.stats()
    e, p2, rho = (np.concatenate([x[k] for x in parts]) for k in range(3))
    return dt, st, e, p2, rho


def drift_fields(st):
    E = abs(st["energy"]) if st["energy"] != 0 else 1.0
    return {"max_fp64_drift": st["max_drift"], "max_fp64_drift_relative": st["max_drift"] / E,
            "resummations": st["resummations"], "resummations_forced": st["resummations_forced"],
            "max_fp64_drift_forced": st["max_drift_forced"], "cap_pairs": st["cap_pairs"],
            "cap_colloid": st["cap_colloid"]}


def ess_leg(args, waves, schedule, T=1.0):
    """E/N and P2 of the plain FP32 sweep on the plain workload's geometry from a nematic start:
    equilibration, then the measured sweeps (graph replays, FP64 re-summation every 100
    sweeps: its drift is reported)."""
    N, rho, R = args.n, 0.32, 5.0
    seed = 20261206
    p = m.gb_params(N, rho, T, colloid_radius=R, precision=32, seed=seed, waves_per_cell=waves,
                    pass_schedule=schedule, energy_check_interval=100)
    pos, u, box = nematic_start(N, rho, R, seed)
    out = {"workload": f"plain FP32 sweep, N={N} rho*=0.32 T*={T} R=5, nematic start",
           "equilibration_sweeps": args.ess_equil, "sweeps": args.ess_sweeps}
    with m.Context(p) as ctx:
        ctx.set_state(pos, u, box)
        ctx.set_kernel_timing(False)
        t0 = time.perf_counter()
        advance(ctx, args.ess_equil, "ess equilibration")
        out["equilibration_s"] = time.perf_counter() - t0
        dt, st, e, p2, _ = timed(ctx, args.ess_sweeps, label="ess measurement")
        out["seconds"] = dt
        out["moves_per_s"] = st["attempted"] / dt
        out["acceptance"] = st["accepted"] / max(1, st["attempted"])
        out["E/N"] = series_report(e, dt)
        out["P2"] = series_report(p2, dt)
        out.update(drift_fields(st))
        pairs = None
        if not args.no_da:
            ctx.pairs_reserve(args.pairs // 2 + 2 * 10 ** 6)
            k = 0
            while ctx.pairs_count() < args.pairs // 2 and k < 8:
                ctx.sweep(10)
                ctx.emit_pairs(stride=8, offset=k % 8)
                k += 1
            pairs = ctx.pairs_get()
    return out, pairs


def ratio_with_error(a, b):
    """(ESS/s of a) / (ESS/s of b) with the relative error of the two tau estimates."""
    if not a or not b or not a["ess_per_s"] or not b["ess_per_s"]:
        return None, None
    r = a["ess_per_s"] / b["ess_per_s"]
    rel = np.hypot(a["tau_err"] / a["tau_int"], b["tau_err"] / b["tau_int"])
    return float(r), float(r * rel)


def da_leg(args, pairs_plain, waves, schedule):
    """Learned delayed acceptance against the plain sweep on the T* = 0.8 (nematic) state."""
    N, rho, T, R = args.n, 0.32, 0.8, 5.0
    seed = 20261207
    p = m.gb_params(N, rho, T, colloid_radius=R, precision=32, seed=seed, waves_per_cell=waves,
                    pass_schedule=schedule, energy_check_interval=100)
    pos, u, box = nematic_start(N, rho, R, seed)
    out = {"workload": f"delayed acceptance vs plain, N={N} rho*=0.32 T*=0.8 R=5, nematic start",
           "T": T, "equilibration_sweeps": args.da_equil, "sweeps": args.da_sweeps}
    with m.Context(p) as ctx:
        ctx.set_state(pos, u, box)
        ctx.set_kernel_timing(False)
        advance(ctx, args.da_equil, "da equilibration")
        # training pairs: this trajectory plus the T* = 1.0 nematic one
        t0 = time.perf_counter()
        ctx.pairs_reserve(int(1.2 * args.pairs) + 4 * 10 ** 6)
        if pairs_plain is not None:
            ctx.pairs_put(pairs_plain)
        k = 0
        while ctx.pairs_count() < args.pairs and k < 16:
            ctx.sweep(10)
            ctx.emit_pairs(stride=8, offset=k % 8)
            k += 1
        out["train_pairs"] = ctx.pairs_count()
        out["emit_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        ctx.surrogate_init(1)
        losses = ctx.surrogate_train(args.train_steps, batch=4096, lr=3e-3, clip=10.0, seed=5,
                                     wfloor=1.0)
        out["train_s"] = time.perf_counter() - t0
        out["train_steps"] = args.train_steps
        out["loss_first_last"] = [float(losses[:10].mean()), float(losses[-200:].mean())]
        q0, u0, L0 = ctx.get_state_fixed()
        res = {}
        steps = [float(x) for x in args.da_steps.split(",")]
        by_step = {"plain": {}, "da": {}}
        da_w = waves
        # every plain run first: the autotuner then tunes W for the delayed-acceptance kernel,
        # which stays for the rest of this context
        for mode in ("plain", "da"):
            for step in steps:
                ctx.set_state_fixed(q0, u0, L0)
                ctx.da_enable(mode == "da")
                if mode == "da" and step == steps[0]:
                    # the delayed-acceptance kernel has its own register budget: tune W for it
                    # (every schedule runs that kernel while it is on), then restart from q0
                    da_w, _, _ = ctx.autotune(2)
                    ctx.set_state_fixed(q0, u0, L0)
                if mode == "da":
                    ctx.da_guard(50, 0.9, 0.02)  # falls back only for a broken network
                ctx.set_step_sizes(step, step)
                ctx.sweep(2)                               # graph capture outside the timing
                dt, st, e, p2, _ = timed(ctx, args.da_sweeps, label=f"da leg, {mode}, step {step}")
                evaluated = st["attempted"] - st["rejected_left_cell"] - st["rejected_hard_core"]
                r = {"step": step, "moves_per_s": st["attempted"] / dt,
                     "ms_per_sweep": 1e3 * dt / args.da_sweeps,
                     "acceptance": st["accepted"] / max(1, evaluated),
                     "E/N": series_report(e, dt), "P2": series_report(p2, dt),
                     "waves_per_cell": da_w if mode == "da" else waves}
                if mode == "da":
                    s1, s2 = st["da_stage1_rejected"], st["da_stage2_rejected"]
                    r["stage1_rejection"] = s1 / max(1, evaluated)
                    r["stage2_rejection"] = s2 / max(1, evaluated - s1)
                    r["exact_evaluations_skipped"] = s1 / max(1, evaluated)
                    r["guard"] = ctx.da_status()
                by_step[mode][step] = r
        ctx.set_step_sizes(steps[0], steps[0])

        # each chain at its own best step size: the largest min(ESS/s E/N, ESS/s P2)
        def score(r):
            v = [r[o]["ess_per_s"] for o in ("E/N", "P2") if r[o] and r[o]["ess_per_s"]]
            return min(v) if len(v) == 2 else 0.0
        for mode in ("plain", "da"):
            best = max(steps, key=lambda x: score(by_step[mode][x]))
            res[mode] = dict(by_step[mode][best], best_of_steps=steps)
        if len(steps) > 1:
            out["by_step"] = {mode: {str(x): {"moves_per_s": by_step[mode][x]["moves_per_s"],
                                              "acceptance": by_step[mode][x]["acceptance"],
                                              "ess_per_s": {o: (by_step[mode][x][o] or {}).get("ess_per_s")
                                                            for o in ("E/N", "P2")},
                                              "tau_int": {o: (by_step[mode][x][o] or {}).get("tau_int")
                                                          for o in ("E/N", "P2")}}
                                      for x in steps} for mode in by_step}
        # the plain kernel as judge: exact stage for every move, surrogate error histogram
        ctx.set_state_fixed(q0, u0, L0)
        ctx.da_enable(True, judge=True)
        ctx.da_histogram(clear=True)
        ctx.sweep(4)
        h = ctx.da_histogram()
        ctx.da_enable(False)
    tot = max(1, int(h.sum()))
    res["da"]["error_hist_kT"] = {"within_0.5": float(h[28:36].sum() / tot),
                                  "within_1": float(h[24:40].sum() / tot),
                                  "below_-4": float(h[64] / tot), "above_4": float(h[65] / tot)}
    out.update(res)
    out["nematic"] = bool(res["plain"]["P2"] and res["plain"]["P2"]["mean"] > 0.3)
    ratios = {}
    for obs in ("E/N", "P2"):
        ratios[obs] = ratio_with_error(res["da"][obs], res["plain"][obs])
    out["ess_ratio_da_over_plain"] = {k: {"ratio": v[0], "err": v[1]} for k, v in ratios.items()}
    valid = [v for v in ratios.values() if v[0] is not None]
    out["ess_ratio_min"] = min(valid)[0] if valid else None
    out["moves_ratio_da_over_plain"] = res["da"]["moves_per_s"] / res["plain"]["moves_per_s"]
    # the two chains sample one distribution: their means within 2 combined errors
    z = {}
    for obs in ("E/N", "P2"):
        a, b = res["da"][obs], res["plain"][obs]
        if a and b:
            z[obs] = float((a["mean"] - b["mean"]) / max(np.hypot(a["se"], b["se"]), 1e-300))
    out["means_z_da_vs_plain"] = z
    return out


def npt_leg(args, waves, schedule):
    """NPT: N = 64000 GB(3,5,2,1), no colloid, P* = 2, T* = 1, start rho* = 0.28 (jittered
    lattice), one isotropic ln V move (max 0.01) per sweep with a full FP64 recompute."""
    N, rho0, T, P = 64000, 0.28, 1.0, 2.0
    seed = 20261204
    p = m.gb_params(N, rho0, T, precision=32, seed=seed, npt=True, pressure=P,
                    npt_max_dlnv=0.01, lattice_jitter=0.05, sublattice_colors=args.npt_colors)
    p.waves_per_cell = waves
    p.pass_schedule = schedule
    pos, u, box = m.init_lattice(N, rho0, jitter=0.05, seed=seed)
    with m.Context(p) as ctx:
        ctx.set_state(pos, u, box)
        ctx.set_kernel_timing(False)
        ctx.set_step_sizes(RELAX_STEP, RELAX_ROT)         # the lattice start's caps
        ctx.sweep(300)
        ctx.set_step_sizes(0.1, 0.1)
        ctx.sweep(2)
        dt, st, e, p2, rho = timed(ctx, args.npt_sweeps, label="npt")
    h = len(rho) // 2
    return {"workload": "NPT N=64000 P*=2 T*=1 from rho*=0.28, one ln V move per sweep",
            "sublattices": args.npt_colors ** 3,
            "sweeps": args.npt_sweeps, "moves_per_s": st["attempted"] / dt,
            "ms_per_sweep": 1e3 * dt / args.npt_sweeps,
            "volume_acceptance": st["volume_accepted"] / max(1, st["volume_attempted"]),
            "density_start_end": [float(rho[0]), float(rho[-1])],
            "density": series_report(rho[h:], dt / 2), "P2": series_report(p2[h:], dt / 2),
            "note": "statistics over the second half of the sweeps"}


def replicas_leg(args):
    """32 NVT replicas of GB(3,5,2,1) at rho* = 0.30: N log-uniform in [2000, 32000] (seeded),
    T* = 0.8 .. 1.6 evenly, replicas with index = 0 mod 4 carry an R = 3 colloid.  Every
    replica advances in the same kernel launches (one launch per sublattice pass covers the
    active cells of all replicas), the sweep sequence replayed as a HIP graph."""
    seed = 20261205
    rng = np.random.default_rng(seed)
    Ns = np.round(np.exp(rng.uniform(np.log(2000), np.log(32000), 32))).astype(int)
    params, states = [], []
    for i in range(32):
        R = 3.0 if i % 4 == 0 else 0.0
        T = 0.8 + i * 0.8 / 31
        params.append(m.gb_params(int(Ns[i]), 0.30, T, colloid_radius=R, precision=32, seed=seed,
                                  lattice_jitter=0.05, sublattice_colors=args.replica_colors,
                                  colloid_exclusion_radius=R + 0.5 if R > 0 else 0.0))
        states.append(m.init_lattice(int(Ns[i]), 0.30, colloid_radius=R,
                                     exclusion_radius=R + 0.5 if R > 0 else -1.0, jitter=0.05,
                                     seed=seed + i))
    with m.Context(params) as ctx:
        for i, (pos, u, box) in enumerate(states):
            ctx.set_state(pos, u, box, replica=i)
        for i in range(32):
            ctx.set_step_sizes(RELAX_STEP, RELAX_ROT, replica=i)
        ctx.sweep(300)
        for i in range(32):
            ctx.set_step_sizes(0.1, 0.1, replica=i)
        ctx.sweep(args.relax // 2)
        best, sched, ms = ctx.autotune(4)
        ctx.sweep(5)
        ctx.set_kernel_timing(False)
        a0 = sum(ctx.stats(replica=i)["attempted"] for i in range(32))
        t0 = time.perf_counter()
        ctx.sweep(args.steps)
        dt = time.perf_counter() - t0
        a1 = sum(ctx.stats(replica=i)["attempted"] for i in range(32))
    return {"workload": "32 NVT replicas, N log-uniform in [2000, 32000], T* 0.8..1.6, rho*=0.30",
            "sublattices": args.replica_colors ** 3,
            "N_total": int(Ns.sum()), "N_min_max": [int(Ns.min()), int(Ns.max())],
            "sweeps": args.steps, "moves_per_s": (a1 - a0) / dt,
            "ms_per_sweep": 1e3 * dt / args.steps, "waves_per_cell": best,
            "schedule": SCHEDULES[sched], "autotune_ms_per_pass": ms}


def n8000_leg(args):
    """The CPU configuration: N = 8000, rho* = 0.30, T* = 1.2, no colloid, 20^3 simple-cubic
    lattice (spacing 1.4938), orientations uniform; 300 untimed sweeps at 0.6 / 0.6 remove the
    lattice's capped overlaps, then `n8000_sweeps` sweeps at 0.1 / 0.1.  The FP64 oracle on the
    host cores (OpenMP over the independent cells of a pass) against the GPU in FP64 (the same
    chain: its E/N series equals the oracle's) and FP32 (a different chain of the same
    distribution: the means compared within 2 combined errors)."""
    from oracle import oracle_ctypes as oc
    N, rho, T = 8000, 0.30, 1.2
    seed = 20261208
    threads, model = host_cpu()
    L = (N / rho) ** (1.0 / 3.0)
    g = (np.arange(20) + 0.5) * (L / 20)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = [X.ravel().copy(), Y.ravel().copy(), Z.ravel().copy()]
    rng = np.random.default_rng(seed)
    uu = rng.normal(size=(N, 3))
    uu /= np.linalg.norm(uu, axis=1, keepdims=True)
    u = [uu[:, d].copy() for d in range(3)]
    box = np.full(3, L)
    S = args.n8000_sweeps
    out = {"workload": "N=8000 rho*=0.30 T*=1.2, no colloid, 20^3 lattice start",
           "relax_sweeps": 300, "sweeps": S}
    # the GPU, both precisions, from the same start
    qstart = None
    for prec in (64, 32):
        p = m.gb_params(N, rho, T, precision=prec, seed=seed)
        with m.Context(p) as ctx:
            ctx.set_state(pos, u, box)
            ctx.set_kernel_timing(False)
            ctx.set_step_sizes(RELAX_STEP, RELAX_ROT)
            ctx.sweep(300)
            ctx.set_step_sizes(0.1, 0.1)
            if prec == 64:
                qstart = ctx.get_state_fixed()
                t_start = int(ctx.stats()["sweeps"])
            ctx.sweep(2)
            dt, st, e, p2, _ = timed(ctx, S, label=f"n8000 gpu fp{prec}")
        out[f"gpu_fp{prec}"] = {"moves_per_s": st["attempted"] / dt, "seconds": dt,
                                "E/N": series_report(e, dt), "P2": series_report(p2, dt)}
        if prec == 64:
            e64 = e
    # the oracle from the GPU FP64 state after the relaxation (same chain from here on)
    c = oc.consts(temperature=T)
    q, uo, Lo = qstart
    s = oc.System(c, q, uo, Lo, seed=seed, nthreads=threads)
    s.t = t_start
    for _ in range(2):
        s.sweep()
    s.counters[:] = 0
    eo, po = np.empty(S), np.empty(S)
    t0 = time.perf_counter()
    for k in range(S):
        s.sweep()
        eo[k] = s.energy_value() / N
        po[k] = s.order_parameter()[0]
    dt = time.perf_counter() - t0
    out["cpu_fp64"] = {"moves_per_s": float(s.counters[0] / dt), "seconds": dt,
                       "cores": threads, "cpu_model": model, "kind": "port",
                       "E/N": series_report(eo, dt), "P2": series_report(po, dt)}
    out["gpu_fp64_equals_cpu_series"] = bool(np.allclose(e64, eo, rtol=1e-9, atol=1e-9))
    a, b = out["gpu_fp32"], out["cpu_fp64"]
    out["z_gpu_fp32_vs_cpu"] = {
        k: float((a[k]["mean"] - b[k]["mean"]) / max(np.hypot(a[k]["se"], b[k]["se"]), 1e-300))
        for k in ("E/N", "P2") if a[k] and b[k]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-da", action="store_true", help="skip the delayed-acceptance leg")
    ap.add_argument("--no-ess", action="store_true", help="skip the plain ESS leg")
    ap.add_argument("--no-n8000", action="store_true", help="skip the N=8000 CPU-configuration leg")
    ap.add_argument("--cpu-sweeps", type=int, default=10)
    ap.add_argument("--n", type=int, default=256000)
    ap.add_argument("--relax-big", type=int, default=300,
                    help=f"untimed overlap-removal sweeps at {RELAX_STEP}/{RELAX_ROT}")
    ap.add_argument("--relax", type=int, default=200,
                    help="untimed sweeps at 0.1/0.1 after the overlap-removal sweeps")
    ap.add_argument("--waves", type=int, default=0, help="wavefronts per cell (0: autotune)")
    ap.add_argument("--pace", type=int, default=-1,
                    help="sequential-pass pacing in percent of the mean move time (0: off; "
                         "default: the library's, re-checked by the autotuner)")
    ap.add_argument("--schedule", type=int, default=1,
                    help="pass schedule with --waves: 0 sequential, 1 speculative, "
                         "2 sequential built for four wavefronts per SIMD, 3 two moves in "
                         "flight (two wavefronts)")
    ap.add_argument("--ess-equil", type=int, default=20000)
    ap.add_argument("--ess-sweeps", type=int, default=20000)
    ap.add_argument("--da-equil", type=int, default=20000)
    ap.add_argument("--da-sweeps", type=int, default=8000)
    ap.add_argument("--da-steps", default="0.1",
                    help="comma-separated step sizes (translation = rotation) at which plain and "
                         "delayed acceptance both run; each chain is reported at its best")
    ap.add_argument("--pairs", type=int, default=10 ** 7, help="training pairs for the surrogate")
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--no-npt", action="store_true", help="skip the NPT leg")
    ap.add_argument("--npt-sweeps", type=int, default=8000)
    ap.add_argument("--no-replicas", action="store_true", help="skip the 32-replica leg")
    ap.add_argument("--n8000-sweeps", type=int, default=2000)
    ap.add_argument("--colors", type=int, default=2, choices=(2, 3),
                    help="sublattice colours per axis of the plain leg "
                         "(2: 8 sublattices of cells >= r_c; 3: 27 of cells >= r_c/2)")
    ap.add_argument("--npt-colors", type=int, default=2, choices=(2, 3),
                    help="sublattice colours per axis of the NPT leg")
    ap.add_argument("--replica-colors", type=int, default=3, choices=(2, 3),
                    help="sublattice colours per axis of the replica leg (small boxes: 27 "
                         "sublattices keep the per-cell move chains short)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sweep is not "
              "sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    if m.device_count() < 1:
        print("no GPU visible to HIP", file=sys.stderr)
        sys.exit(1)
    N, rho, T, R = args.n, 0.32, 1.0, 5.0
    p, pos, u, box = workload(N, rho, T, R, colors=args.colors)
    seed = int(p.seed)
    if args.waves:
        p.waves_per_cell = args.waves
        p.pass_schedule = args.schedule
    with m.Context(p) as ctx:
        if args.pace >= 0:
            ctx.set_pace(args.pace)
        ctx.set_state(pos, u, box)
        t_relax = time.perf_counter()
        ctx.set_kernel_timing(False)
        ctx.set_step_sizes(RELAX_STEP, RELAX_ROT)
        ctx.sweep(args.relax_big)
        caps_after_relax = ctx.stats()["cap_pairs"]
        ctx.set_step_sizes(0.1, 0.1)
        if args.relax:
            ctx.sweep(args.relax)
        ctx.set_kernel_timing(True)
        t_relax = time.perf_counter() - t_relax
        tune = None
        tune_w = args.waves
        tune_s = args.schedule
        if not args.waves:
            best, sched, ms = ctx.autotune(4)
            tune = {"waves_per_cell": best, "schedule": SCHEDULES[sched], "ms_per_pass": ms,
                    "pace_pct": ctx.pace()}
            tune_w, tune_s = best, sched
        pace_pct = ctx.pace()
        if args.warmup:
            ctx.sweep(args.warmup)
        ctx.set_kernel_timing(False)   # the timed sweeps: graph replays only
        dt, st, e, p2, _ = timed(ctx, args.steps)
        ctx.set_kernel_timing(True)    # kernel times from HIP events: 4 x 2 more sweeps, untimed
        kts = []
        for _ in range(4):
            ctx.sweep(2)
            kts.append(ctx.kernel_times())
        kt = {k: float(np.mean([x[k] for x in kts])) for k in kts[0]}
        t_get = time.perf_counter()
        ctx.get_state()                # the whole state device -> host (PCIe), outside the timed loop
        t_get = time.perf_counter() - t_get
        t_get2 = time.perf_counter()
        ctx.get_state()                # again: the pinned record is allocated by the first call
        t_get2 = time.perf_counter() - t_get2
    moves = st["attempted"]
    value = moves / dt
    flops_sweep = (F_CAND * st["candidates"] + F_PAIR * st["pairs"] + F_COLL * st["colloid_evals"]) / args.steps
    flops_pass = flops_sweep / float(args.colors ** 3)
    achieved = flops_pass / (kt["ms_pass"] * 1e-3) / 1e12 if kt["ms_pass"] > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "trial moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain FP32 NVT checkerboard sweep with a colloid",
                   "N": N, "density": rho, "T": T, "colloid_radius": R, "gb": [3, 5, 2, 1],
                   "cutoff": 4.0, "max_displacement": 0.1, "max_rotation": 0.1,
                   "lattice_jitter": 0.05, "seed": seed,
                   "step": f"one sweep ({args.colors ** 3} sublattice passes)",
                   "start": f"jittered lattice; {args.relax_big} untimed sweeps at "
                            f"{RELAX_STEP}/{RELAX_ROT} (capped overlaps left: {caps_after_relax}), "
                            f"{args.relax} at 0.1/0.1"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": (TRAFFIC_PASS_BYTES if (tune_w == 2 and tune_s == 0 and args.colors == 2)
                                 else None),
                     "kernel": "k_pass<float>", "ms_per_launch": kt["ms_pass"],
                     "flop_per_launch": flops_pass},
        "acceptance": st["accepted"] / max(1, moves),
        "ms_rebuild": kt["ms_rebuild"],
        "ms_sweep_device": kt["ms_sweep"],
        "ms_get_state": 1e3 * t_get,
        "ms_get_state_repeat": 1e3 * t_get2,
        "energy_total": st["energy"],
        "relax_s": t_relax,
        "autotune": tune,
        "pace_pct": pace_pct,
    }
    pairs_plain = None
    if not args.no_ess or not args.no_da:
        ess, pairs_plain = ess_leg(args, tune_w or 2, tune_s)
        if not args.no_ess:
            out["ess"] = ess
    if not args.no_da:
        out["da"] = da_leg(args, pairs_plain, tune_w or 2, tune_s)
    if not args.no_npt:
        out["npt"] = npt_leg(args, tune_w or 2, tune_s)
    if not args.no_replicas:
        out["replicas"] = replicas_leg(args)
    if not args.no_n8000:
        out["n8000"] = n8000_leg(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pos, u, box, N, rho, T, R, seed, sweeps=args.cpu_sweeps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
