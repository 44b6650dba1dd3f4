"""Benchmark of the checkerboard sweep on one MI355X: plain FP32 vs learned delayed acceptance.

Plain workload (the `value` of the line): N = 256000 Gay-Berne ellipsoids GB(3, 5, 2, 1) at
rho* = 0.32, T* = 1.0, one colloid of radius 5 sigma0 fixed at the box centre (GBCollRod form),
seeded jittered 64^3 lattice (jitter 0.05, sites within 5.5 of the colloid removed), moves
0.1 sigma0 / 0.1 rad, FP32 pair math with FP64 totals.  A step is one sweep (8 sublattice
passes, ~N trial moves).

Delayed-acceptance workload (the `da` object): the same geometry at T* = 0.8.  The per-pair
MLP 4-64-64-1 (SiLU, BF16 matrix-core inference) is trained on the GPU on ~10^7 exact pair
energies emitted from GPU trajectories of both temperatures; then the plain sweep and the
delayed-acceptance sweep run the same number of sweeps from the same state, and their
trial moves/s and ESS/s of E/N and P2 are reported side by side.

Prints one JSON line: trial moves/s of the plain sweep over the timed sweeps, the roofline of
its dominant kernel (the sublattice-pass kernel: algorithmic FP32 flops per launch over its
HIP-event-timed average duration, against the 157.3 TFLOP/s FP32 vector peak), ESS/s of E/N
and P2, the delayed-acceptance comparison, and the FP64 CPU oracle timed on a bounded sample of
the plain workload.
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
TRAFFIC_PASS_BYTES = 28.37e6
# pass schedules (mcgb_params.pass_schedule)
SCHEDULES = ("sequential", "speculative", "sequential_4waves")


def workload(N=256000, rho=0.32, T=1.0, R=5.0, seed=20261203, colors=2):
    p = m.gb_params(N, rho, T, colloid_radius=R, precision=32, seed=seed,
                    lattice_jitter=0.05, colloid_exclusion_radius=5.5, sublattice_colors=colors)
    pos, u, box = m.init_lattice(N, rho, colloid_radius=R, exclusion_radius=5.5, jitter=0.05, seed=seed)
    return p, pos, u, box


def cpu_baseline(pos, u, box, N, rho, T, R, seed, sweeps=1):
    """The FP64 oracle (project C restatement of the reference) on the same start state."""
    from oracle import oracle_ctypes as oc
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(16, ncpu))
    c = oc.consts(temperature=T, colloid_radius=R, colloid_on=1)
    q = [oc.to_fixed(pos[d], box[d]) for d in range(3)]
    s = oc.System(c, q, u, box, seed=seed, nthreads=threads)
    t0 = time.perf_counter()
    for _ in range(sweeps):
        s.sweep()
    dt = time.perf_counter() - t0
    return {"value": float(s.counters[0] / dt), "unit": "trial moves/s", "cores": threads,
            "kind": "port",
            "sample": f"{sweeps} FP64 sweep(s) ({int(s.counters[0])} trial moves) of the same "
                      f"N={N} colloid workload from the same start state, OpenMP over the "
                      f"independent cells of each pass, {dt:.2f} s wall"}


def ess_per_s(x, seconds):
    if len(x) < 8 or seconds <= 0:
        return None
    tau, _ = m.tau_int(np.asarray(x))
    return float(len(x) / tau / seconds)


def timed(ctx, sweeps):
    ctx.reset_counters()
    ctx.series_clear()
    t0 = time.perf_counter()
    ctx.sweep(sweeps)                  # returns after hipStreamSynchronize
    dt = time.perf_counter() - t0
    st = ctx.stats()
    e, p2, _ = ctx.series()
    return dt, st, e, p2


def da_leg(args, pairs_plain, waves, schedule):
    """Learned delayed acceptance against the plain sweep on the T* = 0.8 state."""
    N, rho, T, R = args.n, 0.32, 0.8, 5.0
    p, pos, u, box = workload(N, rho, T, R)
    p.waves_per_cell = waves
    p.pass_schedule = schedule
    out = {"workload": "delayed acceptance vs plain, same geometry at T*=0.8",
           "T": T, "sweeps": args.da_sweeps}
    with m.Context(p) as ctx:
        ctx.set_state(pos, u, box)
        ctx.sweep(args.relax)
        # training pairs: this trajectory plus the plain workload's (T* = 1.0)
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
        for mode in ("plain", "da"):
            ctx.set_state_fixed(q0, u0, L0)
            ctx.da_enable(mode == "da")
            if mode == "da":
                # the delayed-acceptance kernel has its own register budget: tune W for it
                # (every schedule runs that kernel while it is on), then restart from q0
                da_w, _, _ = ctx.autotune(2)
                ctx.set_state_fixed(q0, u0, L0)
            ctx.sweep(2)                                   # graph capture outside the timing
            dt, st, e, p2 = timed(ctx, args.da_sweeps)
            evaluated = st["attempted"] - st["rejected_left_cell"] - st["rejected_hard_core"]
            ee, ep = ess_per_s(e, dt), ess_per_s(p2, dt)
            r = {"moves_per_s": st["attempted"] / dt, "ms_per_sweep": 1e3 * dt / args.da_sweeps,
                 "acceptance": st["accepted"] / max(1, evaluated),
                 "ess_per_s": {"E/N": ee, "P2": ep},
                 "mean": {"E/N": float(np.mean(e)), "P2": float(np.mean(p2))}}
            r["waves_per_cell"] = da_w if mode == "da" else waves
            if mode == "da":
                s1, s2 = st["da_stage1_rejected"], st["da_stage2_rejected"]
                r["stage1_rejection"] = s1 / max(1, evaluated)
                r["stage2_rejection"] = s2 / max(1, evaluated - s1)
                r["exact_evaluations_skipped"] = s1 / max(1, evaluated)
            res[mode] = r
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
    mins = []
    for mode in ("plain", "da"):
        v = [x for x in res[mode]["ess_per_s"].values() if x is not None]
        mins.append(min(v) if v else None)
    out["ess_ratio_da_over_plain"] = (mins[1] / mins[0]) if (mins[0] and mins[1]) else None
    out["moves_ratio_da_over_plain"] = res["da"]["moves_per_s"] / res["plain"]["moves_per_s"]
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
        ctx.sweep(2)
        ctx.reset_counters()
        ctx.series_clear()
        ctx.set_kernel_timing(False)
        t0 = time.perf_counter()
        ctx.sweep(args.npt_sweeps)
        dt = time.perf_counter() - t0
        st = ctx.stats()
        e, p2, rho = ctx.series(max_n=max(1 << 16, args.npt_sweeps + 16))
    h = len(rho) // 2
    return {"workload": "NPT N=64000 P*=2 T*=1 from rho*=0.28, one ln V move per sweep",
            "sublattices": args.npt_colors ** 3,
            "sweeps": args.npt_sweeps, "moves_per_s": st["attempted"] / dt,
            "ms_per_sweep": 1e3 * dt / args.npt_sweeps,
            "volume_acceptance": st["volume_accepted"] / max(1, st["volume_attempted"]),
            "density_start_end": [float(rho[0]), float(rho[-1])],
            "density_mean_second_half": float(np.mean(rho[h:])),
            "P2_mean_second_half": float(np.mean(p2[h:])),
            "ess_per_s": {"density": ess_per_s(rho[h:], dt / 2), "P2": ess_per_s(p2[h:], dt / 2)}}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-da", action="store_true", help="skip the delayed-acceptance leg")
    ap.add_argument("--cpu-sweeps", type=int, default=10)
    ap.add_argument("--n", type=int, default=256000)
    ap.add_argument("--relax", type=int, default=200,
                    help="untimed sweeps from the lattice start before tuning and warm-up")
    ap.add_argument("--waves", type=int, default=0, help="wavefronts per cell (0: autotune)")
    ap.add_argument("--schedule", type=int, default=1,
                    help="pass schedule with --waves: 0 sequential, 1 speculative, "
                         "2 sequential built for four wavefronts per SIMD")
    ap.add_argument("--da-sweeps", type=int, default=100)
    ap.add_argument("--pairs", type=int, default=10 ** 7, help="training pairs for the surrogate")
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--no-npt", action="store_true", help="skip the NPT leg")
    ap.add_argument("--npt-sweeps", type=int, default=5000)
    ap.add_argument("--no-replicas", action="store_true", help="skip the 32-replica leg")
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
    pairs_plain = None
    with m.Context(p) as ctx:
        ctx.set_state(pos, u, box)
        t_relax = time.perf_counter()
        if args.relax:
            ctx.sweep(args.relax)
        t_relax = time.perf_counter() - t_relax
        tune = None
        tune_w = args.waves
        tune_s = args.schedule
        if not args.waves:
            best, sched, ms = ctx.autotune(4)
            tune = {"waves_per_cell": best, "schedule": SCHEDULES[sched], "ms_per_pass": ms}
            tune_w, tune_s = best, sched
        if args.warmup:
            ctx.sweep(args.warmup)
        ctx.set_kernel_timing(False)   # the timed sweeps: graph replays only
        dt, st, e, p2 = timed(ctx, args.steps)
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
        if not args.no_da:
            ctx.pairs_reserve(args.pairs // 2 + 2 * 10 ** 6)
            k = 0
            while ctx.pairs_count() < args.pairs // 2 and k < 8:
                ctx.sweep(10)
                ctx.emit_pairs(stride=8, offset=k % 8)
                k += 1
            pairs_plain = ctx.pairs_get()
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
                   "step": f"one sweep ({args.colors ** 3} sublattice passes)"},
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
        "ess_per_s": {"E/N": ess_per_s(e, dt), "P2": ess_per_s(p2, dt), "samples": int(len(e))},
        # |running total - FP64 re-summation| at the re-summations (recorded after sweeps whose
        # accepted moves changed the finite energy by <= 1e10 in total), beside the total it is
        # a difference of (caps included)
        "max_fp64_drift": st["max_drift"],
        "energy_total": st["energy"],
        "resummations_timed": st["resummations"],
        "relax_sweeps": args.relax,
        "relax_s": t_relax,
        "autotune": tune,
    }
    if not args.no_da:
        out["da"] = da_leg(args, pairs_plain, tune_w or 2, tune_s)
    if not args.no_npt:
        out["npt"] = npt_leg(args, tune_w or 2, tune_s)
    if not args.no_replicas:
        out["replicas"] = replicas_leg(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pos, u, box, N, rho, T, R, seed, sweeps=args.cpu_sweeps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
