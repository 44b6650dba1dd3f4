"""Benchmark of the plain FP32 checkerboard sweep on one MI355X.

Workload: N = 256000 Gay-Berne ellipsoids GB(3, 5, 2, 1) at rho* = 0.32, T* = 1.0, one colloid
of radius 5 sigma0 fixed at the box centre (GBCollRod form), seeded jittered 64^3 lattice
(jitter 0.05, sites within 5.5 of the colloid removed), moves 0.1 sigma0 / 0.1 rad, FP32 pair
math with FP64 totals.  A step is one sweep (8 sublattice passes, ~N trial moves).

Prints one JSON line: trial moves/s over the timed sweeps, the roofline of the dominant kernel
(the sublattice-pass kernel, algorithmic FP32 flops per launch over its HIP-event-timed average
duration, against the 157.3 TFLOP/s FP32 vector peak), ESS/s of E/N and P2 over the timed
sweeps, and the FP64 CPU oracle timed on a bounded sample of the same workload.
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
TRAFFIC_PASS_BYTES = 27.99e6


def workload(N=256000, rho=0.32, T=1.0, R=5.0, seed=20261203):
    p = m.gb_params(N, rho, T, colloid_radius=R, precision=32, seed=seed,
                    lattice_jitter=0.05, colloid_exclusion_radius=5.5)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sweeps", type=int, default=10)
    ap.add_argument("--n", type=int, default=256000)
    ap.add_argument("--relax", type=int, default=200,
                    help="untimed sweeps from the lattice start before tuning and warm-up")
    ap.add_argument("--waves", type=int, default=0, help="wavefronts per cell (0: autotune)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sweep is not "
              "sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    if m.device_count() < 1:
        print("no GPU visible to HIP", file=sys.stderr)
        sys.exit(1)
    N, rho, T, R = args.n, 0.32, 1.0, 5.0
    p, pos, u, box = workload(N, rho, T, R)
    seed = int(p.seed)
    if args.waves:
        p.waves_per_cell = args.waves
    with m.Context(p) as ctx:
        ctx.set_state(pos, u, box)
        t_relax = time.perf_counter()
        if args.relax:
            ctx.sweep(args.relax)
        t_relax = time.perf_counter() - t_relax
        tune = None
        tune_w = args.waves
        if not args.waves:
            best, ms = ctx.autotune(4)
            tune = {"waves_per_cell": best, "ms_per_pass": ms}
            tune_w = best
        if args.warmup:
            ctx.sweep(args.warmup)
        ctx.reset_counters()
        ctx.series_clear()
        t0 = time.perf_counter()
        ctx.sweep(args.steps)          # returns after hipStreamSynchronize
        dt = time.perf_counter() - t0
        st = ctx.stats()
        kt = ctx.kernel_times()
        e, p2, _ = ctx.series()
        t_get = time.perf_counter()
        ctx.get_state()                # the whole state device -> host (PCIe), outside the timed loop
        t_get = time.perf_counter() - t_get
    moves = st["attempted"]
    value = moves / dt
    flops_sweep = (F_CAND * st["candidates"] + F_PAIR * st["pairs"] + F_COLL * st["colloid_evals"]) / args.steps
    flops_pass = flops_sweep / 8.0
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
                   "lattice_jitter": 0.05, "seed": seed, "step": "one sweep (8 sublattice passes)"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": TRAFFIC_PASS_BYTES if tune_w == 2 else None,
                     "kernel": "k_pass<float>", "ms_per_launch": kt["ms_pass"],
                     "flop_per_launch": flops_pass},
        "acceptance": st["accepted"] / max(1, moves),
        "ms_rebuild": kt["ms_rebuild"],
        "ms_sweep_device": kt["ms_sweep"],
        "ms_get_state": 1e3 * t_get,
        "ess_per_s": {"E/N": ess_per_s(e, dt), "P2": ess_per_s(p2, dt), "samples": int(len(e))},
        "max_fp64_drift": st["max_drift"],
        "resummations_timed": st["resummations"],
        "relax_sweeps": args.relax,
        "relax_s": t_relax,
        "autotune": tune,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pos, u, box, N, rho, T, R, seed, sweeps=args.cpu_sweeps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
