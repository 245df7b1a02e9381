"""Benchmark of the collision pipeline's hard-model step on one MI355X.

Workload (default): one hard-model timestep of a 300,000-cell colony:
lengths U[1,2], diameter 0.5, nematic domains with directions U[0,pi) packed
at area fraction ~0.85 (R ~ 279) with ~1e-3..1e-2 overlaps, contact cutoff
0.05 in every ReLCP phase, BBPGD and ReLCP tolerance 1e-6, dt = 1e-3,
lambda = 1e-2, tau = 1/ln2, seeded and synthetic.  The full ReLCP loop runs
(3 phases on this input).  I.i.d. uniform centres at area fraction 0.9 are
not used: ~16% of their contacts are crossing rods whose normal is the
arbitrary (1,0,0) of ArrayMath::normalize, and BBPGD then does not reach 1e-6
(3M iterations at 5k and 20k cells).

A "step" = restore the initial state from a device-resident snapshot (a
device-to-device copy) + pcc_step.  The state is resident in HBM before the
timed region starts.  `value` is the mixed-precision path's cell-steps/s; the
plain-FP64 path is measured beside it on identical inputs.

Prints ONE JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "proliferatingcellcollectives-idp_amd"))

import numpy as np  # noqa: E402

from pcc import generators as G  # noqa: E402
from pcc.abi_types import PREC_FP64, PREC_MIXED, reference_params  # noqa: E402

METRIC = "cell-steps/s @300k cells, batched mixed vs plain FP64; p50/p99 step latency"
HBM_PEAK_GBS = 8000.0


def params_for(precision):
    return reference_params(dt=1e-3, tolerance=1e-6, tau=1 / math.log(2), lam=1e-2,
                            cutoff_phase0=0.05, cutoff_phaseK=0.05, relcp_tol=1e-6,
                            precision=precision, max_bbpgd_iterations=2_000_000)


def run_gpu(bodies, precision, steps, warmup):
    from pcc.engine import ParticleManager
    pm = ParticleManager(params_for(precision), device=0)
    pm.set_state(bodies)
    pm.snapshot()
    results = []
    for i in range(warmup):
        pm.restore()
        results.append(pm.step(i))
    lat = []
    t0 = time.perf_counter()
    for i in range(steps):
        a = time.perf_counter()
        pm.restore()
        r = pm.step(i)
        lat.append(time.perf_counter() - a)
        results.append(r)
    total = time.perf_counter() - t0
    # kernel timing pass: one more step with HIP events around every launch of
    # the BBPGD kernels, on the context stream
    phases = pm.phase_stats()
    pm.set_timing(True)
    pm.restore()
    pm.step(0)
    tim = pm.timing()
    pm.set_timing(False)
    pm.close()
    last = results[-1]
    return {"total_s": total, "lat": lat, "result": last, "timing": tim, "phases": phases}


def roofline(tim):
    best = None
    for name in ("constraint_apply", "body_apply"):
        t = tim[name]
        if t["launches"] == 0:
            continue
        avg_s = t["total_ms"] / t["launches"] / 1e3
        gbs = t["bytes_per_launch"] / avg_s / 1e9
        entry = {"kernel": name, "avg_us": avg_s * 1e6, "gbs": gbs,
                 "bytes": t["bytes_per_launch"], "share_ms": t["total_ms"]}
        if best is None or entry["share_ms"] > best["share_ms"]:
            best = entry
    return best


def cpu_baseline(bodies, phases, sample_caps=(10, 110)):
    """The FP64 oracle built with the reference's Release flags + OpenMP, on this
    host's cores, on the same step: every ReLCP phase the GPU solved is run
    (detection + assembly + BBPGD) with BBPGD capped at two iteration counts;
    each phase's time is fitted as fixed + per-iteration cost and evaluated at
    the GPU FP64 path's iteration count for that phase (identical algorithm
    and iteration path), then summed."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    p = params_for(PREC_FP64).replace(max_recursive_iterations=max(1, len(phases)))
    per_cap = []
    t_total = 0.0
    for cap in sample_caps:
        t = time.perf_counter()
        oracle.step(p, bodies, fast=True, iteration_cap=cap)
        t_total += time.perf_counter() - t
        per_cap.append(oracle.phase_times(fast=True))
    est = 0.0
    detail = []
    for k, (iters, ncons) in enumerate(phases):
        if k >= len(per_cap[0]) or k >= len(per_cap[1]):
            break
        t0, t1 = per_cap[0][k], per_cap[1][k]
        slope = max(t1 - t0, 0.0) / (sample_caps[1] - sample_caps[0])
        fixed = max(t0 - slope * sample_caps[0], 0.0)
        est += fixed + slope * iters
        detail.append({"phase": k, "iterations": iters, "constraints": ncons,
                       "ms_per_iteration": slope * 1e3, "fixed_s": fixed})
    return {
        "value": bodies.n / est,
        "unit": "cell-steps/s",
        "cores": oracle.threads(fast=True),
        "kind": "port",
        "sample": (f"FP64 CPU restatement (-O3 -ffast-math, OpenMP, {oracle.threads(fast=True)} "
                   f"threads) on the same {bodies.n}-cell step: each of the {len(phases)} ReLCP "
                   f"phases run with BBPGD capped at {sample_caps[0]} and {sample_caps[1]} "
                   f"iterations ({t_total:.1f} s of CPU work), per-phase cost fitted and "
                   f"evaluated at the GPU FP64 iteration counts -> {est:.1f} s per step"),
        "est_step_s": est,
        "phases": detail,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=300_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["both", "fp64", "mixed"], default="both")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    bodies = G.dense_colony(args.n, seed=args.seed)
    print(f"[bench] n={bodies.n} R={G.disc_radius(bodies):.1f} "
          f"area_fraction={G.area_fraction(bodies):.3f}", file=sys.stderr, flush=True)
    out = {}
    for prec, name in ((PREC_FP64, "fp64"), (PREC_MIXED, "mixed")):
        if args.only not in ("both", name):
            continue
        r = run_gpu(bodies, prec, args.steps, args.warmup)
        lat = np.array(r["lat"]) * 1e3
        out[name] = {
            "cell_steps_per_s": bodies.n * args.steps / r["total_s"],
            "ms_per_step": r["total_s"] / args.steps * 1e3,
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "bbpgd_iterations": r["result"]["bbpgd_iterations"],
            "n_constraints": r["result"]["n_constraints"],
            "residual": r["result"]["residual"],
            "roofline": roofline(r["timing"]),
            "kernels": r["timing"],
            "phases": r["phases"],
        }
        print(f"[bench] {name}: {out[name]['ms_per_step']:.1f} ms/step, "
              f"{out[name]['bbpgd_iterations']} BBPGD iterations, "
              f"C={out[name]['n_constraints']}", file=sys.stderr, flush=True)
    head = out.get("mixed") or out.get("fp64")
    rf = head["roofline"]
    line = {
        "metric": METRIC,
        "value": head["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if "mixed" in out else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"hard-model single step (full ReLCP), {bodies.n}-cell colony packing "
                         f"(nematic domains, area fraction {G.area_fraction(bodies):.3f}, "
                         f"R={G.disc_radius(bodies):.0f}, seed {args.seed}), lengths U[1,2], "
                         f"d=0.5, cutoff 0.05, BBPGD/ReLCP tol 1e-6, dt 1e-3"),
            "n_cells": bodies.n,
            "precision": "mixed (FP32 bulk applies, FP64 reductions/integration)"
            if "mixed" in out else "fp64",
            "instances": 1,
        },
        "p50_step_ms": head["p50_ms"],
        "p99_step_ms": head["p99_ms"],
        "roofline": {
            "bound": "hbm",
            "kernel": rf["kernel"],
            "achieved": rf["gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": rf["gbs"] / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": rf["avg_us"],
            "algorithmic_bytes_per_launch": rf["bytes"],
        },
    }
    if "fp64" in out and "mixed" in out:
        line["fp64"] = {k: out["fp64"][k] for k in ("cell_steps_per_s", "ms_per_step", "p50_ms",
                                                    "p99_ms", "bbpgd_iterations", "n_constraints",
                                                    "roofline")}
        line["mixed_over_fp64"] = out["mixed"]["cell_steps_per_s"] / out["fp64"]["cell_steps_per_s"]
        line["mixed_bbpgd_iterations"] = out["mixed"]["bbpgd_iterations"]
    line["phases"] = {k: v["phases"] for k, v in out.items()}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(bodies, (out.get("fp64") or head)["phases"])
    print(json.dumps(line))


if __name__ == "__main__":
    main()
