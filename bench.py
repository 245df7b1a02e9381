"""Benchmark of the collision pipeline's hard-model step on one MI355X.

Headline workload (default): one hard-model timestep of 300,000-cell colonies:
lengths U[1,2], diameter 0.5, nematic domains with directions U[0,pi) packed
at area fraction ~0.85 (R ~ 279) with ~1e-3..1e-2 overlaps, contact cutoff
0.05 in every ReLCP phase, BBPGD and ReLCP tolerance 1e-6, dt = 1e-3,
lambda = 1e-2, tau = 1/ln2, seeded and synthetic.  The full ReLCP loop runs
(3 phases on this input).  `--placement iid` steps i.i.d. uniform centres at
area fraction 0.9 instead (the configured placement): ~15% of their contacts
are crossing rods whose normal is the arbitrary (1,0,0) of
ArrayMath::normalize, and their phase-0 LCP needs ~18,000 BBPGD iterations at
5k and 20k cells (the nematic packing's needs 20).

`--instances B` colonies (seeds seed..seed+B-1) are stepped together by the
batched mixed-precision engine (one context, one step call, captured HIP
graphs; a colony of this size fills the device alone, so the batch step runs
the colonies' ReLCP loops one after another instead of in lockstep); the plain
FP64 single-colony path steps the same colonies one after another in
contexts of their own.  `value` is the batched mixed path's cell-steps/s.

A "step" = restore the initial state from a device-resident snapshot (a
device-to-device copy) + the step call.  The state is resident in HBM before
the timed region starts.

`--workload ensemble`: 128 independent colony packings (disc-shaped nematic
domains, standing in for grown colonies) with log-uniform sizes in
[500, 50000] (~1.5M cells) stepped 50 times (device division, then the step)
in mixed precision by one batched context; the state carries on from step to
step.  The untimed warm-up step relaxes the packings' initial overlaps.

`--workload soft`: the soft Hertz model on one 100,000-cell colony packing,
E* = 1e4, r_eff = 0.125, dt = 1e-3/30, 100 time-loop steps, mixed and FP64,
with the CPU restatement's soft step timed beside it.

`--workload seeds`: one colony per seed (default 8 seeds) stepped alone in
FP64 and in mixed precision; the mixed / FP64 step-time ratio split into the
per-iteration ratio and the BBPGD iteration-count ratio (geometric means).

`--workload serving`: colony-growth requests with Poisson arrivals (20/s for
10 s), 1k-100k cells each, hard or soft 50/50, growth rate U[0.5,1.5] ln2,
200 steps each, served by the continuous-batching loop (pcc.serving) under an
HBM budget and a per-model budget of resident cells; p50/p99 are over every
request's per-step latencies, and per size bucket over the steady steps.  The hard
colonies are relaxed by one untimed device step before the run, as growing
colonies would be (--no-relax submits the raw packings).

Prints ONE JSON line.  --gpus N > 1 exits: the library drives one GPU per
process and the benchmark measures one GPU.
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
    """Hard model of the configured workloads: rotational diffusion off
    (temperature 0), dt 1e-3, tolerances 1e-6, growth at ln 2."""
    return reference_params(dt=1e-3, tolerance=1e-6, tau=1 / math.log(2), lam=1e-2,
                            cutoff_phase0=0.05, cutoff_phaseK=0.05, relcp_tol=1e-6,
                            precision=precision, max_bbpgd_iterations=2_000_000,
                            temperature=0.0)


def run_single(colonies, precision, steps, warmup, timing=True):
    """The single-colony path: one context per colony, colonies one after
    another; per-kernel timing pass on the first colony (if `timing`)."""
    from pcc.engine import ParticleManager
    total, lat, results, phases, tim, tuning = 0.0, [], [], None, None, None
    for q, bodies in enumerate(colonies):
        pm = ParticleManager(params_for(precision), device=0)
        pm.set_state(bodies)
        pm.snapshot()
        for i in range(warmup):
            pm.restore()
            pm.step(i)
        for i in range(steps):
            a = time.perf_counter()
            pm.restore()
            r = pm.step(i)
            dt = time.perf_counter() - a
            lat.append(dt)
            total += dt
            results.append(r)
        if q == 0:
            phases = pm.phase_stats()
            tuning = pm.tuning()
        if q == 0 and timing:
            pm.set_timing(True)
            pm.restore()
            pm.step(0)
            tim = pm.timing()
            pm.set_timing(False)
        pm.close()
    return {"total_s": total, "lat": lat, "results": results, "timing": tim, "phases": phases,
            "tuning": tuning}


def run_batched(colonies, precision, steps, warmup, timing=True, restore=True, report=0,
                divide=False):
    """All colonies in one batched context; one latency per batch step
    (a progress line every `report` steps when report > 0).  divide: every
    step is the time-loop body, device division (seeded per colony and step)
    then the step."""
    from pcc.engine import ColonyBatch
    from pcc.serving import division_seed
    cb = ColonyBatch(params_for(precision), device=0)
    cb.set_state(colonies, [params_for(precision)] * len(colonies))
    if restore:
        cb.snapshot()

    def one(i):
        if divide:
            cb.divide([division_seed(q, i) for q in range(len(colonies))], 0, 0.0, 0.01)
        return cb.step(i)
    for i in range(warmup):
        if restore:
            cb.restore()
        one(i)
    lat, results, cell_steps = [], [], 0
    t0 = time.perf_counter()
    for i in range(steps):
        a = time.perf_counter()
        if restore:
            cb.restore()
        results.append(one(warmup + i))
        lat.append(time.perf_counter() - a)
        cell_steps += cb.n
        if report and (i + 1) % report == 0:
            print(f"[bench] step {i + 1}/{steps}: {lat[-1] * 1e3:.1f} ms", file=sys.stderr,
                  flush=True)
    total = time.perf_counter() - t0
    phases = [cb.phase_stats(q) for q in range(len(colonies))]
    tuning = cb.tuning()
    tim = None
    if timing:
        # per-kernel pass: HIP events around every BBPGD launch on the context
        # stream (plain launches, as event records cannot sit in a graph)
        cb.set_timing(True)
        if restore:
            cb.restore()
        cb.step(0)
        tim = cb.timing()
        cb.set_timing(False)
    cells_end = list(cb.sizes)
    cb.close()
    return {"total_s": total, "lat": lat, "results": results, "timing": tim, "phases": phases,
            "tuning": tuning, "cell_steps": cell_steps, "cells_end": cells_end}


PMC_SUMMARY = os.path.join(REPO, "profiles", "bbpgd_mixed_300k_colony_pmc.txt")


TRACE_SUMMARY = os.path.join(REPO, "profiles", "bbpgd_mixed_300k_colony_kernel_trace.txt")


def traced_from_profile(kernel, tuning):
    """Average duration (us) of `kernel` ("body_apply" with the variant the
    autotuner picked, or "constraint_apply") in the committed rocprofv3
    kernel-trace summary of this workload, or None when the summary holds no
    line for that kernel instantiation.  Events around each launch also time
    the launch's dispatch and its end-of-kernel flush, so the traced average
    is the shorter of the two."""
    if not os.path.exists(TRACE_SUMMARY):
        return None
    if kernel == "body_apply":
        name = (f"k_body_coop<float, true, {tuning['body_chunk']}, {tuning['body_threads']}, "
                f"{'true' if tuning['prefetch'] else 'false'}, 0>")
    else:
        name = "k_constr<float, true>"
    for line in open(TRACE_SUMMARY):
        if line.startswith("pcc::" + name + " ") and "avg_us=" in line:
            return float(line.rsplit("avg_us=", 1)[1])
    return None


def schedule(n_cells):
    """How a batch step runs colonies of n_cells cells: each ReLCP loop alone
    ("solo", from the library's threshold of 150000 cells or PCC_SOLO_CELLS)
    or all colonies in one lockstep loop."""
    solo = int(os.environ.get("PCC_SOLO_CELLS", "150000"))
    return "solo" if solo > 0 and n_cells >= solo else "lockstep"


def traffic_from_profile(kernel, instances, sched):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE passes, each scaled
    by the factor that maps a copy of known size onto its true bytes), or
    None when no summary for this configuration is committed."""
    if not os.path.exists(PMC_SUMMARY):
        return None
    for line in open(PMC_SUMMARY):
        f = line.split()
        if len(f) == 5 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel \
                and f[2] == f"instances={instances}" and f[3] == f"schedule={sched}":
            return float(f[4])
    return None


def latency_stats(lat):
    a = np.array(lat) * 1e3
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))


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


def cpu_baseline_check(seed, n=40_000):
    """The baseline's extrapolation checked against a measurement: on a
    smaller colony of the same generator, one full FP64 step of the CPU
    restatement (uncapped, to convergence) is timed and compared with the
    estimate that the same fit (capped runs of each phase) gives at that
    run's own per-phase iteration counts.  The Release-flag build rounds
    differently from the GPU FP64 path, so its iteration counts are reported
    beside the GPU's."""
    from pcc.engine import ParticleManager
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    b = G.dense_colony(n, seed=seed)
    pm = ParticleManager(params_for(PREC_FP64))
    pm.set_state(b)
    pm.step()
    gpu_phases = pm.phase_stats()
    pm.close()
    t = time.perf_counter()
    st, res, _, _, ph, _ = oracle.step(params_for(PREC_FP64), b, fast=True)
    measured = time.perf_counter() - t
    cpu_phases = [(int(x), 0) for x in ph]
    est = cpu_baseline(b, cpu_phases)
    return {"n_cells": n, "est_step_s": est["est_step_s"], "measured_step_s": measured,
            "est_over_measured": est["est_step_s"] / measured,
            "cpu_phase_iterations": [x for x, _ in cpu_phases],
            "gpu_phase_iterations": [int(x) for x, _ in gpu_phases], "cpu_status": int(st)}


def ensemble_colonies(n_colonies, seed, lo=500, hi=50_000):
    return G.ensemble(n_colonies, seed, lo, hi)


def ensemble_summary(n_colonies, seed, steps, warmup):
    """Config 4: the ensemble stepped `steps` times in one batched context."""
    cols = ensemble_colonies(n_colonies, seed)
    ncell = sum(c.n for c in cols)
    print(f"[bench] ensemble: {len(cols)} colonies, {ncell} cells, sizes "
          f"{min(c.n for c in cols)}..{max(c.n for c in cols)}", file=sys.stderr, flush=True)
    r = run_batched(cols, PREC_MIXED, steps, warmup, timing=True, restore=False, report=10,
                    divide=True)
    p50, p99 = latency_stats(r["lat"])
    its = [sum(x["bbpgd_iterations"] for x in step) for step in r["results"]]
    rf = roofline(r["timing"])
    return {
        "cell_steps_per_s": r["cell_steps"] / r["total_s"],
        "ms_per_step": r["total_s"] / steps * 1e3,
        "p50_step_ms": p50, "p99_step_ms": p99, "steps": steps,
        "colonies": len(cols), "n_cells": ncell, "n_cells_end": int(sum(r["cells_end"])),
        "bbpgd_iterations_per_step": {"min": int(min(its)), "max": int(max(its))},
        "roofline": {"bound": "hbm", "kernel": rf["kernel"], "achieved": rf["gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": rf["gbs"] / HBM_PEAK_GBS,
                     "traffic": None, "avg_launch_us": rf["avg_us"],
                     "algorithmic_bytes_per_launch": rf["bytes"]},
        "tuning": r["tuning"],
    }


def main_ensemble(args):
    steps = args.steps if args.steps_given else 50
    e = ensemble_summary(args.colonies, args.seed, steps, args.warmup)
    line = {
        "metric": "cell-steps/s, batched ensemble of independent colonies; p50/p99 step latency",
        "value": e["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": e["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"{e['colonies']} independent hard-model colony packings (disc "
                                f"nematic domains, relaxed by the warm-up step), sizes "
                                f"log-uniform in [500, 50000] ({e['n_cells']} cells), mixed "
                                f"precision, {steps} consecutive time-loop steps (device "
                                f"division seeded per colony and step, N(0, 0.01 rad) turns, "
                                f"then the step) in one batched context with captured BBPGD "
                                f"graphs and per-colony convergence"),
                   "n_cells": e["n_cells"], "n_cells_end": e["n_cells_end"],
                   "colonies": e["colonies"]},
        "p50_step_ms": e["p50_step_ms"],
        "p99_step_ms": e["p99_step_ms"],
        "bbpgd_iterations_per_step": e["bbpgd_iterations_per_step"],
        "roofline": e["roofline"],
        "tuning": e["tuning"],
    }
    print(json.dumps(line))


def geomean(v):
    return float(np.exp(np.mean(np.log(np.asarray(v, dtype=float)))))


def main_seeds(args):
    """mixed / FP64 over several colony seeds, split into the per-iteration
    time ratio (what the storage precision buys) and the BBPGD iteration
    count ratio (which differs between precisions from seed to seed): each
    seed's colony is stepped alone by the single-colony path in both
    precisions, one untimed step then one timed step."""
    per = []
    for s in range(args.seed, args.seed + args.seeds):
        b = G.dense_colony(args.n, seed=s)
        row = {"seed": s}
        for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
            r = run_single([b], prec, 1, 1, timing=False)
            res = r["results"][0]
            its = int(res["bbpgd_iterations"])
            row[name] = {"ms": r["total_s"] * 1e3, "bbpgd_iterations": its,
                         "us_per_iteration": r["total_s"] * 1e6 / max(its, 1),
                         "phases": r["phases"]}
        print(f"[bench] seed {s}: fp64 {row['fp64']['ms']:.1f} ms / {row['fp64']['bbpgd_iterations']}"
              f" its, mixed {row['mixed']['ms']:.1f} ms / {row['mixed']['bbpgd_iterations']} its",
              file=sys.stderr, flush=True)
        per.append(row)
    step = geomean([x["fp64"]["ms"] / x["mixed"]["ms"] for x in per])
    per_it = geomean([x["fp64"]["us_per_iteration"] / x["mixed"]["us_per_iteration"] for x in per])
    its = geomean([x["fp64"]["bbpgd_iterations"] / x["mixed"]["bbpgd_iterations"] for x in per])
    fp64_cs = args.n * len(per) / (sum(x["fp64"]["ms"] for x in per) / 1e3)
    mixed_cs = args.n * len(per) / (sum(x["mixed"]["ms"] for x in per) / 1e3)
    line = {
        "metric": "mixed / FP64 step-time ratio over colony seeds (single-colony path)",
        "value": step,
        "unit": "ratio",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": sum(x["mixed"]["ms"] for x in per) / len(per),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 vs f64",
        "data": "synthetic",
        "config": {"workload": (f"hard-model single step (full ReLCP) of one {args.n}-cell "
                                f"colony packing per seed (nematic domains, seeds {args.seed}.."
                                f"{args.seed + args.seeds - 1}), stepped alone in FP64 and in "
                                f"mixed precision; cutoff 0.05, BBPGD/ReLCP tol 1e-6, dt 1e-3"),
                   "n_cells": args.n, "seeds": args.seeds},
        "ratio_step_time_geomean": step,
        "ratio_per_iteration_geomean": per_it,
        "ratio_iterations_geomean": its,
        "fp64_cell_steps_per_s": fp64_cs,
        "mixed_cell_steps_per_s": mixed_cs,
        "per_seed": per,
    }
    print(json.dumps(line))


def serving_requests(rate, duration, seed, steps, lo=1_000, hi=100_000):
    """Poisson arrivals at `rate`/s over `duration` s; colony sizes
    log-uniform in [lo, hi]; hard or soft 50/50; growth rate U[0.5,1.5] ln2.
    Colonies are disc packings of nematic domains (the ensemble's generator)."""
    from pcc.abi_types import MODE_SOFT
    from pcc.serving import Request
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        rate_g = rng.uniform(0.5, 1.5) * math.log(2)
        soft = rng.uniform() < 0.5
        if soft:
            p = soft_params(PREC_MIXED).replace(tau=1 / rate_g)
        else:
            p = params_for(PREC_MIXED).replace(tau=1 / rate_g)
        reqs.append(Request(len(reqs), G.packed_colony(n, seed=seed * 100_000 + len(reqs)), p,
                            steps, arrival=t))
    return reqs


def soft_params(precision):
    """Config-3 soft model: F = (4/3) E* sqrt(r_eff) delta^1.5 with E* = 1e4 and
    r_eff = R1 R2/(R1 + R2) = 0.125 for d = 0.5 (kcc = 4/3 E*), dt 30x below
    the hard model's, growth at ln2."""
    from pcc.abi_types import MODE_SOFT
    return reference_params(mode=MODE_SOFT, dt=1e-3 / 30, tolerance=1e-6, tau=1 / math.log(2),
                            lam=1e-2, kcc=4.0 / 3.0 * 1e4, precision=precision,
                            temperature=0.0)


def run_soft(bodies, precision, steps, warmup):
    """`steps` time-loop steps (device division, then the soft step) of one
    colony; per-kernel timing on one extra step."""
    from pcc.engine import ParticleManager
    from pcc.serving import division_seed
    pm = ParticleManager(soft_params(precision), device=0)
    pm.set_state(bodies)
    for i in range(warmup):
        pm.divide(division_seed(0, i), 0, 0.0, 0.01)
        pm.step(i)
    lat, cell_steps = [], 0
    t0 = time.perf_counter()
    for i in range(warmup, warmup + steps):
        a = time.perf_counter()
        pm.divide(division_seed(0, i), 0, 0.0, 0.01)
        r = pm.step(i)
        lat.append(time.perf_counter() - a)
        cell_steps += pm.n
    total = time.perf_counter() - t0
    pm.set_timing(True)
    pm.step(warmup + steps)
    tim = pm.timing()
    pm.set_timing(False)
    n_end = pm.n
    pm.close()
    return {"total_s": total, "lat": lat, "cell_steps": cell_steps, "timing": tim,
            "n_constraints": r["n_constraints"], "max_overlap": r["max_overlap"],
            "n_end": n_end}


def kernel_table(tim):
    out = {}
    for name, t in tim.items():
        if not t["launches"]:
            continue
        avg = t["total_ms"] / t["launches"] / 1e3
        row = {"launches": t["launches"], "avg_us": avg * 1e6, "total_ms": t["total_ms"]}
        if t["bytes_per_launch"] > 0:
            row["algorithmic_bytes_per_launch"] = t["bytes_per_launch"]
            row["gbs"] = t["bytes_per_launch"] / avg / 1e9
            row["frac"] = row["gbs"] / HBM_PEAK_GBS
        out[name] = row
    return out


def cpu_soft_baseline(bodies, reps=3):
    """The FP64 oracle (Release flags + OpenMP) on this host's cores: `reps`
    soft steps of the same colony, timed."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    p = soft_params(PREC_FP64)
    b = bodies
    oracle.step(p, b, 0, fast=True)  # warm
    t = time.perf_counter()
    for i in range(reps):
        b = oracle.step(p, b, i, fast=True)[2]
    dt = (time.perf_counter() - t) / reps
    return {"value": bodies.n / dt, "unit": "cell-steps/s", "cores": oracle.threads(fast=True),
            "kind": "port",
            "sample": (f"FP64 CPU restatement (-O3 -ffast-math, OpenMP, "
                       f"{oracle.threads(fast=True)} threads): {reps} soft steps of the same "
                       f"{bodies.n}-cell colony, {dt * 1e3:.1f} ms per step (no division)")}


def soft_summary(n, seed, steps, warmup, fp64=True, cpu=False):
    """Config 3: `steps` soft time-loop steps of an n-cell colony packing."""
    b = G.dense_colony(n, seed=seed)
    print(f"[bench] soft: n={b.n}, {steps} steps", file=sys.stderr, flush=True)
    rm = run_soft(b, PREC_MIXED, steps, warmup)
    p50, p99 = latency_stats(rm["lat"])
    km = kernel_table(rm["timing"])
    h = km.get("hertz", {})
    out = {
        "cell_steps_per_s": rm["cell_steps"] / rm["total_s"],
        "ms_per_step": rm["total_s"] / steps * 1e3,
        "p50_step_ms": p50, "p99_step_ms": p99, "steps": steps,
        "n_cells": b.n, "n_cells_end": rm["n_end"], "n_contacts": rm["n_constraints"],
        "max_overlap": rm["max_overlap"],
        "roofline": {"bound": "hbm", "kernel": "hertz", "achieved": h.get("gbs"),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": h.get("frac"),
                     "traffic": None, "avg_launch_us": h.get("avg_us"),
                     "algorithmic_bytes_per_launch": h.get("algorithmic_bytes_per_launch")},
        "kernels": km,
    }
    if fp64:
        r64 = run_soft(b, PREC_FP64, steps, warmup)
        out["fp64"] = {"cell_steps_per_s": r64["cell_steps"] / r64["total_s"],
                       "ms_per_step": r64["total_s"] / steps * 1e3,
                       "kernels": kernel_table(r64["timing"])}
    if cpu:
        out["cpu_baseline"] = cpu_soft_baseline(b)
    return out


def main_soft(args):
    from pcc.engine import copy_bandwidth
    n = args.n if args.n != 300_000 else 100_000
    steps = args.steps if args.steps_given else 100
    copy_gbs = copy_bandwidth(1 << 30, 10)
    sm = soft_summary(n, args.seed, steps, args.warmup, fp64=True, cpu=not args.no_cpu)
    rf = dict(sm["roofline"])
    rf["achievable_copy_gbs"] = copy_gbs
    line = {
        "metric": "cell-steps/s, soft Hertz model; p50/p99 step latency",
        "value": sm["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": sm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"soft Hertz model, one {sm['n_cells']}-cell colony packing "
                                f"(nematic domains), F = (4/3) E* sqrt(r_eff) delta^1.5 with "
                                f"E* = 1e4, r_eff = 0.125, dt = 1e-3/30, {steps} time-loop steps "
                                f"(device division, then the step), Hertz sums FP32, "
                                f"integration FP64"),
                   "n_cells": sm["n_cells"], "n_cells_end": sm["n_cells_end"],
                   "n_contacts": sm["n_contacts"]},
        "p50_step_ms": sm["p50_step_ms"],
        "p99_step_ms": sm["p99_step_ms"],
        "roofline": rf,
        "kernels": sm["kernels"],
        "fp64": sm["fp64"],
        "max_overlap": sm["max_overlap"],
    }
    if "cpu_baseline" in sm:
        line["cpu_baseline"] = sm["cpu_baseline"]
    print(json.dumps(line))


def relax_requests(reqs, group=16):
    """One untimed step of every hard request's initial colony on the device
    (batches of `group` colonies of similar size, each with its own
    parameters), standing in for colonies that have been growing: the
    synthetic packings start with 1e-3..1e-2 overlaps whose first step costs
    tens of thousands of BBPGD iterations.  Returns the seconds it took."""
    from pcc.abi_types import MODE_HARD
    from pcc.engine import ColonyBatch
    hard = sorted((r for r in reqs if r.params.mode == MODE_HARD), key=lambda r: r.bodies.n)
    t = time.perf_counter()
    for k in range(0, len(hard), group):
        part = hard[k:k + group]
        cb = ColonyBatch(part[0].params, device=0)
        cb.set_state([r.bodies for r in part], [r.params for r in part])
        cb.step(0)
        for q, r in enumerate(part):
            r.bodies = cb.get_state(q)
        cb.close()
        print(f"[bench] relaxed hard requests {k + len(part)}/{len(hard)}", file=sys.stderr,
              flush=True)
    return time.perf_counter() - t


def main_serving(args):
    from pcc.serving import ColonyServer
    reqs = serving_requests(args.rate, args.duration, args.seed, args.serve_steps)
    relax_s = relax_requests(reqs) if args.relax else 0.0
    ncell = sum(r.bodies.n for r in reqs)
    print(f"[bench] serving: {len(reqs)} requests over {args.duration:.0f} s, {ncell} cells, "
          f"{sum(r.params.mode == 1 for r in reqs)} soft", file=sys.stderr, flush=True)
    srv = ColonyServer(int(args.hbm_budget_gb * 2**30), precision=PREC_MIXED,
                       warm_steps=args.warm_steps)
    st = srv.run(reqs, time_limit=args.duration + args.serve_limit,
                 progress=lambda m: print(f"[serving] {m}", file=sys.stderr, flush=True))
    batches = srv.step_summary()
    srv.close()
    waits = [r.admitted - r.arrival for r in srv.results.values()]
    # per-step latency of completed requests, split into warm-up and steady steps
    warm = [x for r in srv.results.values() for x in r.step_ms[:args.warm_steps]]
    steady = [x for r in srv.results.values() for x in r.step_ms[args.warm_steps:]]
    grown = [r.state.n / r.n_cells for r in srv.results.values() if r.n_cells]
    # every recorded step (completed and still-resident requests) per model
    per_mode = {0: [], 1: []}
    for r in srv.results.values():
        per_mode[r.mode].extend(r.step_ms)
    for key, ms in srv.members.items():
        for m in ms:
            per_mode[key[0]].extend(m.step_ms)
    modes = {}
    for mode, name in ((0, "hard"), (1, "soft")):
        x = per_mode[mode]
        done = [r for r in srv.results.values() if r.mode == mode]
        # steady steps (after the warm-up batch) per size bucket: the batches
        # whose step times a colony of that size sees
        buckets = {}
        for bk in range(len(srv.bucket_edges) + 1):
            y = [v for r in srv.results.values() if r.mode == mode and srv._bucket(r.n_cells) == bk
                 for v in r.step_ms[args.warm_steps:]]
            y += [v for key, ms in srv.members.items() if key[0] == mode and key[1] == bk
                  for m in ms for v in m.step_ms[args.warm_steps:]]
            if y:
                buckets[str(bk)] = {"steps": len(y), "p50_step_ms": float(np.percentile(y, 50)),
                                    "p99_step_ms": float(np.percentile(y, 99))}
        mw = [r.admitted - r.arrival for r in done]
        modes[name] = {"requests": sum(q.params.mode == mode for q in reqs),
                       "completed": len(done), "steps_recorded": len(x),
                       "p50_step_ms": float(np.percentile(x, 50)) if x else 0.0,
                       "p99_step_ms": float(np.percentile(x, 99)) if x else 0.0,
                       "steady_by_bucket": buckets,
                       "admission_wait_s": {"p50": float(np.percentile(mw, 50)) if mw else 0.0,
                                            "max": float(max(mw)) if mw else 0.0}}
    line = {
        "metric": "cell-steps/s, continuous-batching colony server; p50/p99 per-step latency",
        "value": st["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.serve_steps,
        "warmup": 0,
        "ms_per_step": st["p50_step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"Poisson arrivals at {args.rate:g} requests/s for "
                                f"{args.duration:g} s, colony packings of 1k-100k cells "
                                f"(log-uniform), "
                                f"hard or soft 50/50, growth rate U[0.5,1.5] ln2, "
                                f"{args.serve_steps} steps each (divide, then step), HBM "
                                f"budget {args.hbm_budget_gb:g} GiB, resident-cell compute "
                                f"budget {srv.max_resident}, mixed precision"),
                   "requests": len(reqs), "n_cells": ncell, "warm_steps": args.warm_steps,
                   "initial_states": ("hard colonies relaxed by one untimed device step "
                                      f"({relax_s:.1f} s before the run)" if args.relax
                                      else "unrelaxed packings"),
                   "division": "device, seeded per request and step, exact halves, "
                               "N(0, 0.01 rad) turns"},
        "p50_step_ms": st["p50_step_ms"],
        "p99_step_ms": st["p99_step_ms"],
        "serving": {k: st[k] for k in ("completed", "cell_steps", "wall_s", "batch_steps",
                                       "peak_reserved_bytes", "peak_device_bytes",
                                       "budget_bytes", "peak_resident_cells",
                                       "max_resident_cells")},
        "bucket_edges_cells": list(srv.bucket_edges),
        "admission_wait_s": {"p50": float(np.percentile(waits, 50)) if waits else 0.0,
                             "max": float(max(waits)) if waits else 0.0},
        "completed_step_ms": {
            "warm_up": {"p50": float(np.percentile(warm, 50)) if warm else 0.0,
                        "p99": float(np.percentile(warm, 99)) if warm else 0.0},
            "steady": {"p50": float(np.percentile(steady, 50)) if steady else 0.0,
                       "p99": float(np.percentile(steady, 99)) if steady else 0.0}},
        "growth_factor": {"mean": float(np.mean(grown)) if grown else 0.0,
                          "max": float(max(grown)) if grown else 0.0},
        "per_model": modes,
        # per batch "model-bucket-stage" (model 0 hard, 1 soft; stage 0 warm-up, 1 steady)
        "batches": batches,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--warm-steps", type=int, default=2,
                    help="serving: steps a new colony runs in the warm-up batch")
    ap.add_argument("--no-relax", dest="relax", action="store_false",
                    help="serving: submit the hard colonies unrelaxed")
    ap.add_argument("--n", type=int, default=300_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--instances", type=int, default=2)
    ap.add_argument("--colonies", type=int, default=128)
    ap.add_argument("--placement", choices=["nematic", "iid"], default="nematic",
                    help="headline colonies: nematic-domain packings, or i.i.d. centres at "
                         "area fraction 0.9")
    ap.add_argument("--seeds", type=int, default=8,
                    help="--workload seeds: number of colony seeds")
    ap.add_argument("--workload", choices=["colony300k", "ensemble", "serving", "soft", "seeds"],
                    default="colony300k")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=10.0)
    ap.add_argument("--serve-steps", type=int, default=200)
    ap.add_argument("--hbm-budget-gb", type=float, default=128.0)
    ap.add_argument("--serve-limit", type=float, default=120.0,
                    help="seconds of serving after the last arrival before the run stops")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--no-extras", dest="extras", action="store_false",
                    help="headline: skip the soft (100k) and ensemble (128 colonies) summaries")
    args = ap.parse_args()
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 3
    if args.gpus != 1:
        print(f"bench.py: the benchmark measures one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        return main_ensemble(args)
    if args.workload == "serving":
        return main_serving(args)
    if args.workload == "soft":
        return main_soft(args)
    if args.workload == "seeds":
        return main_seeds(args)
    from pcc.engine import copy_bandwidth, read_bandwidth
    copy_gbs = copy_bandwidth(1 << 30, 10)
    read_gbs = read_bandwidth(1 << 31, 10)
    make = G.random_disc if args.placement == "iid" else G.dense_colony
    cols = [make(args.n, seed=args.seed + q) for q in range(args.instances)]
    bodies = cols[0]
    ncell = sum(c.n for c in cols)
    print(f"[bench] {len(cols)} x n={bodies.n} R={G.disc_radius(bodies):.1f} "
          f"area_fraction={G.area_fraction(bodies):.3f}", file=sys.stderr, flush=True)
    out = {}
    if args.only in ("both", "fp64"):
        r = run_single(cols, PREC_FP64, args.steps, args.warmup)
        p50, p99 = latency_stats(r["lat"])
        out["fp64"] = {
            "cell_steps_per_s": ncell * args.steps / r["total_s"],
            "ms_per_step": r["total_s"] / (args.steps * len(cols)) * 1e3,
            "p50_ms": p50, "p99_ms": p99,
            "bbpgd_iterations": [x["bbpgd_iterations"] for x in r["results"][::args.steps]],
            "n_constraints": [x["n_constraints"] for x in r["results"][::args.steps]],
            "roofline": roofline(r["timing"]),
            "phases": r["phases"],
            "tuning": r["tuning"],
        }
        print(f"[bench] fp64 single-colony path: {out['fp64']['ms_per_step']:.1f} ms per colony-step, "
              f"iterations {out['fp64']['bbpgd_iterations']}", file=sys.stderr, flush=True)
    if args.only in ("both", "mixed"):
        r = run_batched(cols, PREC_MIXED, args.steps, args.warmup)
        p50, p99 = latency_stats(r["lat"])
        last = r["results"][-1]
        out["mixed"] = {
            "cell_steps_per_s": ncell * args.steps / r["total_s"],
            "ms_per_step": r["total_s"] / args.steps * 1e3,
            "p50_ms": p50, "p99_ms": p99,
            "bbpgd_iterations": [x["bbpgd_iterations"] for x in last],
            "n_constraints": [x["n_constraints"] for x in last],
            "roofline": roofline(r["timing"]),
            "kernels": r["timing"],
            "phases": r["phases"][0],
            "tuning": r["tuning"],
        }
        print(f"[bench] mixed batched ({len(cols)} colonies): {out['mixed']['ms_per_step']:.1f} ms "
              f"per batch step, iterations {out['mixed']['bbpgd_iterations']}",
              file=sys.stderr, flush=True)
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
            "workload": (f"hard-model single step (full ReLCP) of {len(cols)} independent "
                         f"{bodies.n}-cell colony packings ("
                         f"{'i.i.d. centres' if args.placement == 'iid' else 'nematic domains'}"
                         f", area fraction "
                         f"{G.area_fraction(bodies):.3f}, R={G.disc_radius(bodies):.0f}, seeds "
                         f"{args.seed}..{args.seed + len(cols) - 1}) batched in one context "
                         f"(colonies of >= 150000 cells run their ReLCP loops one after "
                         f"another inside the batch step); "
                         f"lengths U[1,2], d=0.5, cutoff 0.05, BBPGD/ReLCP tol 1e-6, dt 1e-3"),
            "n_cells": bodies.n,
            "instances": len(cols),
            "precision": "mixed (FP32 bulk applies, FP64 reductions/integration)"
            if "mixed" in out else "fp64",
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
            "traffic": traffic_from_profile(rf["kernel"], len(cols), schedule(bodies.n)),
            "avg_launch_us": rf["avg_us"],
            "algorithmic_bytes_per_launch": rf["bytes"],
            "achievable_copy_gbs": copy_gbs,
            "achievable_read_gbs": read_gbs,
            "frac_of_achievable_read": rf["gbs"] / read_gbs,
        },
    }
    if "mixed" in out:
        k = out["mixed"]["kernels"]
        line["kernels"] = {}
        for name in ("body_apply", "constraint_apply"):
            t = k[name]
            if t["launches"]:
                avg = t["total_ms"] / t["launches"] / 1e3
                line["kernels"][name] = {"launches": t["launches"], "avg_us": avg * 1e6,
                                         "algorithmic_bytes_per_launch": t["bytes_per_launch"],
                                         "gbs": t["bytes_per_launch"] / avg / 1e9,
                                         "frac": t["bytes_per_launch"] / avg / 1e9 / HBM_PEAK_GBS}
                # the same kernel's average in the committed rocprofv3 trace
                # of this workload, when that trace has its instantiation
                tr = traced_from_profile(name, out["mixed"]["tuning"]) \
                    if schedule(bodies.n) == "solo" and len(cols) == 2 else None
                if tr:
                    line["kernels"][name]["traced_avg_us"] = tr
                    line["kernels"][name]["traced_frac"] = \
                        t["bytes_per_launch"] / (tr * 1e-6) / 1e9 / HBM_PEAK_GBS
        kr = line["kernels"].get(rf["kernel"], {})
        if "traced_avg_us" in kr:
            line["roofline"]["traced_avg_us"] = kr["traced_avg_us"]
            line["roofline"]["traced_frac"] = kr["traced_frac"]
    if "fp64" in out and "mixed" in out:
        line["fp64"] = {k: out["fp64"][k] for k in ("cell_steps_per_s", "ms_per_step", "p50_ms",
                                                    "p99_ms", "bbpgd_iterations", "n_constraints",
                                                    "roofline")}
        line["mixed_over_fp64"] = out["mixed"]["cell_steps_per_s"] / out["fp64"]["cell_steps_per_s"]
        line["mixed_bbpgd_iterations"] = out["mixed"]["bbpgd_iterations"]
    line["phases"] = {k: v["phases"] for k, v in out.items()}
    # body-kernel variant picked by the autotuner, with every candidate's time
    line["tuning"] = {k: v["tuning"] for k, v in out.items()}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(bodies, (out.get("fp64") or head)["phases"])
        line["cpu_baseline"]["check"] = cpu_baseline_check(args.seed)
    if args.extras:
        # the other configured workloads, at their configured sizes and step
        # counts, as summaries beside the headline
        sm = soft_summary(100_000, args.seed, 100, 1, fp64=False)
        line["soft_100k"] = {k: sm[k] for k in ("cell_steps_per_s", "ms_per_step",
                                                "p50_step_ms", "p99_step_ms", "steps",
                                                "n_cells", "n_contacts", "roofline")}
        line["soft_100k"]["workload"] = ("soft Hertz model, 100,000-cell colony packing, "
                                         "E* = 1e4, r_eff = 0.125, dt = 1e-3/30, 100 time-loop "
                                         "steps (device division, then the step), mixed")
        e = ensemble_summary(128, args.seed, 50, 1)
        line["ensemble_128"] = {k: e[k] for k in ("cell_steps_per_s", "ms_per_step",
                                                  "p50_step_ms", "p99_step_ms", "steps",
                                                  "colonies", "n_cells", "n_cells_end",
                                                  "roofline")}
        line["ensemble_128"]["workload"] = ("128 hard-model colony packings, log-uniform "
                                            "[500, 50000] cells, 50 time-loop steps (device "
                                            "division, then the step), mixed, one batched "
                                            "context with captured BBPGD graphs")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
