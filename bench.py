#!/usr/bin/env python3
"""Benchmark of the per-step physics path: particle-steps/s of the full KDK step.

Headline workload (one GPU): the batched sweep of 64 independent impact scenarios
(per-scenario N log-uniform in [20k, 500k], 9,602,616 particles in all; angle U[0, 75] deg,
v U[1, 3] v_esc, impactor/target mass ratio U[0.01, 0.2], Fe/granite layering, seeds
splitmix64(42 + k)), advanced together: one hipGraph-replayed KDK step over all segments,
each scenario with its own Morton box, tree and adaptive dt, FP32 arithmetic with FP64
leapfrog position/velocity accumulators (mixed precision).  A step is one full KDK step
(dt -> kick -> drift -> sort/tree -> pass 1 + EOS -> gravity -> pass 2 -> kick + EOS) of every
particle; inputs are resident in HBM when the timed region starts.

Beside it, the plain FP32 forms the metric is compared with:
  * the same 64 scenarios run one after another as single-scenario plain FP32 runs, each
    replayed from its own hipGraph (the sum of their timed regions);
  * the 1,000,000-particle giant impact (900k + 100k, 45 deg, 1.1 v_esc) as one plain FP32
    scenario.
Timed regions: K steps bracketed by stream synchronisation.  A second region of K eager steps
of the batched run records HIP events around every stage on the context stream; it gives the
per-stage times and the gravity kernel's average launch duration used for the roofline.  The
CPU baseline times the FP64 OpenMP oracle on a bounded sample of the batched workload (the
first scenarios of the sweep, one step after their initial force evaluation).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "astraea-impact-sim_amd"))

METRIC = "particle-steps/s (full KDK step), batched mixed vs plain FP32, ≤1e-4 energy drift"
MAX_DT = 5.0 / 1.0e3          # Simulation::step default maxDtSec = 5 s (PSU time)
PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 peak
FLOP_PER_INTERACTION = 23     # 20 acceleration (GPU Gems 3 N-body convention) + 3 potential
PEAK_HBM_GBS = 8000.0         # MI355X published HBM3E bandwidth
# algorithmic bytes per particle of the stages the HBM roofline is asked for (compulsory traffic,
# neighbour records cache-served, FP32 state): sort = 8 + 24 per 8-bit digit pass
ALG_BYTES = {"sort": 200, "pass1": 126, "pass2": 151, "drift": 88}
TRAFFIC_PROFILE = ROOT / "profiles" / "pmc_gravity_batched_sweep.md"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def energy_total(e):
    """K + Phi + U + radiated, per segment (rows of astraea_energy)."""
    e = np.atleast_2d(e)
    return e[:, 1] + e[:, 2] + e[:, 3] + e[:, 10]


def gravity_traffic():
    """HBM bytes per gravity launch of this workload from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-kernel calibration), or None."""
    try:
        text = TRAFFIC_PROFILE.read_text()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def time_steps(ctx, steps, graph=True):
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(MAX_DT, steps, graph=graph)
    ctx.synchronize()
    return time.perf_counter() - t0


def run_batched(A, state, counts, steps, warmup, autotune=True):
    from astraea_amd import scenario as S

    n = int(sum(counts))
    ctx = A.Context(capacity=n, max_segments=len(counts))
    ctx.set_params(S.plain_params(precision=A.PREC_FP64_ACCUM))
    ctx.set_layout(counts)
    ctx.load_state(state)
    log(f"[bench] batched: initial forces ({len(counts)} scenarios, {n:,} particles)")
    ctx.initial_forces()
    ctx.synchronize()
    e_init = energy_total(ctx.energy())
    tune = None
    if autotune:
        # untimed like the warm-up: four eager steps, one per launch variant of each pass
        log("[bench] batched: autotune (8 steps)")
        tune = ctx.autotune(MAX_DT, steps_per_variant=1)
        log(f"[bench] batched: tuning {tune['chosen']}")
    ctx.step(MAX_DT, warmup, graph=True)
    ctx.synchronize()
    e_start = energy_total(ctx.energy())
    ctx.counters()
    log(f"[bench] batched: timed {steps} steps")
    wall = time_steps(ctx, steps)
    e_end = energy_total(ctx.energy())
    cnt = ctx.counters()
    drift = (e_end - e_start) / np.maximum(np.abs(e_init), 1.0)

    log(f"[bench] batched: instrumented {steps} steps")
    ctx.enable_stage_timing(True)
    ctx.step(MAX_DT, steps, graph=False)
    ctx.synchronize()
    stages = ctx.stage_times()
    ctx.enable_stage_timing(False)
    cnt2 = ctx.counters()
    return dict(wall=wall, drift=drift, cnt=cnt, cnt2=cnt2, stages=stages, n=n, tune=tune)


def run_sequential_plain(A, states, counts, steps, warmup):
    """Each scenario alone, plain FP32, one after another; returns the summed timed seconds."""
    from astraea_amd import scenario as S

    ctx = A.Context(capacity=max(counts))
    ctx.set_params(S.plain_params())
    total = 0.0
    for st, n in zip(states, counts):
        ctx.set_layout([n])
        ctx.load_state(st)
        ctx.initial_forces()
        ctx.step(MAX_DT, warmup, graph=True)
        total += time_steps(ctx, steps)
    return total


def run_single_plain(A, steps, warmup, n_target, n_impactor):
    from astraea_amd import scenario as S

    state = S.config_giant(n_target, n_impactor)
    n = n_target + n_impactor
    ctx = A.Context(capacity=n)
    ctx.set_params(S.plain_params())
    ctx.set_layout([n])
    ctx.load_state(state)
    ctx.initial_forces()
    ctx.step(MAX_DT, warmup, graph=True)
    return n * steps / time_steps(ctx, steps)


def cpu_baseline(states, counts, params, min_particles):
    """FP64 OpenMP oracle on the first scenarios of the sweep (at least `min_particles` in all):
    initial forces, then one full step timed."""
    import oracle

    k, n = 0, 0
    while k < len(counts) and n < min_particles:
        n += counts[k]
        k += 1
    merged = {f: np.ascontiguousarray(np.concatenate([s[f] for s in states[:k]])) for f in states[0]}
    o = oracle.Oracle(params, group_size=params.group_size or 64)
    o.set_layout(counts[:k])
    o.load_state(merged)
    o.initial_forces()
    t0 = time.perf_counter()
    o.step(MAX_DT, 1)
    dt = time.perf_counter() - t0
    return {
        "value": n / dt,
        "unit": "particle-steps/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"1 full KDK step of the first {k} scenarios of the sweep ({n:,} particles) after their initial "
                  f"force evaluation, FP64, {oracle.num_threads()} OpenMP threads ({dt:.1f} s)",
    }


def drift_vs_oracle(A, steps, specs):
    """Energy drift of the batched mixed-precision step against the FP64 oracle over `steps` steps,
    on a scaled-down sweep the oracle finishes in about a minute.  Both start from the same state
    and take their own adaptive dt; per scenario, drift = (E(t) - E(0)) / max(|E(0)|, 1) with
    E = K + Phi + U + E_rad, and the figure of merit is |drift_gpu - drift_fp64|."""
    import oracle
    from astraea_amd import scenario as S

    state, counts = S.build_sweep(specs)
    params = S.plain_params(precision=A.PREC_FP64_ACCUM)
    ctx = A.Context(capacity=int(sum(counts)), max_segments=len(counts))
    ctx.set_params(params)
    ctx.set_layout(counts)
    ctx.load_state(state)
    ctx.initial_forces()
    ctx.synchronize()
    g0 = energy_total(ctx.energy())
    ctx.step(MAX_DT, steps, graph=True)
    ctx.synchronize()
    g1 = energy_total(ctx.energy())
    o = oracle.Oracle(params, group_size=params.group_size or 64)
    o.set_layout(counts)
    o.load_state(state)
    o.initial_forces()
    r0 = energy_total(o.energy())
    t0 = time.perf_counter()
    o.step(MAX_DT, steps)
    t_oracle = time.perf_counter() - t0
    r1 = energy_total(o.energy())
    dg = (g1 - g0) / np.maximum(np.abs(g0), 1.0)
    dr = (r1 - r0) / np.maximum(np.abs(r0), 1.0)
    diff = np.abs(dg - dr)
    return {
        "steps": steps,
        "scenarios": len(counts),
        "particles": int(sum(counts)),
        "max_abs_vs_fp64": float(diff.max()),
        "median_abs_vs_fp64": float(np.median(diff)),
        "gpu_max_abs": float(np.abs(dg).max()),
        "fp64_max_abs": float(np.abs(dr).max()),
        "oracle_seconds": t_oracle,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scenarios", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 comparison runs")
    ap.add_argument("--cpu-particles", type=int, default=1_000_000)
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="steps of the energy-drift comparison with the FP64 oracle (0: skip)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch variants instead of timing them before the warm-up")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the batched scenarios would shard "
              "whole scenarios across GPUs, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    import astraea_amd as A
    from astraea_amd import scenario as S

    K, W = args.steps, args.warmup
    specs = S.sweep_scenarios(args.scenarios)
    log(f"[bench] building {len(specs)} scenarios")
    states, counts = S.build_sweep(specs, split=True)
    merged = {f: np.ascontiguousarray(np.concatenate([s[f] for s in states])) for f in states[0]}

    b = run_batched(A, merged, counts, K, W, autotune=not args.no_autotune)
    n = b["n"]
    value = n * K / b["wall"]
    stages = b["stages"]
    grav_ms, grav_n = stages["gravity"]
    grav_avg_ms = grav_ms / max(grav_n, 1)
    inter_per_launch = b["cnt2"]["interactions"] / max(grav_n, 1)
    achieved = FLOP_PER_INTERACTION * inter_per_launch / (grav_avg_ms * 1e-3) / 1e12
    drift = b["drift"]

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * b["wall"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 leapfrog accumulators)",
        "data": "synthetic",
        "config": {
            "workload": f"batched sweep of {len(counts)} impact scenarios ({min(counts):,}-{max(counts):,} particles "
                        f"each, {n:,} in all), Fe/granite, angle U[0,75] deg, v U[1,3] v_esc, mass ratio "
                        "U[0.01,0.2]; one hipGraph-replayed KDK step over all segments, FP32 with FP64 "
                        "position/velocity accumulators",
            "scenarios": len(counts),
            "particles": n,
            "theta": 0.5,
            "softening": "h",
            "cfl": 0.3,
            "gravity_group": 64,
            "parallelism": "single GPU",
        },
        "energy_drift_rel": {"max_abs": float(np.max(np.abs(drift))), "median_abs": float(np.median(np.abs(drift))),
                             "steps": K},
        "interactions_per_particle": b["cnt"]["interactions"] / (n * K),
        "gravity_forced_accepts": int(b["cnt"]["forced_accepts"]),
        "neighbour_candidates_per_particle": b["cnt2"]["nbr_entries"] / n,
        "stage_ms": {k: (v[0] / max(v[1], 1)) for k, v in stages.items()},
        "hbm": {k: {"GB/s": n * ALG_BYTES[k] / (stages[k][0] / max(stages[k][1], 1) * 1e-3) / 1e9,
                    "frac": n * ALG_BYTES[k] / (stages[k][0] / max(stages[k][1], 1) * 1e-3) / 1e9 / PEAK_HBM_GBS,
                    "bytes_per_particle": ALG_BYTES[k]} for k in ALG_BYTES},
        "gravity_interactions_per_s": inter_per_launch / (grav_avg_ms * 1e-3),
        "tuning": b["tune"] and {"chosen": b["tune"]["chosen"], "steps": 8,
                                 "ms_per_step": {k: b["tune"][k] for k in ("pass1_ms", "pass2_ms", "drift_ms")}},
        "roofline": {
            "bound": "valu",
            "kernel": "kGravity",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": gravity_traffic(),
            "launch_ms": grav_avg_ms,
            "flop_per_launch": FLOP_PER_INTERACTION * inter_per_launch,
        },
    }
    if not args.no_plain:
        log("[bench] plain FP32: scenarios one after another")
        seq = n * K / run_sequential_plain(A, states, counts, K, W)
        log("[bench] plain FP32: 1,000,000-particle giant impact")
        single = run_single_plain(A, K, W, 900_000, 100_000)
        out["plain_fp32"] = {
            "sequential_scenarios": seq,
            "giant_impact_1m": single,
            "unit": "particle-steps/s",
        }
        out["speedup_vs_sequential_plain"] = value / seq
        out["speedup_vs_giant_impact_plain"] = value / single
    if args.drift_steps > 0:
        log(f"[bench] energy drift vs the FP64 oracle, {args.drift_steps} steps")
        out["energy_drift_vs_fp64_oracle"] = drift_vs_oracle(
            A, args.drift_steps, S.sweep_scenarios(8, n_min=1500, n_max=3000))
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(states, counts, S.plain_params(precision=A.PREC_FP64_ACCUM),
                                           args.cpu_particles)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
