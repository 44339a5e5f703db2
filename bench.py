#!/usr/bin/env python3
"""Benchmark of the per-step physics path: particle-steps/s of the full KDK step.

Workload (one GPU): the 1,000,000-particle giant impact (900k Fe/granite target + 100k
impactor, 45 deg, 1.1 v_esc, Plummer softening = h, CFL 0.3, strength off), plain FP32
state.  A step is one full KDK step (dt -> kick -> drift -> sort/tree -> pass 1 + EOS ->
gravity -> pass 2 -> kick + EOS) of every particle; inputs are resident in HBM.

Timed region: K steps replayed from a captured hipGraph on the context stream, bracketed by
stream synchronisation.  A second region of K eager steps records HIP events around every
stage on the same stream; it gives the per-stage times and the gravity kernel's average
launch duration used for the roofline.  The CPU baseline times the FP64 OpenMP oracle on a
bounded sample of the same workload (rank 0, N = 1).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "astraea-impact-sim_amd"))

METRIC = "particle-steps/s (full KDK step), batched mixed vs plain FP32, ≤1e-4 energy drift"
MAX_DT = 5.0 / 1.0e3          # Simulation::step default maxDtSec = 5 s (PSU time)
PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 peak
PEAK_HBM_GBS = 8000.0
FLOP_PER_INTERACTION = 23     # 20 acceleration (GPU Gems 3 N-body convention) + 3 potential


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def energy_total(e):
    return e[1] + e[2] + e[3] + e[10]


def cpu_baseline(state, counts, params, sample_steps=1):
    """FP64 OpenMP oracle on the same workload: initial forces, then `sample_steps` full steps timed."""
    import oracle

    o = oracle.Oracle(params, group_size=64)
    o.set_layout(counts)
    o.load_state(state)
    o.initial_forces()
    t0 = time.perf_counter()
    o.step(MAX_DT, sample_steps)
    dt = time.perf_counter() - t0
    n = int(sum(counts))
    return {
        "value": n * sample_steps / dt,
        "unit": "particle-steps/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"{sample_steps} full KDK step(s) of the {n:,}-particle giant impact after the initial force "
                  f"evaluation, FP64, {oracle.num_threads()} OpenMP threads ({dt:.1f} s)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--n-target", type=int, default=900_000)
    ap.add_argument("--n-impactor", type=int, default=100_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the batched scenarios would shard "
              "whole scenarios across GPUs, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    import astraea_amd as A
    from astraea_amd import scenario as S

    n_t, n_i = args.n_target, args.n_impactor
    n = n_t + n_i
    log(f"[bench] building giant impact {n_t}+{n_i}")
    state = S.config_giant(n_t, n_i)
    params = S.plain_params()

    ctx = A.Context(capacity=n)
    ctx.set_params(params)
    ctx.set_layout([n])
    ctx.load_state(state)
    log("[bench] initial force evaluation")
    ctx.initial_forces()
    ctx.synchronize()
    e0 = ctx.energy()[0]
    log(f"[bench] warmup {args.warmup} steps")
    ctx.step(MAX_DT, args.warmup, graph=True)
    ctx.synchronize()
    e_start = ctx.energy()[0]
    ctx.counters()

    # ---- timed region: graph replay ----
    K = args.steps
    log(f"[bench] timed {K} steps")
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(MAX_DT, K, graph=True)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    e_end = ctx.energy()[0]
    cnt = ctx.counters()
    ms_per_step = 1e3 * wall / K
    value = n * K / wall
    drift = (energy_total(e_end) - energy_total(e_start)) / max(abs(energy_total(e0)), 1.0)

    # ---- instrumented region: per-stage HIP events on the context stream ----
    log(f"[bench] instrumented {K} steps")
    ctx.enable_stage_timing(True)
    ctx.step(MAX_DT, K, graph=False)
    ctx.synchronize()
    stages = ctx.stage_times()
    ctx.enable_stage_timing(False)
    cnt2 = ctx.counters()
    grav_ms, grav_n = stages["gravity"]
    grav_avg_ms = grav_ms / max(grav_n, 1)
    inter_per_launch = cnt2["interactions"] / max(grav_n, 1)
    achieved = FLOP_PER_INTERACTION * inter_per_launch / (grav_avg_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"giant impact {n_t:,}+{n_i:,} particles, Fe/granite, 45 deg, 1.1 v_esc, "
                        "plain FP32 single-scenario KDK step (hipGraph)",
            "particles": n,
            "theta": 0.5,
            "softening": "h",
            "cfl": 0.3,
            "parallelism": "single GPU",
        },
        "energy_drift_rel": drift,
        "interactions_per_particle": cnt["interactions"] / (n * K),
        "neighbour_candidates_per_particle": cnt2["nbr_entries"] / n,
        "neighbour_candidates_max": cnt2["nbr_max"],
        "stage_ms": {k: (v[0] / max(v[1], 1)) for k, v in stages.items()},
        "roofline": {
            "bound": "valu",
            "kernel": "kGravity",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None,
            "launch_ms": grav_avg_ms,
            "flop_per_launch": FLOP_PER_INTERACTION * inter_per_launch,
        },
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(state, [n], params, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
