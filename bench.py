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
per-stage times and the gravity kernel's average launch duration used for the roofline.

Two rooflines are printed.  `roofline` is the dominant kernel's (gravity: FP32 VALU, 23 FLOP per
interaction).  `step_roofline` is the whole step's: achieved = sum over stages of
max(B_s / 8 TB/s, F_s / 157.3 TFLOP/s) / measured step time, with the algorithmic bytes B_s and
FLOPs F_s per particle of each stage (hydro step, 1,183 B + 120 B for the FP64 accumulators;
F = 64 n1 + 100 n2 + 40 n2 + 23 I + 2 x 60) and n1, n2, I measured in the same run.  The
neighbour-table build has no counterpart in the algorithm (B = F = 0): its time counts against
the step.  `traffic_ratio` gives, per kernel, HBM bytes read and written (rocprofv3 FETCH_SIZE +
WRITE_SIZE, calibrated) over its algorithmic bytes, from the committed counter summary.

The energy drift of the timed run is reported normalised by |E0| and by |Phi0| (at 2-3 v_esc,
K + Phi + U nearly cancels in some scenarios, so |E0| can be 1 % of |Phi0|).  The comparison
with the FP64 oracle over 1000 steps reports each drift twice: with the Barnes-Hut potential
(1/2 sum m phi_BH, what the step computes) and with the exact N^2 potential of the same
positions (the oracle's direct sum), which separates the potential's measurement error from
the integration's own energy error.

The CPU baseline runs the FP64 OpenMP oracle on the host: config 0 in full (2 x 8,192
particles, 100 adaptive KDK steps) and the 1,000,000-particle giant impact for 5 steps after
its initial force evaluation (the reported value), with the core count, OpenMP threads and CPU
model.

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
KERNEL_TRAFFIC_PROFILE = ROOT / "profiles" / "pmc_step_batched_sweep.md"
# Algorithmic bytes per particle-step of each timed stage (hydro step, FP32 state): kick 1 61,
# drift + XSPH 88, bounds + Morton 36, radix sort 200, permutation 88 + Karras/leaves 123, moment
# climb 123, pass 1 + EOS 126, gravity 68, pass 2 151, kick 2 + EOS 78 and the dt candidates 41
# (computed in kick 2; the dt stage only clamps per segment), 1,183 in all.  The FP64
# position/velocity accumulators add 120 (kick 1 24, drift 48, permutation 24, kick 2 24).
STEP_BYTES = {"dt": 0, "kick1": 61 + 24, "drift": 88 + 48, "morton": 36, "sort": 200,
              "permute_tree": 88 + 123 + 24, "moments": 123, "neighbours": 0, "pass1": 126, "gravity": 68,
              "pass2": 151, "kick2": 78 + 41 + 24}


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


def step_flops(stage, n1, n2, inter):
    """Algorithmic FLOPs per particle of a stage: 64 per pass-1 pair, 100 per pass-2 pair, 40 per
    XSPH pair, 23 per gravity interaction, 60 per EOS evaluation (after pass 1 and in kick 2)."""
    return {"drift": 40.0 * n2, "pass1": 64.0 * n1 + 60.0, "gravity": 23.0 * inter, "pass2": 100.0 * n2,
            "kick2": 60.0}.get(stage, 0.0)


def step_roofline(stages, n, n1, n2, inter, step_ms):
    """Sum over stages of max(B_s / HBM peak, F_s / FP32 peak) over the measured step time."""
    rows, t_min = {}, 0.0
    for k, b in STEP_BYTES.items():
        if k not in stages:
            continue
        f = step_flops(k, n1, n2, inter)
        t_b = n * b / (PEAK_HBM_GBS * 1e9) * 1e3
        t_f = n * f / (PEAK_FP32_TFLOPS * 1e12) * 1e3
        t = stages[k][0] / max(stages[k][1], 1)
        rows[k] = {"bytes_per_particle": b, "flop_per_particle": f, "t_min_ms": max(t_b, t_f),
                   "bound": "hbm" if t_b >= t_f else "fp32", "t_ms": t, "frac": max(t_b, t_f) / t if t > 0 else None}
        t_min += max(t_b, t_f)
    return {"definition": "sum_s max(B_s/8 TB/s, F_s/157.3 TFLOP/s) / step time", "achieved": t_min / step_ms,
            "t_min_ms": t_min, "step_ms": step_ms, "n1": n1, "n2": n2, "I": inter,
            "bytes_per_particle": sum(STEP_BYTES[k] for k in rows),
            "flop_per_particle": sum(r["flop_per_particle"] for r in rows.values()), "stages": rows}


def kernel_traffic_ratios():
    """Per kernel: counter bytes (FETCH_SIZE + WRITE_SIZE, calibrated) over algorithmic bytes per
    launch, from the committed rocprofv3 --pmc summary of this workload (rows
    `ratio: <kernel> <counter bytes> <algorithmic bytes>`), or None."""
    try:
        text = KERNEL_TRAFFIC_PROFILE.read_text()
    except OSError:
        return None
    out = {}
    for m in re.finditer(r"^ratio:\s*(\S+)\s+([0-9.eE+]+)\s+([0-9.eE+]+)", text, re.M):
        c, a = float(m.group(2)), float(m.group(3))
        out[m.group(1)] = {"counter_bytes": c, "algorithmic_bytes": a, "ratio": c / a if a > 0 else None}
    return out or None


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
        log("[bench] batched: autotune (12 steps)")
        tune = ctx.autotune(MAX_DT, steps_per_variant=1)
        log(f"[bench] batched: tuning {tune['chosen']}")
    ctx.step(MAX_DT, warmup, graph=True)
    ctx.synchronize()
    e_start = energy_total(ctx.energy())
    ctx.counters()
    log(f"[bench] batched: timed {steps} steps")
    wall = time_steps(ctx, steps)
    e_full = ctx.energy()
    e_end = energy_total(e_full)
    cnt = ctx.counters()
    drift = (e_end - e_start) / np.maximum(np.abs(e_init), 1.0)
    drift_phi = (e_end - e_start) / np.abs(e_full[:, 2])

    log(f"[bench] batched: instrumented {steps} steps")
    ctx.enable_stage_timing(True)
    ctx.step(MAX_DT, steps, graph=False)
    ctx.synchronize()
    stages = ctx.stage_times()
    ctx.enable_stage_timing(False)
    cnt2 = ctx.counters()
    return dict(wall=wall, drift=drift, drift_phi=drift_phi, cnt=cnt, cnt2=cnt2, stages=stages, n=n, tune=tune)


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


def run_config(A, state, counts, params, steps, warmup):
    """Throughput of another benchmark configuration in its own context: autotuned like the
    headline run, then `warmup` untimed and `steps` timed graph steps, then the same number of
    eager steps with stage events."""
    n = int(sum(counts))
    ctx = A.Context(capacity=n, max_segments=len(counts))
    try:
        ctx.set_params(params)
        ctx.set_layout(counts)
        ctx.load_state(state)
        ctx.initial_forces()
        tune = ctx.autotune(MAX_DT, steps_per_variant=1)
        ctx.step(MAX_DT, warmup, graph=True)
        ctx.counters()
        wall = time_steps(ctx, steps)
        cnt = ctx.counters()
        ctx.enable_stage_timing(True)
        ctx.step(MAX_DT, steps, graph=False)
        ctx.synchronize()
        stages = ctx.stage_times()
        return {"particles": n, "scenarios": len(counts), "particle_steps_per_s": n * steps / wall,
                "ms_per_step": 1e3 * wall / steps, "steps": steps, "warmup": warmup,
                "gravity_forced_accepts": int(cnt["forced_accepts"]),
                "stage_ms": {k: v[0] / max(v[1], 1) for k, v in stages.items()},
                "tuning": tune["chosen"]}
    finally:
        ctx.close()


def host_info():
    import os
    import oracle

    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "omp_max_threads": oracle.num_threads(), "cpu_model": model}


def cpu_baseline(A):
    """FP64 OpenMP oracle on the host: config 0 in full (2 x 8,192, 100 adaptive KDK steps) and the
    1,000,000-particle giant impact for 5 steps after its initial force evaluation.  Wall clock
    around the step loop only."""
    import oracle
    from astraea_amd import scenario as S

    params = S.plain_params(precision=A.PREC_FP64_ACCUM)
    info = host_info()
    log("[bench] CPU baseline: config 0, 2 x 8,192 particles, 100 steps (FP64 oracle)")
    st = S.config_small()
    n0 = len(st["mass"])
    o = oracle.Oracle(params, group_size=params.group_size or 64)
    o.set_layout([n0])
    o.load_state(st)
    o.initial_forces()
    e0 = o.energy()
    t0 = time.perf_counter()
    o.step(MAX_DT, 100)
    c0_s = time.perf_counter() - t0
    e1 = o.energy()
    tot0, tot1 = float(energy_total(e0)[0]), float(energy_total(e1)[0])
    c0 = {"particles": n0, "steps": 100, "seconds": c0_s, "particle_steps_per_s": n0 * 100 / c0_s,
          "energy_drift_rel_E0": (tot1 - tot0) / abs(tot0), "energy_drift_rel_Phi0": (tot1 - tot0) / abs(e0[0, 2])}
    del o
    log("[bench] CPU baseline: 1,000,000-particle giant impact, 5 steps (FP64 oracle)")
    st = S.config_giant()
    n1 = len(st["mass"])
    o = oracle.Oracle(params, group_size=params.group_size or 64)
    o.set_layout([n1])
    o.load_state(st)
    o.initial_forces()
    t0 = time.perf_counter()
    o.step(MAX_DT, 5)
    c1_s = time.perf_counter() - t0
    del o
    return {
        "value": n1 * 5 / c1_s,
        "unit": "particle-steps/s",
        "cores": info["omp_max_threads"],
        "kind": "port",
        "sample": f"FP64 OpenMP oracle, {info['omp_max_threads']} threads: the 1,000,000-particle giant impact, 5 full "
                  f"KDK steps after its initial force evaluation ({c1_s:.1f} s, {1e3 * c1_s / 5:.0f} ms per step); "
                  f"config 0 in full beside it",
        "config1": {"particles": n1, "steps": 5, "seconds": c1_s, "ms_per_step": 1e3 * c1_s / 5,
                    "particle_steps_per_s": n1 * 5 / c1_s},
        "config0": c0,
        "host": info,
    }


def exact_potential(oracle, state64, params, counts):
    """1/2 sum m phi of the direct N^2 sum (the oracle's gravity at theta = 0, Plummer softening
    soft_factor x h0 as in the step), per segment, for positions given in FP64."""
    q = oracle.default_params(theta=0.0, soft_factor=params.soft_factor)
    e = oracle.Oracle(q, 64)
    e.set_layout(counts)
    e.load_state({"pos_x": state64["pos_x"].astype(np.float32), "pos_y": state64["pos_y"].astype(np.float32),
                  "pos_z": state64["pos_z"].astype(np.float32), "mass": state64["mass"], "h": state64["h0"],
                  "h0": state64["h0"]})
    for k in ("x", "y", "z"):
        e.upload("pos64_" + k, state64["pos_" + k])
    e.sort_and_build()
    e.gravity()
    return e.energy()[:, 2]


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
    ge0 = ctx.energy()

    def gpu_state():
        # the segment's particles in original-id order, FP64 positions (the accumulators)
        ids = ctx.download("original_id")
        out = {}
        for k, src in (("pos_x", "pos64_x"), ("pos_y", "pos64_y"), ("pos_z", "pos64_z"), ("mass", "mass"),
                       ("h0", "h0")):
            v = ctx.download(src)
            w = np.empty_like(v)
            w[ids] = v
            out[k] = w
        return out

    px0 = exact_potential(oracle, gpu_state(), params, counts)
    ctx.step(MAX_DT, steps, graph=True)
    ctx.synchronize()
    g1 = energy_total(ctx.energy())
    ge1 = ctx.energy()
    px1 = exact_potential(oracle, gpu_state(), params, counts)
    o = oracle.Oracle(params, group_size=params.group_size or 64)
    o.set_layout(counts)
    o.load_state(state)
    o.initial_forces()
    r0 = energy_total(o.energy())
    re0 = o.energy()

    def oracle_state():
        ids = o.download("original_id")
        out = {}
        for k, src in (("pos_x", "pos64_x"), ("pos_y", "pos64_y"), ("pos_z", "pos64_z"), ("mass", "mass"),
                       ("h0", "h0")):
            v = o.download64(src) if src.startswith("pos64") else o.download(src)
            w = np.empty_like(v)
            w[ids] = v
            out[k] = w
        return out

    rx0 = exact_potential(oracle, oracle_state(), params, counts)
    t0 = time.perf_counter()
    o.step(MAX_DT, steps)
    t_oracle = time.perf_counter() - t0
    r1 = energy_total(o.energy())
    re1 = o.energy()
    rx1 = exact_potential(oracle, oracle_state(), params, counts)
    dg = (g1 - g0) / np.maximum(np.abs(g0), 1.0)
    dr = (r1 - r0) / np.maximum(np.abs(r0), 1.0)
    diff = np.abs(dg - dr)
    phi0 = np.abs(re0[:, 2])
    # drifts relative to |Phi0|, with the BH potential and with the exact one
    bh_g = (g1 - g0) / phi0
    bh_r = (r1 - r0) / phi0
    ex_g = ((g1 - ge1[:, 2] + px1) - (g0 - ge0[:, 2] + px0)) / phi0
    ex_r = ((r1 - re1[:, 2] + rx1) - (r0 - re0[:, 2] + rx0)) / phi0
    return {
        "steps": steps,
        "scenarios": len(counts),
        "particles": int(sum(counts)),
        "max_abs_vs_fp64": float(diff.max()),
        "median_abs_vs_fp64": float(np.median(diff)),
        "gpu_max_abs": float(np.abs(dg).max()),
        "fp64_max_abs": float(np.abs(dr).max()),
        "relative_to_abs_phi0": {
            "bh_potential": {"gpu_max_abs": float(np.abs(bh_g).max()), "fp64_max_abs": float(np.abs(bh_r).max()),
                             "max_abs_gpu_minus_fp64": float(np.abs(bh_g - bh_r).max())},
            "exact_potential": {"gpu_max_abs": float(np.abs(ex_g).max()), "fp64_max_abs": float(np.abs(ex_r).max()),
                                "max_abs_gpu_minus_fp64": float(np.abs(ex_g - ex_r).max())},
            "bh_potential_error_at_start_max": float(np.abs((re0[:, 2] - rx0) / phi0).max()),
        },
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
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="steps of the energy-drift comparison with the FP64 oracle (0: skip)")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the strength (4M) and gas-giant (16 x 500k) throughput runs")
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
    drift_phi = b["drift_phi"]
    cnt = b["cnt2"]
    steps_instr = K
    n1 = cnt["pass1_pairs"] / n
    n2 = cnt["pass2_pairs"] / n
    inter_pp = cnt["interactions"] / (n * steps_instr)

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
                             "steps": K, "normalised_by": "max(|E0|, 1)",
                             "by_abs_phi": {"max_abs": float(np.max(np.abs(drift_phi))),
                                            "median_abs": float(np.median(np.abs(drift_phi)))}},
        "interactions_per_particle": b["cnt"]["interactions"] / (n * K),
        "gravity_forced_accepts": int(b["cnt"]["forced_accepts"]),
        "neighbour_candidates_per_particle": b["cnt2"]["nbr_entries"] / n,
        "stage_ms": {k: (v[0] / max(v[1], 1)) for k, v in stages.items()},
        "hbm": {k: {"GB/s": n * ALG_BYTES[k] / (stages[k][0] / max(stages[k][1], 1) * 1e-3) / 1e9,
                    "frac": n * ALG_BYTES[k] / (stages[k][0] / max(stages[k][1], 1) * 1e-3) / 1e9 / PEAK_HBM_GBS,
                    "bytes_per_particle": ALG_BYTES[k]} for k in ALG_BYTES},
        "gravity_interactions_per_s": inter_per_launch / (grav_avg_ms * 1e-3),
        "tuning": b["tune"] and {"chosen": b["tune"]["chosen"], "steps": 12,
                                 "ms_per_step": {k: b["tune"][k] for k in ("pass1_ms", "pass2_ms", "drift_ms", "sort_ms")},
                                 "ms_per_launch": {k: b["tune"][k] for k in ("gravity_ms", "build_ms")}},
        "step_roofline": step_roofline(stages, n, n1, n2, inter_pp, 1e3 * b["wall"] / K),
        "traffic_ratio": kernel_traffic_ratios(),
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
    if not args.no_other_configs:
        # the other GPU configurations at their own sizes, reported beside the headline (not `value`)
        log("[bench] precision form: 4,000,000 particles with strength and damage")
        st4, p4 = S.config_strength()
        other = {"strength_4m": run_config(A, st4, [len(st4["mass"])], p4, K, W)}
        del st4
        other["strength_4m"]["workload"] = ("4,000,000-particle impact (Fe/granite/ice, strength and damage on for "
                                            "granite and ice), FP32 with FP64 leapfrog accumulators")
        log("[bench] gas giants: 16 scenarios of 500,000 particles")
        stg, cg = S.build_gas_giant_sweep(S.gas_giant_sweep(16, 500_000))
        other["gas_giants_16x500k"] = run_config(A, stg, cg, S.gas_giant_params(), K, W)
        del stg
        other["gas_giants_16x500k"]["workload"] = ("16 gas-giant impacts of 500,000 particles (H/He envelope, "
                                                   "cooling above 2,000 K), one batched step, FP64 accumulators")
        out["other_configs"] = other
    if args.drift_steps > 0:
        log(f"[bench] energy drift vs the FP64 oracle, {args.drift_steps} steps")
        out["energy_drift_vs_fp64_oracle"] = drift_vs_oracle(
            A, args.drift_steps, S.sweep_scenarios(8, n_min=1500, n_max=3000))
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(A)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
