"""Benchmark impact scenarios built on the host.

Bodies are differentiated spheres on a seed-jittered hexagonal-close-packed lattice
(jitter 0.1 spacing, h = 1.3 spacing) with uniform density per layer and u from the
closed-form hydrostatic pressure (astraea_build_body).  Two bodies are placed on
the reference's Keplerian approach orbit (engine/simulation.cpp:1389-1436) at
b = sin(angle) and v = ratio * mutual v_esc, 2 combined radii apart.

Seeds follow the benchmark plan: target seed s, impactor seed s + 1
(engine/simulation.cpp:1362), scenario k of a sweep uses splitmix64(42 + k).
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import PREC_FP32, PREC_FP64_ACCUM, build_body, default_params, place_two_bodies, splitmix64

R_EARTH_M = 6.371e6
IRON, GRANITE, WATER_ICE, HHE, REGOLITH = 0, 1, 2, 3, 4
FE_GRANITE = [(IRON, 0.3, 0.0), (GRANITE, 0.7, 0.0)]
FE_GRANITE_ICE = [(IRON, 0.3, 0.0), (GRANITE, 0.6, 0.0), (WATER_ICE, 0.1, 0.0)]
GAS_GIANT = [(IRON, 0.075, 0.0), (WATER_ICE, 0.075, 0.0), (HHE, 0.85, 0.0)]
R_GAS_GIANT_M = 7.0e7  # surface of the n = 1 polytrope envelope (its mass, ~0.47 M_Jupiter, follows)


def body_radius(mass_ratio: float, r_target_m: float) -> float:
    """Radius of a body with the same layering and mass_ratio times the target's mass."""
    return r_target_m * mass_ratio ** (1.0 / 3.0)


def impact_pair(n_target: int, n_impactor: int, angle_deg: float, v_ratio: float, seed: int = 42,
                layers=FE_GRANITE, r_target_m: float = R_EARTH_M, mass_ratio: float = 0.1, sep_radii: float = 2.0,
                flaws: bool = False) -> dict:
    """State dict (field name -> array) of a target + impactor approach, concatenated target first."""
    r_i = body_radius(mass_ratio, r_target_m)
    t = build_body(n_target, r_target_m, layers, seed=seed, flaws=flaws)
    i = build_body(n_impactor, r_i, layers, seed=seed + 1, flaws=flaws)
    n = n_target + n_impactor

    def cat(k):
        return np.ascontiguousarray(np.concatenate([t[k], i[k]]).astype(np.float32))

    st = {
        "pos_x": cat("x"), "pos_y": cat("y"), "pos_z": cat("z"),
        "vel_x": np.zeros(n, np.float32), "vel_y": np.zeros(n, np.float32), "vel_z": np.zeros(n, np.float32),
        "mass": cat("m"), "h": cat("h"), "h0": cat("h"), "rho": cat("rho"), "u": cat("u"),
        "eps_fail": cat("eps_fail"),
        "material": np.ascontiguousarray(np.concatenate([t["mat"], i["mat"]])),
        "body": np.concatenate([np.zeros(n_target, np.uint8), np.ones(n_impactor, np.uint8)]),
    }
    place_two_bodies(st, n_target, n_impactor, r_target_m / 1.0e6, r_i / 1.0e6, v_ratio,
                     math.sin(math.radians(angle_deg)), sep_radii)
    return st


def plain_params(**kw):
    """Benchmark physics: Plummer softening = h, CFL 0.3, strength and damage off, gravity
    opening decisions shared by a whole wavefront of 64 particles."""
    base = dict(soft_factor=1.0, cfl_factor=0.3, strength_mask=0, precision=PREC_FP32, group_size=64)
    base.update(kw)
    return default_params(**base)


def config_small(seed: int = 42) -> dict:
    """Two 8,192-particle Fe/granite bodies, 45 deg, 1.1 v_esc (the CPU-runnable case)."""
    return impact_pair(8192, 8192, 45.0, 1.1, seed=seed)


def config_strength(n_target: int = 3_600_000, n_impactor: int = 400_000, seed: int = 42, sep_radii: float = 2.0):
    """Precision form: Fe 30 % / granite 60 % / water ice 10 % bodies with Weibull flaws (m = 9,
    k = 4e29 per m^3), 30 deg, 1.5 v_esc.  Returns (state, params): strength and damage on for
    granite and ice (Y0 = 3.5 GPa for granite in the strength table), FP64 leapfrog accumulators."""
    st = impact_pair(n_target, n_impactor, 30.0, 1.5, seed=seed, layers=FE_GRANITE_ICE, sep_radii=sep_radii,
                     flaws=True)
    return st, plain_params(strength_mask=(1 << GRANITE) | (1 << WATER_ICE), precision=PREC_FP64_ACCUM)


def uniform_body_radius(mass_kg: float, layers) -> float:
    """Radius of a body of uniform-density layers with the given mass: such a body's mass scales
    as R^3, so one small trial build fixes it."""
    r0 = 1.0e6
    m0 = build_body(64, r0, layers)["mass_kg"]
    return r0 * (mass_kg / m0) ** (1.0 / 3.0)


def gas_giant_pair(n_target: int, n_impactor: int, angle_deg: float, v_ratio: float, seed: int = 42,
                   mass_ratio: float = 0.05, sep_radii: float = 2.0) -> dict:
    """Gas-giant form: a target with an H/He envelope (85 % of the mass, n = 1 polytrope) around an
    iron + water-ice core, hit by an iron/granite impactor of `mass_ratio` times its mass."""
    t = build_body(n_target, R_GAS_GIANT_M, GAS_GIANT, seed=seed)
    r_i = uniform_body_radius(mass_ratio * t["mass_kg"], FE_GRANITE)
    i = build_body(n_impactor, r_i, FE_GRANITE, seed=seed + 1)
    n = n_target + n_impactor

    def cat(k):
        return np.ascontiguousarray(np.concatenate([t[k], i[k]]).astype(np.float32))

    st = {
        "pos_x": cat("x"), "pos_y": cat("y"), "pos_z": cat("z"),
        "vel_x": np.zeros(n, np.float32), "vel_y": np.zeros(n, np.float32), "vel_z": np.zeros(n, np.float32),
        "mass": cat("m"), "h": cat("h"), "h0": cat("h"), "rho": cat("rho"), "u": cat("u"),
        "eps_fail": cat("eps_fail"),
        "material": np.ascontiguousarray(np.concatenate([t["mat"], i["mat"]])),
        "body": np.concatenate([np.zeros(n_target, np.uint8), np.ones(n_impactor, np.uint8)]),
    }
    place_two_bodies(st, n_target, n_impactor, R_GAS_GIANT_M / 1.0e6, r_i / 1.0e6, v_ratio,
                     math.sin(math.radians(angle_deg)), sep_radii)
    return st


def gas_giant_params(**kw):
    """Gas-giant physics: H/He gamma = 5/3, surface cooling only above 2,000 K, FP64 accumulators."""
    base = dict(gamma_hhe=5.0 / 3.0, cooling_min_temp_k=2000.0, precision=PREC_FP64_ACCUM)
    base.update(kw)
    return plain_params(**base)


def gas_giant_sweep(count: int = 16, n_per: int = 500_000, master_seed: int = 42):
    """Batched gas-giant cases: per scenario angle U[0, 60] deg, v U[1, 2] v_esc, impactor 0.05 of
    the target's mass with n/20 particles, seeds splitmix64(master + k)."""
    out = []
    for k in range(count):
        s = splitmix64(master_seed + k)
        rng = np.random.default_rng(s)
        n_i = max(16, n_per // 20)
        out.append(dict(n_target=n_per - n_i, n_impactor=n_i, angle_deg=rng.uniform(0.0, 60.0),
                        v_ratio=rng.uniform(1.0, 2.0), seed=int(s % (2**63))))
    return out


def build_gas_giant_sweep(specs, threads: int | None = None):
    """Concatenated state of the gas-giant scenarios `specs` (gas_giant_sweep) and their particle
    counts, built on `threads` host threads."""
    from concurrent.futures import ThreadPoolExecutor

    def one(sp):
        return gas_giant_pair(sp["n_target"], sp["n_impactor"], sp["angle_deg"], sp["v_ratio"], seed=sp["seed"])

    threads = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=threads) as pool:
        states = list(pool.map(one, specs))
    counts = [sp["n_target"] + sp["n_impactor"] for sp in specs]
    merged = {k: np.ascontiguousarray(np.concatenate([s[k] for s in states])) for k in states[0]}
    return merged, counts


def config_giant(n_target: int = 900_000, n_impactor: int = 100_000, seed: int = 42) -> dict:
    """Giant impact, 900k + 100k particles, 45 deg, 1.1 v_esc."""
    return impact_pair(n_target, n_impactor, 45.0, 1.1, seed=seed)


def sweep_scenarios(count: int = 64, n_min: int = 20_000, n_max: int = 500_000, master_seed: int = 42,
                    scale: float = 1.0):
    """Batched sweep: per scenario N log-uniform in [n_min, n_max] (times `scale`), angle U[0,75],
    v U[1,3] v_esc, mass ratio U[0.01,0.2], seeds splitmix64(master + k)."""
    out = []
    for k in range(count):
        s = splitmix64(master_seed + k)
        rng = np.random.default_rng(s)
        n = int(math.exp(rng.uniform(math.log(n_min), math.log(n_max))) * scale)
        ratio = rng.uniform(0.01, 0.2)
        angle = rng.uniform(0.0, 75.0)
        v = rng.uniform(1.0, 3.0)
        n_i = max(16, int(round(n * ratio / (1.0 + ratio))))
        n_t = max(16, n - n_i)
        out.append(dict(n_target=n_t, n_impactor=n_i, angle_deg=angle, v_ratio=v, mass_ratio=ratio,
                        seed=int(s % (2**63))))
    return out


def build_sweep(specs, threads: int | None = None, split: bool = False):
    """Concatenated state of the scenarios `specs` and their particle counts (segment order).
    Bodies are built on `threads` host threads (the builder releases the GIL).  With `split`,
    the per-scenario states are returned instead of the concatenation."""
    from concurrent.futures import ThreadPoolExecutor

    def one(sp):
        return impact_pair(sp["n_target"], sp["n_impactor"], sp["angle_deg"], sp["v_ratio"], seed=sp["seed"],
                           mass_ratio=sp["mass_ratio"])

    threads = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=threads) as pool:
        states = list(pool.map(one, specs))
    counts = [sp["n_target"] + sp["n_impactor"] for sp in specs]
    if split:
        return states, counts
    merged = {k: np.ascontiguousarray(np.concatenate([s[k] for s in states])) for k in states[0]}
    return merged, counts
