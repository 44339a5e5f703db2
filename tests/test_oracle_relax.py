"""Oracle damped relaxation (engine/simulation.cpp:585-781) on closed forms.

* steps = 0: only the initial forces run, and the result is v_rms / v_esc of the loaded state,
  v_esc = sqrt(2 G M / max(nominal radius, largest |x|)).
* Force-free particles (far apart, negligible masses): each relaxation step removes the mean
  velocity once and multiplies what is left by 0.85; the four closing steps are undamped, so
  v_rms ends at 0.85^steps times the initial v_rms about the mean.
* After the final re-anchor u = u_cold(rho) + c_v T_anchor exactly, and stress and damage are 0.
* A real body keeps its centre of mass at the origin and no net momentum.
"""
import numpy as np
import pytest

import oracle
from astraea_amd import build_body, default_params

G = 66.7430
GRANITE = 1


def _state_from_body(b):
    n = len(b["m"])
    return {"pos_x": b["x"], "pos_y": b["y"], "pos_z": b["z"], "vel_x": np.zeros(n, np.float32),
            "vel_y": np.zeros(n, np.float32), "vel_z": np.zeros(n, np.float32), "mass": b["m"], "h": b["h"],
            "h0": b["h"], "rho": b["rho"], "u": b["u"], "eps_fail": b["eps_fail"], "material": b["mat"],
            "body": np.zeros(n, np.uint8)}


def free_particles(seed=3):
    """Eight granite particles 200 PSU apart with masses of 1e-12 PSU and random velocities."""
    rng = np.random.default_rng(seed)
    n = 8
    pos = (np.indices((2, 2, 2)).reshape(3, -1).T * 200.0 - 100.0).astype(np.float32)
    vel = rng.normal(0.0, 1.0, (n, 3)).astype(np.float32)
    f = lambda v: np.full(n, v, np.float32)
    return {"pos_x": pos[:, 0].copy(), "pos_y": pos[:, 1].copy(), "pos_z": pos[:, 2].copy(),
            "vel_x": vel[:, 0].copy(), "vel_y": vel[:, 1].copy(), "vel_z": vel[:, 2].copy(), "mass": f(1e-12),
            "h": f(0.05), "h0": f(0.05), "rho": f(2.68e-3), "u": f(1.0), "material": np.full(n, GRANITE, np.uint8),
            "body": np.zeros(n, np.uint8)}


def _oracle(st, counts=None):
    n = len(st["mass"])
    o = oracle.Oracle(default_params(), group_size=64)
    o.set_layout(counts or [n])
    o.load_state(st)
    return o


def test_relax_zero_steps_is_vrms_over_vesc():
    b = build_body(600, 2.0e6, [(GRANITE, 1.0, 0.0)], seed=5)
    st = _state_from_body(b)
    rng = np.random.default_rng(1)
    for k in ("vel_x", "vel_y", "vel_z"):
        st[k] = rng.normal(0.0, 0.3, len(b["m"])).astype(np.float32)
    o = _oracle(st)
    m = st["mass"].astype(np.float64)
    v2 = sum(st[k].astype(np.float64) ** 2 for k in ("vel_x", "vel_y", "vel_z"))
    r = np.sqrt(sum(st[k].astype(np.float64) ** 2 for k in ("pos_x", "pos_y", "pos_z")))
    for nominal in (1.0, 10.0):
        expect = np.sqrt((m * v2).sum() / m.sum()) / np.sqrt(2 * G * m.sum() / max(nominal, r.max()))
        assert o.relax(0, nominal)[0] == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("steps", [1, 3])
def test_relax_damps_force_free_particles(steps):
    st = free_particles()
    o = _oracle(st)
    m = st["mass"].astype(np.float64)
    v = np.stack([st[k] for k in ("vel_x", "vel_y", "vel_z")], 1).astype(np.float64)
    dv = v - (m[:, None] * v).sum(0) / m.sum()
    nominal = 1.0e6  # larger than every |x|, so v_esc is fixed
    vesc = np.sqrt(2 * G * m.sum() / nominal)
    expect = 0.85 ** steps * np.sqrt((m * (dv ** 2).sum(1)).sum() / m.sum()) / vesc
    assert o.relax(steps, nominal)[0] == pytest.approx(expect, rel=1e-9)


def test_relax_reanchors_and_keeps_body_centred():
    b = build_body(800, 2.0e6, [(0, 0.3, 0.0), (GRANITE, 0.7, 0.0)], seed=9)
    st = _state_from_body(b)
    n = len(b["m"])
    anchor = 300.0 + 10.0 * (np.arange(n) % 7)
    o = _oracle(st)
    o.relax(2, 2.0, anchor_temp=anchor)
    oid = o.download("original_id")
    rho, u, mat = o.download64("rho"), o.download64("u"), o.download("material")
    cv = {k: oracle.material(k)["cv"] for k in (0, GRANITE)}
    want = np.array([oracle.cold_energy_psu(int(mat[i]), rho[i]) + cv[int(mat[i])] / 1e6 * anchor[oid[i]]
                     for i in range(n)])
    np.testing.assert_allclose(u, want, rtol=1e-12)
    for k in ("sxx", "syy", "szz", "sxy", "sxz", "syz", "damage", "damage_cbrt"):
        assert not np.any(o.download64(k)), k
    m = o.download64("mass")
    x = np.stack([o.download64(k) for k in ("pos_x", "pos_y", "pos_z")], 1)
    v = np.stack([o.download64(k) for k in ("vel_x", "vel_y", "vel_z")], 1)
    vesc = np.sqrt(2 * G * m.sum() / 2.0)
    assert np.linalg.norm((m[:, None] * x).sum(0)) / m.sum() < 1e-4 * 2.0
    assert np.linalg.norm((m[:, None] * v).sum(0)) / (m.sum() * vesc) < 1e-4
