"""Hydrostatic planet builder (astraea_build_planet / astraea_planet_profile), checked against
the reference's acceptance properties (tests/cpp/test_planet_builder.cpp:138-260, AC2) and
independently of its own code:

* the profile integrates back: P(r) from the surface inward by the trapezoid rule on
  -G m rho / r^2, and m(r) from the centre by the trapezoid rule on 4 pi r^2 rho, agree with the
  stored pressure and mass to 2e-3 of the central pressure and the total mass;
* every sample's (rho, T) gives its pressure back through the oracle's EOS
  (or_eos_from_rho_temp_si, pinned on the reference's EOS closed forms), to 1e-6 of P_c;
* the particles' u is the oracle's u(rho, T), h0 = 1.2 (m / rho)^(1/3), and the masses add up.
"""
import numpy as np
import pytest

import oracle
from astraea_amd import build_planet, planet_profile, planet_spec

G = 6.67430e-11
ARCH = {  # test_planet_builder.cpp:24-66 with the compositions of presets.cpp:9-88
    "LivingTerra": (5.97219e24, [(0, .32), (1, .675), (2, .005)], 288.0, 5500.0, 0.0),
    "IronCore": (3.3011e23, [(0, .68), (1, .32)], 440.0, 3200.0, 0.0),
    "OceanIce": (4.0e24, [(0, .15), (1, .45), (2, .40)], 275.0, 3800.0, 0.0),
    "LavaIo": (2.0e24, [(0, .25), (1, .75)], 1550.0, 4200.0, 0.0),
    "GasGiant": (1.2e26, [(0, .05), (1, .10), (3, .85)], 165.0, 18000.0, 1.0e5),
    "ProtoTheia": (6.417e23, [(0, .22), (1, .58), (4, .20)], 240.0, 2200.0, 0.0),
}


def _spec(name, seed=42):
    m, lay, ts, tc, ps = ARCH[name]
    return planet_spec(m, lay, ts, tc, ps, seed=seed)


@pytest.mark.parametrize("name", list(ARCH))
def test_profile_is_hydrostatic(name):
    m0, lay, ts, tc, ps = ARCH[name]
    pr = planet_profile(_spec(name))
    r, rho, p, m, t, mat = (pr[k] for k in ("r", "rho", "p", "m", "t", "mat"))
    pc, R = pr["central_pressure_pa"], pr["radius_m"]
    # AC2 properties
    assert len(r) >= 16 and R > 1e5 and pc > 1e7
    assert abs(m[-1] - m0) <= 1e-5 * m0 and abs(r[-1] - R) <= 1e-9 * R
    assert np.all(np.diff(r) >= 0) and np.all(np.diff(m) >= 0)
    assert np.all(np.diff(p) <= 1e-4 * pc)
    same = mat[1:] == mat[:-1]
    assert np.all(rho[1:][same] <= rho[:-1][same] * (1 + 1e-5))
    assert set(np.unique(mat)) == {k for k, _ in lay}
    assert p[-1] == pytest.approx(ps, abs=1e-6 * pc)
    # integrate the profile back
    dr = np.diff(r)
    g = np.where(r > 0, G * m * rho / np.maximum(r, 1e-30) ** 2, 0.0)
    p_back = ps + np.concatenate([np.cumsum((0.5 * (g[1:] + g[:-1]) * dr)[::-1])[::-1], [0.0]])
    assert np.max(np.abs(p_back - p)) < 2e-3 * pc
    dm = 4 * np.pi * r ** 2 * rho
    m_back = np.concatenate([[0.0], np.cumsum(0.5 * (dm[1:] + dm[:-1]) * dr)])
    assert np.max(np.abs(m_back - m)) < 2e-3 * m0
    # geotherm linear in pressure between surface and centre temperatures
    np.testing.assert_allclose(t, ts + (tc - ts) * np.clip((p - ps) / (pc - ps), 0, 1), rtol=1e-12, atol=1e-9)
    # each sample is on the EOS: P(rho, T) of the oracle
    p_eos = np.array([oracle.eos_from_rho_temp_si(int(mat[i]), rho[i], t[i])[1] for i in range(len(r))])
    inside = p > ps  # at the surface the density is the EOS's own root
    assert np.max(np.abs(p_eos[inside] - p[inside])) < 1e-6 * pc


@pytest.mark.parametrize("name", list(ARCH))
def test_planet_particles(name):
    m0, lay, ts, tc, ps = ARCH[name]
    n = 6000
    b = build_planet(_spec(name), n)
    for k in ("x", "y", "z", "m", "h", "rho", "u", "temp_k", "eps_fail"):
        assert np.all(np.isfinite(b[k])), k
    assert np.all(b["m"] > 0) and np.all(b["h"] > 0) and np.all(b["rho"] > 0) and np.all(b["temp_k"] > 0)
    assert b["m"].astype(np.float64).sum() * 1e24 == pytest.approx(m0, rel=1e-5)
    # particles per layer follow the mass fractions
    counts = [int((b["mat"] == k).sum()) for k, _ in lay]
    assert sum(counts) == n
    for (k, f), c in zip(lay, counts):
        assert abs(c - f * n) <= 1.0, (k, c, f * n)
    # u(rho, T) and h0 = 1.2 (m / rho)^(1/3)
    rho_si = b["rho"].astype(np.float64) * 1e6
    u_eos = np.array([oracle.eos_from_rho_temp_si(int(b["mat"][i]), rho_si[i], float(b["temp_k"][i]))[0]
                      for i in range(0, n, 7)])
    np.testing.assert_allclose(b["u"][::7] * 1e6, u_eos, rtol=2e-6)
    np.testing.assert_allclose(b["h"], 1.2 * np.cbrt(b["m"] / b["rho"]), rtol=1e-5)
    # denser inside than outside within each material (AC2)
    r = np.sqrt(b["x"].astype(np.float64) ** 2 + b["y"] ** 2 + b["z"] ** 2)
    for k, _ in lay:
        sel = np.flatnonzero(b["mat"] == k)
        if len(sel) < 20:
            continue
        o = sel[np.argsort(r[sel])]
        q = len(o) // 4
        assert b["rho"][o[:q]].mean() >= 0.99 * b["rho"][o[-q:]].mean()
    assert r.max() <= b["radius_m"] / 1e6 * (1 + 1e-6)


def test_planet_is_deterministic_and_seeded():
    a = build_planet(_spec("LivingTerra"), 4000)
    b = build_planet(_spec("LivingTerra"), 4000)
    c = build_planet(_spec("LivingTerra", seed=7), 4000)
    for k in ("x", "y", "z", "rho", "u", "eps_fail"):
        assert np.array_equal(a[k], b[k]), k
    assert not np.array_equal(a["x"], c["x"])
    assert np.array_equal(a["rho"], c["rho"])  # the profile does not depend on the seed
