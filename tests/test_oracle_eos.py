"""Pins the oracle's EOS to the closed forms of the reference's tests/cpp/test_eos.cpp.

Each check restates one reference assertion (file:line) against analytic formulas written
independently here, so a dropped term or wrong branch in the oracle fails.
"""
import math

import numpy as np
import pytest

U_UNIT = 1.0e6   # J/kg per PSU (engine/units.hpp:22)
RHO_UNIT = 1.0e6  # kg/m^3 per PSU
P_UNIT = 1.0e12  # Pa per PSU
MATS = range(5)
HHE = 3


def cold_si(m, rho):
    """tests/cpp/test_eos.cpp:26-38"""
    if m["kH2"] > 0:
        return m["kH2"] * rho
    eta = rho / m["rho0"]
    if eta <= 1.0:
        return 0.0
    ln = math.log(eta)
    return (m["A"] * (ln + 1 / eta - 1) + m["B"] * (eta - 1 / eta - 2 * ln)) / m["rho0"]


def pc_si(m, rho, u):
    """tests/cpp/test_eos.cpp:40-46"""
    eta = rho / m["rho0"]
    mu = eta - 1
    b = m["B"] if mu >= 0 else 0.0
    return (m["a"] + m["b"] / (u / (m["E0"] * eta * eta) + 1)) * rho * u + m["A"] * mu + b * mu * mu


def pe_si(m, rho, u):
    """tests/cpp/test_eos.cpp:48-56"""
    eta = rho / m["rho0"]
    mu = eta - 1
    x = m["rho0"] / rho - 1
    den = u / (m["E0"] * eta * eta) + 1
    return m["a"] * rho * u + (m["b"] * rho * u / den + m["A"] * mu * math.exp(-m["beta"] * x)) * math.exp(
        -m["alpha"] * x * x)


def test_parameter_table(orc):
    """tests/cpp/test_eos.cpp:76-162 (published constants)."""
    fe, gr, ice = orc.material(0), orc.material(1), orc.material(2)
    assert (fe["rho0"], fe["a"], fe["b"], fe["A"], fe["B"]) == (7800.0, 0.5, 1.5, 1.28e11, 1.05e11)
    assert (fe["E0"], fe["Eiv"], fe["Ecv"], fe["alpha"], fe["beta"]) == (9.5e6, 2.4e6, 8.67e6, 5.0, 5.0)
    assert (fe["cv"], fe["tMelt"], fe["tRef"], fe["mu"], fe["Y0"], fe["epsFailMin"]) == (449.0, 1811.0, 293.0, 7.7e10, 6.8e8, 1e-3)
    assert (gr["rho0"], gr["a"], gr["b"], gr["A"], gr["B"]) == (2680.0, 0.5, 1.3, 1.8e10, 1.8e10)
    assert (gr["E0"], gr["Eiv"], gr["Ecv"], gr["cv"], gr["tMelt"], gr["Y0"]) == (1.6e7, 3.5e6, 1.8e7, 790.0, 1600.0, 3.5e9)
    assert (ice["rho0"], ice["a"], ice["b"], ice["A"], ice["epsFailMin"]) == (1000.0, 0.7, 0.15, 2.18e9, 5e-4)


@pytest.mark.parametrize("mat", MATS)
def test_cold_energy_exact_and_monotone(orc, mat):
    """tests/cpp/test_eos.cpp:164-192: u_cold vs the analytic integral, 1e-10 relative."""
    m = orc.material(mat)
    if mat == HHE:
        prev = -1.0
        for rho in (0.178, 10.0, 100.0, 1000.0, 5000.0):
            v = orc.cold_energy_psu(mat, rho / RHO_UNIT)
            e = cold_si(m, rho) / U_UNIT
            assert v > prev and abs(v - e) <= 1e-10 * e + 1e-12
            prev = v
        return
    for eta in (0.5, 0.9, 1.0):
        assert orc.cold_energy_psu(mat, eta * m["rho0"] / RHO_UNIT) == 0.0
    prev = 0.0
    for eta in (1.05, 1.2, 1.5, 2.0, 3.0):
        v = orc.cold_energy_psu(mat, eta * m["rho0"] / RHO_UNIT)
        e = cold_si(m, eta * m["rho0"]) / U_UNIT
        assert v > prev and abs(v - e) <= 1e-10 * e + 1e-12
        prev = v


def cold_si_exact(m, rho):
    """cold_si in 60-digit decimal arithmetic (the log form cancels to ~mu^2/2 near rho0)."""
    from decimal import Decimal, localcontext

    with localcontext() as ctx:
        ctx.prec = 60
        A, B, r0 = Decimal(m["A"]), Decimal(m["B"]), Decimal(m["rho0"])
        eta = Decimal(rho) / r0
        ln = eta.ln()
        return float((A * (ln + 1 / eta - 1) + B * (eta - 1 / eta - 2 * ln)) / r0)


def test_cold_energy_series_branch_continuous(orc):
    """The mu < 1e-5 series (eos.cu:304-310) joins the log form: both sides of the switch agree
    with the integral evaluated in 60-digit arithmetic."""
    for mat in (0, 1, 2, 4):
        m = orc.material(mat)
        r0 = m["rho0"] / RHO_UNIT
        for f in (0.999e-5, 1.001e-5):
            v = orc.cold_energy_psu(mat, r0 * (1 + f))
            exact = cold_si_exact(m, m["rho0"] * (1 + f)) / U_UNIT
            # the series is exact to O(mu^6); the log form loses ~1e-16 / mu^2 to cancellation
            tol = 1e-9 if f < 1e-5 else 1e-5
            assert abs(v - exact) <= tol * exact, (mat, f, v, exact)
        assert orc.cold_energy_psu(mat, r0 * (1 + 1.001e-5)) > orc.cold_energy_psu(mat, r0 * (1 + 0.999e-5))


@pytest.mark.parametrize("mat", [0, 1, 2, 4])
def test_regimes_and_damaged_tension(orc, mat):
    """tests/cpp/test_eos.cpp:194-298: compressed P_c, two-phase blend, expanded P_e, damage."""
    m = orc.material(mat)
    r0 = m["rho0"] / RHO_UNIT
    # compressed, eta = 1.5, T = T_ref + 1000
    rho = np.float32(1.5 * r0)
    rho_si = float(rho) * RHO_UNIT
    ucold = orc.cold_energy_psu(mat, float(rho))
    ueos_si = ucold * U_UNIT + m["cv"] * 1000.0
    u = np.float32((ucold * U_UNIT + m["cv"] * (m["tRef"] + 1000.0)) / U_UNIT)
    p, cs, t, _ = orc.eos_eval(mat, float(rho), float(u), 0.0)
    assert p > 0 and cs > 0 and t >= 3.0
    assert abs(p * P_UNIT - pc_si(m, rho_si, ueos_si)) <= 1e-4 * abs(pc_si(m, rho_si, ueos_si)) + 1e3
    pd = orc.eos_eval(mat, float(rho), float(u), 1.0)[0]
    assert pd > 0 and abs(pd - p) <= 1e-5 * abs(p) + 1e-12
    # two-phase, eta = 0.7, u_EOS midway between E_iv and E_cv
    rho = np.float32(0.7 * r0)
    rho_si = float(rho) * RHO_UNIT
    ue = 0.5 * (m["Eiv"] + m["Ecv"])
    u = np.float32((m["cv"] * m["tRef"] + ue) / U_UNIT)
    p = orc.eos_eval(mat, float(rho), float(u), 0.0)[0]
    exp = ((ue - m["Eiv"]) * pe_si(m, rho_si, ue) + (m["Ecv"] - ue) * pc_si(m, rho_si, ue)) / (m["Ecv"] - m["Eiv"])
    assert abs(p * P_UNIT - exp) <= 1e-4 * abs(exp) + 1e3
    # expanded, eta = 0.2, u_EOS = 2 E_cv
    rho = np.float32(0.2 * r0)
    rho_si = float(rho) * RHO_UNIT
    ue = 2 * m["Ecv"]
    u = np.float32((m["cv"] * m["tRef"] + ue) / U_UNIT)
    p, cs, t, _ = orc.eos_eval(mat, float(rho), float(u), 0.0)
    assert p > 0 and cs > 0
    assert abs(p * P_UNIT - pe_si(m, rho_si, ue)) <= 1e-4 * abs(pe_si(m, rho_si, ue)) + 1e3
    # damaged tension, eta = 0.9, T = T_ref: intact P < 0, D = 1 -> 0, D = 0.5 -> half
    rho = np.float32(0.9 * r0)
    u = np.float32(m["cv"] * m["tRef"] / U_UNIT)
    pi = orc.eos_eval(mat, float(rho), float(u), 0.0)[0]
    assert pi < 0
    assert abs(orc.eos_eval(mat, float(rho), float(u), 1.0)[0]) <= 1e-12
    ph = orc.eos_eval(mat, float(rho), float(u), 0.5)[0]
    assert abs(ph - 0.5 * pi) <= 1e-5 * abs(pi) + 1e-12


def test_hhe_identity(orc):
    """tests/cpp/test_eos.cpp:300-339: P = K rho^2 + (gamma-1) rho u_th, c_s^2 = 2 K rho + gamma(gamma-1) u_th."""
    m = orc.material(HHE)
    for rho_si, uth in ((0.178, 2.0e6), (50.0, 0.0), (200.0, 5.0e6), (1200.0, 2.5e7), (3500.0, 6.0e7)):
        rho = np.float32(rho_si / RHO_UNIT)
        u = np.float32((m["kH2"] * rho_si + uth) / U_UNIT)
        p, cs, _, _ = orc.eos_eval(HHE, float(rho), float(u), 0.0, 1.4)
        exp = m["kH2"] * rho_si ** 2 + 0.4 * rho_si * uth
        assert abs(p * P_UNIT - exp) <= 1e-4 * exp + 1.0
        ecs = max(0.05, math.sqrt(2 * m["kH2"] * rho_si + 1.4 * 0.4 * uth) / 1.0e3)
        assert abs(cs - ecs) <= 1e-4 * ecs + 1e-6 and cs >= 0.05
    lo = orc.eos_eval(HHE, float(np.float32(0.001 / RHO_UNIT)), float(np.float32(m["kH2"] * 0.001 / U_UNIT)), 0.0)
    assert lo[1] >= 0.05 - 1e-6


def test_hhe_gamma_knob(orc):
    """The configurable H/He gamma enters P and c_s exactly as the fixed 1.4 does."""
    m = orc.material(HHE)
    rho_si, uth = 300.0, 1.0e7
    rho = np.float32(rho_si / RHO_UNIT)
    u = np.float32((m["kH2"] * rho_si + uth) / U_UNIT)
    g = 5.0 / 3.0
    p, cs, _, _ = orc.eos_eval(HHE, float(rho), float(u), 0.0, g)
    exp = m["kH2"] * rho_si ** 2 + (g - 1) * rho_si * uth
    assert abs(p * P_UNIT - exp) <= 1e-4 * exp


def test_water_zero_pressure_and_stp(orc):
    """tests/cpp/test_eos.cpp:341-366."""
    rho = orc.invert_density_si(2, 0.0, 288.0, 1000.0)
    assert 940.0 <= rho <= 1005.0
    u, p = orc.eos_from_rho_temp_si(2, rho, 288.0)
    assert abs(p) <= 1.0e4
    _, _, t, ph = orc.eos_eval(2, float(np.float32(rho / RHO_UNIT)), float(np.float32(u / U_UNIT)), 0.0)
    assert ph == 4 and abs(t - 288.0) <= 0.5
    for mat in (0, 1, 2, 4):
        m = orc.material(mat)
        _, p = orc.eos_from_rho_temp_si(mat, m["rho0"], m["tRef"])
        assert abs(p) <= 1.0


def test_phase_branches(orc):
    """tests/cpp/test_eos.cpp:368-428."""
    g = orc.material(1)
    r0 = g["rho0"] / RHO_UNIT
    f = lambda v: float(np.float32(v))
    assert orc.eos_eval(1, f(r0), f(g["cv"] * 300 / U_UNIT), 0.0)[3] == 0
    assert orc.eos_eval(1, f(r0), f(g["cv"] * 2200 / U_UNIT), 0.0)[3] == 1
    he = orc.eos_eval(1, f(0.5 * r0), f((g["cv"] * g["tRef"] + 1.2 * g["Ecv"]) / U_UNIT), 0.0)
    assert he[2] > 12000 and he[3] in (2, 3)
    assert orc.eos_eval(1, f(r0), f(g["cv"] * 15000 / U_UNIT), 0.0)[3] == 3
    w = orc.material(2)
    r0 = w["rho0"] / RHO_UNIT
    assert orc.eos_eval(2, f(r0), f(w["cv"] * 200 / U_UNIT), 0.0)[3] == 0
    assert orc.eos_eval(2, f(r0), f(w["cv"] * 300 / U_UNIT), 0.0)[3] == 4
    assert orc.eos_eval(2, f(0.1 * r0), f((w["cv"] * 1000 + w["Ecv"]) / U_UNIT), 0.0)[3] == 2
    assert orc.eos_eval(2, f(r0), f(w["cv"] * 18000 / U_UNIT), 0.0)[3] == 3
    h = orc.material(HHE)
    rho = f(100.0 / RHO_UNIT)
    uc = orc.cold_energy_psu(HHE, rho)
    assert orc.eos_eval(HHE, rho, f(uc + h["cv"] * 500 / U_UNIT), 0.0)[3] == 5
    assert orc.eos_eval(HHE, rho, f(uc + h["cv"] * 12000 / U_UNIT), 0.0)[3] == 3


def test_temperature_floor(orc):
    """tests/cpp/test_eos.cpp:430-437: T >= 3 K at zero internal energy."""
    for mat in MATS:
        m = orc.material(mat)
        assert orc.eos_eval(mat, float(np.float32(1.4 * m["rho0"] / RHO_UNIT)), 0.0, 0.0)[2] >= 3.0


def test_sound_speed_closed_form_compressed(orc):
    """c_s^2 = dP/drho + P/rho^2 dP/du with the analytic Tillotson derivatives (compressed branch),
    checked against central finite differences of the closed-form P_c."""
    for mat in (0, 1, 2, 4):
        m = orc.material(mat)
        rho_si = 1.3 * m["rho0"]
        ue = m["E0"] * 0.2
        u_total = (cold_si(m, rho_si) + m["cv"] * m["tRef"] + ue) / U_UNIT
        rho = np.float32(rho_si / RHO_UNIT)
        u = np.float32(u_total)
        p, cs, _, _ = orc.eos_eval(mat, float(rho), float(u), 0.0)
        rho_si = float(rho) * RHO_UNIT
        uE = float(u) * U_UNIT - m["cv"] * m["tRef"]  # u_EOS = u_cold + excess
        hr, hu = rho_si * 1e-6, uE * 1e-6
        # P(rho, u_EOS) where u_EOS = u_cold(rho) + (u - u_cold(rho) - cv T_ref): u_EOS = u - cv T_ref at fixed u
        dpdr = (pc_si(m, rho_si + hr, uE) - pc_si(m, rho_si - hr, uE)) / (2 * hr)
        dpdu = (pc_si(m, rho_si, uE + hu) - pc_si(m, rho_si, uE - hu)) / (2 * hu)
        ps = pc_si(m, rho_si, uE)
        cs_exp = math.sqrt(dpdr + ps / rho_si ** 2 * dpdu) / 1.0e3
        floor = max(0.05, math.sqrt(m["A"] / m["rho0"]) / 1.0e3)
        assert abs(cs - max(floor, cs_exp)) <= 1e-3 * cs_exp
