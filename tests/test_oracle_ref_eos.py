"""The oracle's EOS restatement against the reference's own host EOS (engine/eos.cu compiled as
C++ from the reference tree by `make -C oracle ref` into oracle/_ref/libref_eos.so, wrapped by
oracle/ref_eos.cpp).  Both evaluate in FP64 from float32 inputs: pressure, sound speed and
temperature must agree to the float32 result's last bit (one unit in the last place allowed for
a different order of FP64 operations), phases exactly; u_cold, the (rho, T) -> (u, P) map of the
profile builder and its density inversion to FP64 rounding.  Skipped where the reference tree is
not present to build it (the GPU box), since the comparison needs the reference's code."""
import ctypes as C
import itertools
from pathlib import Path

import numpy as np
import pytest

LIB = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "libref_eos.so"
pytestmark = pytest.mark.skipif(not LIB.exists(), reason="oracle/_ref/libref_eos.so is built only where the "
                                "reference tree exists (make -C oracle ref)")


@pytest.fixture(scope="module")
def ref():
    lib = C.CDLL(str(LIB))
    lib.ref_eos_eval.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float] + [C.c_void_p] * 4
    lib.ref_cold_energy_psu.restype = C.c_double
    lib.ref_cold_energy_psu.argtypes = [C.c_int, C.c_double]
    lib.ref_eos_from_rho_temp_si.argtypes = [C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    lib.ref_invert_density_si.restype = C.c_double
    lib.ref_invert_density_si.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    return lib


def ref_eval(lib, mat, rho, u, dmg):
    p, cs, t = C.c_float(), C.c_float(), C.c_float()
    ph = C.c_uint8()
    lib.ref_eos_eval(mat, rho, u, dmg, C.byref(p), C.byref(cs), C.byref(t), C.byref(ph))
    return p.value, cs.value, t.value, ph.value


def ulps(a, b):
    a, b = np.float32(a), np.float32(b)
    if a == b:
        return 0
    ia, ib = np.array([a, b]).view(np.int32).astype(np.int64)
    ia = ia if ia >= 0 else -(ia & 0x7FFFFFFF)
    ib = ib if ib >= 0 else -(ib & 0x7FFFFFFF)
    return abs(int(ia - ib))


def grid(orc, mat):
    m = orc.material(mat)
    rho0 = m["rho0"] / 1e6
    etas = [0.3, 0.6, 0.9, 0.95, 0.999, 0.99999, 1.0, 1.000001, 1.00001, 1.001, 1.05, 1.3, 2.0, 3.5]
    if mat == 3:
        us = [0.0, 0.01, 0.5, 2.0, 10.0, 100.0]
    else:
        e = [m["Eiv"] / 1e6, m["Ecv"] / 1e6, m["E0"] / 1e6]
        us = [0.0, 1e-4, 0.05, 0.3, 0.9 * e[0], e[0], 0.5 * (e[0] + e[1]), e[1], 1.5 * e[1], 3 * e[2], 40.0]
    for eta, u, d in itertools.product(etas, us, (0.0, 0.5, 1.0)):
        yield np.float32(eta * rho0), np.float32(u), np.float32(d)


@pytest.mark.parametrize("mat", [0, 1, 2, 3, 4])
def test_eos_matches_the_reference_code(orc, ref, mat):
    worst = [0, 0, 0]
    n = 0
    for rho, u, d in grid(orc, mat):
        got = orc.eos_eval(mat, float(rho), float(u), float(d))
        want = ref_eval(ref, mat, float(rho), float(u), float(d))
        assert got[3] == want[3], (mat, rho, u, d, got, want)
        for k in range(3):
            worst[k] = max(worst[k], ulps(got[k], want[k]))
        n += 1
    assert n >= 252 and max(worst) <= 1, worst  # measured: 0 ulp for every material


@pytest.mark.parametrize("mat", [0, 1, 2, 3, 4])
def test_cold_energy_and_profile_maps_match_the_reference_code(orc, ref, mat):
    m = orc.material(mat)
    rho0 = m["rho0"] / 1e6
    for eta in (0.5, 1.0, 1.0 + 1e-7, 1.0 + 1e-6, 1.00002, 1.01, 1.5, 3.0):
        a, b = orc.cold_energy_psu(mat, eta * rho0), ref.ref_cold_energy_psu(mat, eta * rho0)
        assert abs(a - b) <= 1e-14 * max(abs(b), 1e-300), (mat, eta, a, b)
    for eta, T in itertools.product((0.7, 1.0, 1.2, 2.0), (50.0, 293.0, 1500.0, 6000.0)):
        u, p = orc.eos_from_rho_temp_si(mat, eta * m["rho0"], T)
        ur, pr = C.c_double(), C.c_double()
        ref.ref_eos_from_rho_temp_si(mat, eta * m["rho0"], T, C.byref(ur), C.byref(pr))
        assert abs(u - ur.value) <= 1e-12 * abs(ur.value) and abs(p - pr.value) <= 1e-12 * abs(pr.value) + 1e-6
        if mat != 3 and pr.value > 0:
            g = orc.invert_density_si(mat, pr.value, T, m["rho0"])
            gr = ref.ref_invert_density_si(mat, pr.value, T, m["rho0"])
            assert abs(g - gr) <= 1e-9 * gr, (mat, eta, T, g, gr)


ARCHETYPES = [  # (enum value, name) of engine/types.hpp:27-34
    (0, "LivingTerra"), (1, "IronCore"), (2, "OceanIce"), (3, "LavaIo"), (4, "GasGiant"), (5, "ProtoTheia")]


@pytest.mark.parametrize("code,name", ARCHETYPES)
def test_planet_profile_matches_the_reference_solver(ref, code, name):
    """The library's hydrostatic planet builder (astraea_planet_profile, the product's own code) against
    the reference's PlanetBuilder::solveHydrostatic1D on the test_planet_builder.cpp:24-66 bodies:
    radius and central pressure (both solve the same shooting problem to 1e-6 in mass, by
    different root finders), and the enclosed mass and pressure at the reference's samples."""
    from test_planet_builder import ARCH, _spec
    from astraea_amd import planet_profile

    m0, _, ts, tc, _ = ARCH[name]
    ref.ref_solve_hydrostatic.restype = C.c_int
    ref.ref_solve_hydrostatic.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                          C.c_int] + [C.c_void_p] * 5
    R, Pc = C.c_double(), C.c_double()
    cap = 20000
    arr = {k: np.zeros(cap) for k in ("r", "rho", "p", "m", "t")}
    n = ref.ref_solve_hydrostatic(code, m0, ts, tc, C.byref(R), C.byref(Pc), cap,
                                  *[arr[k].ctypes.data for k in ("r", "rho", "p", "m", "t")])
    assert 16 <= n <= cap
    pr = planet_profile(_spec(name))
    assert abs(pr["radius_m"] / R.value - 1.0) < 2e-4, (pr["radius_m"], R.value)
    assert abs(pr["central_pressure_pa"] / Pc.value - 1.0) < 2e-4, (pr["central_pressure_pa"], Pc.value)
    r_ref, m_ref, p_ref = arr["r"][:n], arr["m"][:n], arr["p"][:n]
    inside = r_ref < 0.999 * min(R.value, pr["radius_m"])
    m_got = np.interp(r_ref[inside], pr["r"], pr["m"])
    p_got = np.interp(r_ref[inside], pr["r"], pr["p"])
    assert np.max(np.abs(m_got - m_ref[inside])) < 2e-4 * m0
    assert np.max(np.abs(p_got - p_ref[inside])) < 2e-4 * Pc.value
