"""CPU checks of the C-ABI library and of the host-side initial conditions.

- the shared library loads and exports every function that include/astraea_hip.h declares
  (no device calls);
- field sizes and the default parameters equal the reference's compile-time constants;
- without a device, creating a context fails with a status code and a message;
- bodies: exact particle count, mass fractions per layer, centre of mass at the origin,
  the hydrostatic energy reproduces the analytic pressure P(r) through the EOS, determinism;
- placement: the Keplerian approach of engine/simulation.cpp:1389-1436 (separation, speed at
  contact from energy conservation, zero total momentum).
"""
import math
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "astraea_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(astraea_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib._lib, n)]
    assert not missing, missing
    assert sorted(lib.EXPORTED_SYMBOLS) == names


def ctypes_size(t):
    import ctypes

    return ctypes.sizeof(t)


def test_field_sizes_and_defaults(lib):
    f = lib.FIELD
    assert lib._lib.astraea_field_size(f["pos_x"]) == 4
    assert lib._lib.astraea_field_size(f["material"]) == 1
    assert lib._lib.astraea_field_size(f["phase"]) == 1
    assert lib._lib.astraea_field_size(f["pos64_x"]) == 8
    assert lib._lib.astraea_field_size(f["original_id"]) == 4
    assert lib._lib.astraea_field_size(999) == 0
    p = lib.default_params()
    # sph_gravity.cu:842 (0.8 h0), 1510 (0.25), eos.cu:93 (1.4), simulation.cpp:1729 (theta 0.5)
    assert (p.theta, p.cfl_factor, p.strength_mask, p.precision) == (0.5, 0.25, 0x17, 0)
    assert abs(p.soft_factor - 0.8) < 1e-7 and abs(p.gamma_hhe - 1.4) < 1e-7
    assert p.group_size == 32 and lib.GROUP_SIZE == 64  # the reference's warps; 0 means one wavefront
    from astraea_amd import scenario as S

    assert S.plain_params().group_size == 64  # the benchmark configurations: one wavefront
    assert ctypes_size(lib.Params) == 48 == ctypes_size(__import__("oracle").Params)


def test_context_without_device_reports_error(lib):
    import ctypes

    h = ctypes.c_void_p()
    rc = lib._lib.astraea_create(ctypes.byref(h), 0, 1000, 1)
    if rc == 0:
        lib._lib.astraea_destroy(h)
        pytest.skip("a device is present")
    assert rc < 0
    assert len(lib._lib.astraea_last_error()) > 0


def test_splitmix64_known_values(lib):
    # splitmix64 (Steele et al. 2014) reference outputs for state 0 -> first value
    assert lib.splitmix64(0) == 0xE220A8397B1DCDAF
    assert lib.splitmix64(42) != lib.splitmix64(43)


@pytest.fixture(scope="module")
def body(lib):
    from astraea_amd import scenario as S

    return lib.build_body(20000, 6.371e6, S.FE_GRANITE, seed=42)


def test_body_count_mass_layers(body, orc):
    n = len(body["x"])
    assert n == 20000
    m = body["m"].astype(np.float64)
    fe = body["mat"] == 0
    assert abs(m.sum() * 1e24 - body["mass_kg"]) <= 1e-5 * body["mass_kg"]
    assert abs(m[fe].sum() / m.sum() - 0.3) < 0.01
    # total mass of the two uniform layers: (4/3) pi R^3 (rho_c x^3 + rho_m (1 - x^3)), x^3 from the core fraction
    rc, rm = 7800.0, 2680.0
    x3 = 0.3 * rm / (0.7 * rc + 0.3 * rm)
    mass = 4 / 3 * math.pi * 6.371e6 ** 3 * (rc * x3 + rm * (1 - x3))
    assert abs(body["mass_kg"] - mass) <= 1e-9 * mass
    com = [(m * body[k]).sum() / m.sum() for k in ("x", "y", "z")]
    assert max(abs(c) for c in com) < 1e-5
    r = np.sqrt(body["x"] ** 2 + body["y"] ** 2 + body["z"] ** 2)
    assert r.max() < 6.371 * 1.02
    # h = 1.3 * lattice spacing; HCP: n s^3 / sqrt(2) = volume
    s = (math.sqrt(2) * 4 / 3 * math.pi * 6.371e6 ** 3 / n) ** (1 / 3)
    assert abs(body["spacing_m"] - s) < 1e-6 * s
    assert np.allclose(body["h"], 1.3 * s / 1e6)


def test_body_hydrostatic_energy(body, orc):
    """The EOS pressure of every particle equals the closed-form hydrostatic pressure at its radius."""
    G = 6.6743e-11
    R = 6.371e6
    rc_, rm_ = 7800.0, 2680.0
    x3 = 0.3 * rm_ / (0.7 * rc_ + 0.3 * rm_)
    Rc = R * x3 ** (1 / 3)
    Mc = 4 / 3 * math.pi * rc_ * Rc ** 3

    def p_of_r(r):
        km = 4 / 3 * math.pi * rm_
        if r >= Rc:
            return rm_ * G * ((Mc - km * Rc ** 3) * (1 / r - 1 / R) + 0.5 * km * (R * R - r * r))
        pc = rm_ * G * ((Mc - km * Rc ** 3) * (1 / Rc - 1 / R) + 0.5 * km * (R * R - Rc * Rc))
        kc = 4 / 3 * math.pi * rc_
        return pc + rc_ * G * 0.5 * kc * (Rc * Rc - r * r)

    rng = np.random.default_rng(0)
    # the lattice is centred at the origin before the COM shift, which is below 1e-5 PSU
    for k in rng.choice(len(body["x"]), 200, replace=False):
        r = math.sqrt(body["x"][k] ** 2 + body["y"][k] ** 2 + body["z"][k] ** 2) * 1e6
        mat = int(body["mat"][k])
        if abs(r - Rc) < 2e4:
            continue  # particles within the COM shift of the interface
        p, _, _, _ = orc.eos_eval(mat, float(body["rho"][k]), float(body["u"][k]), 0.0)
        exp = p_of_r(r)
        assert abs(p * 1e12 - exp) <= 2e-5 * max(exp, 1e8) + 2e4 * 1e4, (k, p * 1e12, exp)


def test_body_deterministic_and_seeded(lib):
    from astraea_amd import scenario as S

    a = lib.build_body(3000, 3.0e6, S.FE_GRANITE, seed=7)
    b = lib.build_body(3000, 3.0e6, S.FE_GRANITE, seed=7)
    c = lib.build_body(3000, 3.0e6, S.FE_GRANITE, seed=8)
    assert all(np.array_equal(a[k], b[k]) for k in ("x", "y", "z", "m", "u"))
    assert not np.array_equal(a["x"], c["x"])


def test_weibull_flaws(lib):
    from astraea_amd import scenario as S

    a = lib.build_body(5000, 3.0e6, S.FE_GRANITE, seed=3, flaws=True)
    e = a["eps_fail"].astype(np.float64)
    v = 4 / 3 * math.pi * 3.0e6 ** 3 / 5000
    # k V eps^m = -ln U is Exp(1): its sample mean is 1
    t = 4.0e29 * v * e ** 9
    assert abs(t.mean() - 1.0) < 0.08


def test_two_body_placement(lib):
    from astraea_amd import scenario as S

    st = S.impact_pair(3000, 600, 45.0, 1.1)
    m = st["mass"].astype(np.float64)
    x = np.stack([st["pos_x"], st["pos_y"], st["pos_z"]], 1).astype(np.float64)
    v = np.stack([st["vel_x"], st["vel_y"], st["vel_z"]], 1).astype(np.float64)
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-6 * m.sum()
    assert np.abs((m[:, None] * x).sum(0)).max() < 1e-5 * m.sum()
    nt = 3000
    mt, mi = m[:nt].sum(), m[nt:].sum()
    ct = (m[:nt, None] * x[:nt]).sum(0) / mt
    ci = (m[nt:, None] * x[nt:]).sum(0) / mi
    vt = (m[:nt, None] * v[:nt]).sum(0) / mt
    vi = (m[nt:, None] * v[nt:]).sum(0) / mi
    rt, ri = 6.371, 6.371 * 0.1 ** (1 / 3)
    rc = rt + ri
    d0 = np.linalg.norm(ci - ct)
    assert abs(d0 - 2.0 * rc) < 1e-4 * rc
    vrel = np.linalg.norm(vi - vt)
    G = 66.743
    vesc = math.sqrt(2 * G * (mt + mi) / rc)
    vc2 = vrel ** 2 + 2 * G * (mt + mi) * (1 / rc - 1 / d0)
    assert abs(math.sqrt(vc2) - 1.1 * vesc) < 1e-4 * vesc
    # angular momentum of the relative orbit: |d x v| = rc * v_contact * b, b = sin 45 deg
    lrel = np.linalg.norm(np.cross(ci - ct, vi - vt))
    assert abs(lrel - rc * 1.1 * vesc * math.sin(math.radians(45))) < 1e-4 * lrel


def test_gas_giant_envelope_hydrostatic(lib, orc):
    """H/He envelope (85 % of the mass) around an iron + water-ice core: the n = 1 polytrope
    P = K rho^2 of the H/He EOS in hydrostatic equilibrium, re-solved here independently
    (scipy's RK45 and a root find on the density at the core's top) for the mass the builder
    returned: the surface lands at the requested radius, and every envelope particle's density
    matches the profile at its radius, with u = K rho (no thermal part, so P_EOS = K rho^2)."""
    from scipy.integrate import solve_ivp
    from scipy.optimize import brentq
    from astraea_amd import scenario as S

    n, R = 12000, S.R_GAS_GIANT_M
    b = lib.build_body(n, R, S.GAS_GIANT, seed=5)
    G, K = 6.6743e-11, 2.08e5
    M = b["mass_kg"]
    m = b["m"].astype(np.float64)
    assert abs(m.sum() * 1e24 - M) <= 1e-5 * M
    counts = np.bincount(b["mat"], minlength=4)
    assert abs(counts[0] - 0.075 * n) <= 1 and abs(counts[2] - 0.075 * n) <= 1 and abs(counts[3] - 0.85 * n) <= 1
    rc = (3 / (4 * math.pi) * (0.075 * M / 7800.0 + 0.075 * M / 1000.0)) ** (1 / 3)

    def envelope(rho_top):
        def rhs(r, y):
            return [-G * y[1] / (2 * K * r * r), 4 * math.pi * r * r * max(y[0], 0.0)]
        surf = lambda r, y: y[0]
        surf.terminal = True
        return solve_ivp(rhs, [rc, 3 * R], [rho_top, 0.15 * M], events=surf, rtol=1e-10, atol=[1e-9, 1e10],
                         dense_output=True)

    top = brentq(lambda x: envelope(math.exp(x)).y_events[0][0][1] - M, math.log(1.0), math.log(1e5), xtol=1e-12)
    sol = envelope(math.exp(top))
    r_surface = sol.t_events[0][0]
    assert abs(r_surface - R) <= 1e-3 * R
    env = b["mat"] == 3
    r = np.sqrt(b["x"] ** 2 + b["y"] ** 2 + b["z"] ** 2).astype(np.float64) * 1e6
    inner = env & (r < 0.95 * R)  # the outer rim holds the density floor of the builder
    rho_ref = np.array([sol.sol(x)[0] for x in r[inner]])
    rho = b["rho"].astype(np.float64)[inner] * 1e6  # PSU density unit: 1e24 kg / (1e6 m)^3
    # the COM shift moves particles by a fraction of the spacing (measured: median 4.5e-5, max 6.3e-4)
    assert np.median(np.abs(rho / rho_ref - 1)) < 2e-4
    assert np.max(np.abs(rho / rho_ref - 1)) < 3e-3
    for k in np.flatnonzero(env)[::97]:
        u_cold = orc.cold_energy_psu(3, float(b["rho"][k]))
        assert abs(b["u"][k] - u_cold) <= 1e-6 * u_cold
        p, _, t, _ = orc.eos_eval(3, float(b["rho"][k]), float(b["u"][k]), 0.0, gamma=5 / 3)
        assert abs(p - K * float(b["rho"][k]) ** 2) <= 1e-5 * p and t == 3.0
    # core: uniform density at each material's reference density, inside the core radius
    core = ~env
    assert r[core].max() <= rc * 1.01 and r[env].min() >= rc * 0.99


@pytest.mark.parametrize("which", ["rocky", "gas_giant"])
def test_body_radius_for_mass_inverts_the_builder(lib, orc, which):
    """astraea_body_radius_for_mass is the inverse of astraea_build_body's radius -> mass: a body
    built at radius R has mass M, and the radius for M is R again; for uniform layers it is the
    closed form R^3 = 3 M sum(f / rho0) / (4 pi)."""
    from astraea_amd import scenario as S

    layers = S.FE_GRANITE if which == "rocky" else S.GAS_GIANT
    r = 2.0e6 if which == "rocky" else S.R_GAS_GIANT_M
    m = lib.build_body(512, r, layers)["mass_kg"]
    assert lib.body_radius_for_mass(layers, m) == pytest.approx(r, rel=1e-6)
    if which == "rocky":
        v = sum(f / orc.material(mat)["rho0"] for mat, f, _ in layers)
        assert lib.body_radius_for_mass(layers, 1.0e24) ** 3 == pytest.approx(3e24 * v / (4 * np.pi), rel=1e-12)
    with pytest.raises(lib.AstraeaError):
        lib.body_radius_for_mass(layers, -1.0)


def test_pcls_wire_format():
    """The PCLS frame of server/binary_protocol.cpp (include/astraea/binary_protocol.hpp): header
    bytes at their offsets (magic "SLCP", version, flags, f64 time, counts, raw payload size),
    payload order particles | tracers | speed, a speed array of the wrong size dropped, wrong
    magic / version / short headers rejected, zlib round trips (no GPU is used)."""
    import subprocess

    exe = Path(__file__).resolve().parent / "cpp" / "test_simulation"
    r = subprocess.run([str(exe), "pcls"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr


def test_tuning_table_round_trip(lib, tmp_path):
    """tune_cached runs the autotuner once per workload key and applies the stored choice after."""

    class FakeContext:
        def __init__(self):
            self.applied, self.tuned = None, 0

        def set_tuning(self, p1, p2, d, sort_bits=8, gravity_waves=4, gravity_stack=1024, build_waves=4,
                       build_stack=1024):
            self.applied = (p1, p2, d, sort_bits, gravity_waves, gravity_stack, build_waves, build_stack)

        def autotune(self, max_dt, steps_per_variant):
            self.tuned += 1
            return dict(chosen={"pass1_round": 128, "pass2_round": 160, "drift_round": 192, "sort_bits": 9,
                                "gravity_waves": 2, "gravity_stack": 1024, "build_waves": 8, "build_stack": 512})

    path = tmp_path / "tuning.json"
    key = lib.tuning_key(9_602_616, 64, False, True)
    assert key == "gfx950/n24/s64/plain/fp64"
    ctx = FakeContext()
    first = lib.tune_cached(ctx, 5e-3, path, key)
    assert not first["cached"] and ctx.tuned == 1
    again = FakeContext()
    second = lib.tune_cached(again, 5e-3, path, key)
    assert second["cached"] and again.tuned == 0 and again.applied == (128, 160, 192, 9, 2, 1024, 8, 512)
    other = FakeContext()
    assert not lib.tune_cached(other, 5e-3, path, lib.tuning_key(1_000_000, 1, False, False))["cached"]
    import json

    assert set(json.loads(path.read_text())) == {key, "gfx950/n20/s1/plain/fp32"}

    class StaleContext(FakeContext):
        def set_tuning(self, p1, p2, d, *rest):
            if d == 64:  # a drift round this library no longer offers
                raise lib.AstraeaError(lib.E_INVALID, "invalid tuning")
            super().set_tuning(p1, p2, d, *rest)

    table = json.loads(path.read_text())
    table[key]["drift_round"] = 64
    path.write_text(json.dumps(table))
    stale = StaleContext()
    third = lib.tune_cached(stale, 5e-3, path, key)
    assert not third["cached"] and stale.tuned == 1 and json.loads(path.read_text())[key]["drift_round"] == 192
