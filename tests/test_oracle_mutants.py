"""Tests of the oracle's pinning tests: each mutant below is a copy of oracle/oracle.cpp with one
constant, sign or index of the SPH physics changed by hand (the kind of slip a restatement can
make), built into a temporary directory.  tests/test_oracle_sph_pins.py, run against the mutant
(through ASTRAEA_ORACLE_LIB), must fail in the test named for it; the unchanged copy must pass
there.  CPU only: a few seconds per build, built in parallel."""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "oracle.cpp"
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared"]

# (name, text in oracle.cpp, replacement, the pinning test that must fail)
MUTANTS = [
    ("viscosity beta 1", "pi = ((-cb * mu + 2.0 * mu * mu) / rb)", "pi = ((-cb * mu + 1.0 * mu * mu) / rb)",
     "test_pass2_artificial_viscosity_and_balsara"),
    ("viscosity alpha 2", "pi = ((-cb * mu + 2.0 * mu * mu) / rb)", "pi = ((-2.0 * cb * mu + 2.0 * mu * mu) / rb)",
     "test_pass2_artificial_viscosity_and_balsara"),
    ("no Balsara switch", "fB[i] = balsara(", "fB[i] = 1.0 + 0.0 * balsara(",
     "test_pass2_artificial_viscosity_and_balsara"),
    ("tensile normaliser 0.6", "/ 0.6328125", "/ 0.6", "test_pass2_tensile_correction"),
    ("delta-SPH 0.1", "0.05 * Hi * csI * mask * dDf", "0.1 * Hi * csI * mask * dDf",
     "test_pass2_delta_sph_is_a_laplacian"),
    ("Jaumann rotation sign",
     "const double rxy = 0.5 * (g[0][1] - g[1][0]), rxz = 0.5 * (g[0][2] - g[2][0]), ryz = 0.5 * (g[1][2] - g[2][1]);",
     "const double rxy = 0.5 * (g[1][0] - g[0][1]), rxz = 0.5 * (g[2][0] - g[0][2]), ryz = 0.5 * (g[2][1] - g[1][2]);",
     "test_pass2_jaumann_rate"),
    ("stress force sign", "a[0] += mj * (st[0]", "a[0] -= mj * (st[0]", "test_pass2_stress_force_and_work"),
    ("XSPH 0.5", "0.25 * xs[0]", "0.5 * xs[0]", "test_drift_xsph"),
    ("velocity gradient transposed", "g[a][b] += volJ * dv[a] * gg[b]", "g[b][a] += volJ * dv[a] * gg[b]",
     "test_pass1_velocity_gradient_of_a_linear_field"),
    ("neighbour target 50", "std::cbrt(56.0 / (double)ngb)", "std::cbrt(50.0 / (double)ngb)",
     "test_pass1_h_update_counts_against_56"),
    ("surface factor 1.5", "1.6 * std::max(1.0 - shep, nn / hi)", "1.5 * std::max(1.0 - shep, nn / hi)",
     "test_pass1_shepard_and_surface_factor"),
    ("Grady-Kipp 0.5", "0.4 * cc / hi", "0.5 * cc / hi", "test_grady_kipp_rate_at_the_failure_strain"),
    ("J2 without the half", "const double j2 = 0.5 * (t[0]", "const double j2 = 1.0 * (t[0]",
     "test_kick_von_mises_radial_return"),
    ("friction 0.5 P", "0.6 * std::max(0.0, s.P[i])", "0.5 * std::max(0.0, s.P[i])",
     "test_kick_von_mises_radial_return"),
    ("thermal floor 2 c_v", "coldEnergy(mi, rhoN) + 3.0 * mp.cv", "coldEnergy(mi, rhoN) + 2.0 * mp.cv",
     "test_kick_thermal_clamp"),
    ("thermal floor at the old density", "coldEnergy(mi, rhoN) + 3.0 * mp.cv",
     "coldEnergy(mi, rhoN - hdt * s.drho[i]) + 3.0 * mp.cv", "test_kick_thermal_clamp"),
    ("no EOS after the clamp", "        if (clampThermal) {\n            const EosOut o",
     "        if (false) {\n            const EosOut o", "test_kick_thermal_clamp"),
    ("living ambient 273 K", "? 255.0 : 3.0", "? 273.0 : 3.0", "test_cooling_ambient_and_threshold"),
    ("cooling threshold ignored", "if (tI > coolMin) {", "if (true) {", "test_cooling_ambient_and_threshold"),
    ("diffusion across materials", "if (s.mat[j] == matI) {", "if (true) {", "test_pass2_delta_sph_is_a_laplacian"),
]


def _build(src_text, out):
    with tempfile.NamedTemporaryFile("w", suffix=".cpp", dir=out.parent, delete=False) as f:
        f.write(src_text)
    try:
        subprocess.run(["g++", *FLAGS, "-I", str(SRC.parent), "-o", str(out), f.name], check=True,
                       capture_output=True, text=True)
    finally:
        os.unlink(f.name)


def _pins(lib, test):
    env = dict(os.environ, ASTRAEA_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    return subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_oracle_sph_pins.py"), "-q",
                           "-p", "no:cacheprovider", "-k", test], env=env, capture_output=True, text=True, cwd=ROOT)


@pytest.fixture(scope="module")
def builds():
    src = SRC.read_text()
    with tempfile.TemporaryDirectory() as d:
        d = Path(d)
        jobs = [("unchanged", src, d / "lib_unchanged.so")]
        for k, (name, old, new, _) in enumerate(MUTANTS):
            assert src.count(old) == 1, f"{name}: the text to mutate must occur once in oracle.cpp"
            jobs.append((name, src.replace(old, new), d / f"lib_mutant{k}.so"))
        runs = [("unchanged", "test_")] + [(m[0], m[3]) for m in MUTANTS]
        libs = {name: lib for name, _, lib in jobs}
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            list(ex.map(lambda j: _build(j[1], j[2]), jobs))
            results = dict(zip([r[0] for r in runs], ex.map(lambda r: _pins(libs[r[0]], r[1]), runs)))
        yield results


def test_unchanged_copy_passes_the_pins(builds):
    r = builds["unchanged"]
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("name,old,new,target", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_fails_its_pin(builds, name, old, new, target):
    r = builds[name]
    assert r.returncode == 1, f"{name} survived {target}:\n{r.stdout[-1500:]}"
    assert "failed" in r.stdout and "error" not in r.stdout.split("\n")[-2], r.stdout[-1500:]
