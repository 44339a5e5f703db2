"""Oracle largest-remnant classifier (engine/simulation.cpp:1000-1227) on hand-solvable cases.

A test particle of negligible mass moving about a heavy central particle follows the two-body
closed forms: circular speed sqrt(GM/r), escape speed sqrt(2GM/r), and for a tangential launch
at distance r with speed k v_c a periapsis r k^2 / (2 - k^2).  Rigid rotation about Y gives
L_y / I_yy = omega exactly.  Each expectation below comes from those forms, not from the code.
"""
import numpy as np
import pytest

import oracle

G = 66.7430  # PSU gravitational constant (engine/units.hpp:26)
IDX = dict(body0_mass=0, body1_mass=7, seed=14, rem_mass=20, rem_pos=21, rem_radius=27, planet=28, disk=29,
           ejecta=30, imp_bound=31, spin_l=32, spin_i=33, max_damage=34, iterations=35)


def _system(particles):
    """particles: list of (pos, vel, mass, body, phi, damage)."""
    cols = list(zip(*particles))
    pos = np.array(cols[0], np.float64)
    vel = np.array(cols[1], np.float64)
    return dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], vx=vel[:, 0], vy=vel[:, 1], vz=vel[:, 2],
                m=np.array(cols[2], np.float64), body=np.array(cols[3], np.uint8),
                phi=np.array(cols[4], np.float32), damage=np.array(cols[5], np.float32))


def _classify(s, rt):
    return oracle.classify_arrays(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["m"], s["phi"], s["body"],
                                  s["damage"], rt)


M = 1.0
EPS = 1.0e-9  # test-particle mass


def _test_particle(r, speed, body=0, tangential=True, damage=0.0):
    vel = (0.0, speed, 0.0) if tangential else (-speed, 0.0, 0.0)
    return ((r, 0.0, 0.0), vel, EPS, body, -1.0, damage)


def test_orbits_about_point_remnant():
    vc = lambda r: np.sqrt(G * M / r)
    k_ecc_in, k_ecc_out = 0.5, 0.7  # periapsis 3 k^2/(2-k^2) = 0.43 (inside 0.5) and 0.97 (outside)
    parts = [((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), M, 0, -100.0, 0.3)]
    parts += [
        _test_particle(3.0, vc(3.0)),                                   # circular at 3: disk
        _test_particle(4.0, vc(4.0), body=1),                           # circular, impactor: disk + bound
        _test_particle(3.0, 1.01 * np.sqrt(2.0) * vc(3.0)),             # above escape: ejecta
        _test_particle(3.0, 0.99 * np.sqrt(2.0) * vc(3.0), body=1),     # just bound, periapsis ~2.9: disk
        _test_particle(3.0, 0.5 * vc(3.0), tangential=False),           # radial infall: planet
        _test_particle(3.0, k_ecc_in * vc(3.0), damage=0.7),            # periapsis 0.43 < 0.5: planet
        _test_particle(3.0, k_ecc_out * vc(3.0)),                       # periapsis 0.97 > 0.5: disk
        _test_particle(0.3, 0.0, body=1),                               # inside the remnant: planet
    ]
    out = _classify(_system(parts), rt=1.0)
    # remnant radius: max(0.5 R_t, sqrt(5/3) rms r) = 0.5 since the rms radius is ~1e-4
    assert out[IDX["rem_radius"]] == pytest.approx(0.5, abs=1e-12)
    assert out[IDX["iterations"]] == 5
    assert out[IDX["disk"]] == pytest.approx(4 * EPS, rel=1e-12)
    assert out[IDX["ejecta"]] == pytest.approx(1 * EPS, rel=1e-12)
    assert out[IDX["planet"]] == pytest.approx(M + 3 * EPS, rel=1e-12)
    assert out[IDX["imp_bound"]] == pytest.approx(3 * EPS, rel=1e-12)
    assert out[IDX["max_damage"]] == pytest.approx(0.7, rel=1e-7)
    assert out[IDX["rem_mass"]] == pytest.approx(M + 7 * EPS, rel=1e-12)
    np.testing.assert_allclose(out[IDX["rem_pos"]:IDX["rem_pos"] + 3], [0, 0, 0], atol=1e-7)


def test_periapsis_threshold_is_sharp():
    # periapsis r k^2/(2-k^2) against remnant radius 0.5 at launch radius 3: k* = sqrt(2 * 0.5 / 3.5)
    k_star = np.sqrt(1.0 / 3.5)
    vc = np.sqrt(G * M / 3.0)
    core = ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), M, 0, -100.0, 0.0)
    for k, disk in [(k_star * (1 + 1e-6), True), (k_star * (1 - 1e-6), False)]:
        out = _classify(_system([core, _test_particle(3.0, k * vc)]), rt=1.0)
        assert (out[IDX["disk"]] > 0) == disk, k


def test_rigid_rotation_spin_and_radius_cut():
    rng = np.random.default_rng(7)
    n = 2000
    pts = rng.uniform(-1, 1, (n, 3))
    pts = pts[np.linalg.norm(pts, axis=1) < 1.0]
    omega = 0.37
    vel = np.stack([omega * pts[:, 2], np.zeros(len(pts)), -omega * pts[:, 0]], axis=1)
    parts = [(tuple(p), tuple(v), 1.0 / len(pts), 0, -1.0 - np.dot(p, p), 0.0) for p, v in zip(pts, vel)]
    # a body-0 pair beyond 1.5 R_t spinning the other way (centre and mean velocity unchanged)
    # must be left out of the spin
    parts.append(((1.7, 0.0, 0.0), (0.0, 0.0, 5.0), 1e-3, 0, 0.0, 0.0))
    parts.append(((-1.7, 0.0, 0.0), (0.0, 0.0, -5.0), 1e-3, 0, 0.0, 0.0))
    s = _system(parts)
    out = _classify(s, rt=1.0)
    assert out[IDX["spin_l"]] / out[IDX["spin_i"]] == pytest.approx(omega, rel=1e-9)
    # the seed is the particle of lowest phi, i.e. nearest the centre
    j = int(np.argmin(s["phi"]))
    np.testing.assert_array_equal(out[IDX["seed"]:IDX["seed"] + 3], [s["x"][j], s["y"][j], s["z"][j]])


def test_seed_ties_take_first_index_and_nonfinite_phi_is_skipped():
    parts = [((5.0, 0, 0), (0, 0, 0), 1.0, 0, np.nan, 0.0),
             ((1.0, 0, 0), (0, 0, 0), 1.0, 0, -2.0, 0.0),
             ((2.0, 0, 0), (0, 0, 0), 1.0, 0, -2.0, 0.0),
             ((3.0, 0, 0), (0, 0, 0), 1.0, 1, -1.0, 0.0)]
    out = _classify(_system(parts), rt=1.0)
    assert out[IDX["seed"]] == 1.0
    # every phi non-finite: the reference keeps seed index 0
    parts = [((4.0, 0, 0), (0, 0, 0), 1.0, 0, np.inf, 0.0), ((1.0, 0, 0), (0, 0, 0), 1.0, 0, np.nan, 0.0)]
    out = _classify(_system(parts), rt=1.0)
    assert out[IDX["seed"]] == 4.0


def test_iteration_stops_when_nothing_is_bound():
    # Guess mass max(m0, M/2) = 51 binds the counter-moving pair P1, P2 (2 u^2 < G 51 / 2); their
    # own mass 2 does not (u^2 / 2 > G 2 / 1), so the second iteration finds nothing and stops
    # with the first iteration's remnant: mass 2 at the origin.  Its rms radius is taken about the
    # guess centre, the seed P1 (simulation.cpp:1151-1163): sqrt((0 + 2^2) / 2), times sqrt(5/3).
    u = 20.0
    parts = [((-1.0, 0, 0), (0, -u, 0), 1.0, 0, -3.0, 0.0),
             ((1.0, 0, 0), (0, u, 0), 1.0, 0, -2.0, 0.0),
             ((1000.0, 0, 0), (1e4, 0, 0), 100.0, 1, -1.0, 0.0)]
    out = _classify(_system(parts), rt=1.0)
    assert out[IDX["iterations"]] == 1
    assert out[IDX["rem_mass"]] == 2.0
    np.testing.assert_allclose(out[IDX["rem_pos"]:IDX["rem_pos"] + 3], [0, 0, 0], atol=1e-15)
    assert out[IDX["rem_radius"]] == pytest.approx(np.sqrt(5.0 / 3.0 * 2.0), rel=1e-14)
    assert out[IDX["ejecta"]] == 102.0 and out[IDX["planet"]] == 0.0 and out[IDX["disk"]] == 0.0
    # all masses zero: nothing is ever bound, the guess (mass 0, seed state, radius max(R_t, .25)) stays
    parts0 = [((2.0, 0, 0), (1.0, 0, 0), 0.0, 0, -1.0, 0.0), ((3.0, 0, 0), (0, 0, 0), 0.0, 1, -2.0, 0.0)]
    out0 = _classify(_system(parts0), rt=0.1)
    assert out0[IDX["iterations"]] == 0 and out0[IDX["rem_radius"]] == 0.25
    assert out0[IDX["rem_pos"]] == 3.0


def test_body_sums():
    parts = [((1.0, 2.0, 3.0), (0.1, 0.2, 0.3), 2.0, 0, -1.0, 0.0),
             ((3.0, 2.0, 1.0), (0.3, 0.2, 0.1), 2.0, 0, -1.0, 0.0),
             ((-1.0, 0.0, 0.0), (0.0, 0.0, -1.0), 1.0, 2, -1.0, 0.0)]
    out = _classify(_system(parts), rt=1.0)
    assert out[IDX["body0_mass"]] == 4.0 and out[IDX["body1_mass"]] == 1.0
    np.testing.assert_allclose(out[1:7], [2.0, 2.0, 2.0, 0.2, 0.2, 0.2])
    np.testing.assert_allclose(out[8:14], [-1.0, 0.0, 0.0, 0.0, 0.0, -1.0])
