"""Render-record oracle (oracle/render.py) pinned on what the format fixes:

* decoding with the reference's own decoder (tests/cpp/test_conservation.cpp:320-343) returns
  each quantity to within half a quantisation step of the log scale;
* the ends of every range map to the first and last code (T = 10 K and 200,000 K; P <= 1e-3 GPa
  is code 0, P = 1e4 GPa the last code; radius 1 km and 2,000 km; speed 0.05 km/s and 100 km/s);
* the byte layout: float32 km about the centre at 0, 4, 8; u16 at 12 and 14; body in the high
  nibble of byte 17.
"""
import numpy as np
import pytest

from oracle import render as R


def test_temperature_round_trip_and_ends():
    t = np.exp(np.random.default_rng(1).uniform(np.log(10.0), np.log(2.0e5), 5000))
    dec = 10.0 * np.exp(R.encode_temp(t) / 65535.0 * R.LN20000)
    half = 0.5 * R.LN20000 / 65535.0
    assert np.max(np.abs(np.log(dec / t))) <= half * (1 + 1e-9)
    assert list(R.encode_temp([5.0, 10.0, 2.0e5, 1e9, np.nan])) == [0, 0, 65535, 65535, 0]


def test_pressure_round_trip_and_ends():
    p_gpa = np.exp(np.random.default_rng(2).uniform(np.log(1.1e-3), np.log(1.0e4), 5000))
    code = R.encode_pressure(p_gpa / 1000.0)  # PSU pressure is TPa
    assert np.all(code >= 1)
    dec = 1.0e-3 * np.exp((code - 1.0) / 65534.0 * R.LN1E7)
    assert np.max(np.abs(np.log(dec / p_gpa))) <= 0.5 * R.LN1E7 / 65534.0 * (1 + 1e-9)
    assert list(R.encode_pressure(np.array([0.0, 1.0e-6, -1.0, 10.0]))) == [0, 0, 0, 65535]


def test_radius_speed_damage_ends():
    rho = np.array([1.0, 1.0, 1.0])
    # volume-equivalent radius r km: mass = 4/3 pi (r / 1000)^3 rho
    m = 4.0 / 3.0 * np.pi * (np.array([0.5, 1.0, 2000.0]) / 1000.0) ** 3
    assert list(R.encode_radius(m, rho)) == [0, 0, 255]
    v = np.array([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0], [100.0, 0.0, 0.0], [0.0, 3.0, 4.0]])
    s = R.encode_speed(v)
    assert list(s[:3]) == [0, 0, 255]
    assert R.decode_speed(s[3]) == pytest.approx(5.0, rel=0.5 * R.LN2000 / 255.0 * 1.01)
    assert list(R.encode_damage([-1.0, 0.0, 0.5, 1.0, 2.0, np.nan])) == [0, 0, 128, 255, 255, 0]


def test_record_layout():
    pos = np.array([[1.0, 2.0, 3.0], [0.5, 0.25, -1.0]], np.float32)
    com = np.array([0.5, 0.5, 0.5], np.float32)
    rec, spd = R.pack_records(pos, np.zeros((2, 3)), [1e-6, 1e-6], [1.0, 1.0], [10.0, 2.0e5], [0.0, 10.0],
                              [1, 4], [0, 1], [2, 5], [0.0, 1.0], com)
    assert rec.shape == (2, 20) and rec.dtype == np.uint8
    xyz = rec[:, :12].copy().view("<f4").reshape(2, 3)
    np.testing.assert_array_equal(xyz, (pos - com) * np.float32(1000.0))
    assert rec[0, 12] == 0 and rec[0, 13] == 0 and rec[1, 12] == 0xFF and rec[1, 13] == 0xFF
    assert rec[0, 14:16].tolist() == [0, 0] and rec[1, 14:16].tolist() == [0xFF, 0xFF]
    assert rec[:, 16].tolist() == [1, 4] and rec[:, 17].tolist() == [0x02, 0x15] and rec[:, 18].tolist() == [0, 255]
    d = R.decode_records(rec)
    assert d["body"].tolist() == [0, 1] and d["phase"].tolist() == [2, 5]
    assert d["temp_k"][1] == pytest.approx(2.0e5) and d["pressure_gpa"][1] == pytest.approx(1.0e4)


def test_tracer_record_fields():
    # one tracer on a particle displaced 10 km outward along +x from rest radius 1 PSU about the centre
    pos = np.array([[1.01, 0.0, 0.0], [0.0, 2.0, 0.0]], np.float32)
    rest = np.array([[1.0, 0.0, 1.0, 0.0]], np.float32)  # rest normal +y: the particle turned by pi/2
    out = R.tracer_records(pos, [0, -1], rest, [1], [0.9, 0.0], [2.0, 300.0], [0, 0], [-2.0, 0.0], [0.1, 0.1],
                           [0.0, 0.0, 0.0], [0.01, 0.0, 0.0])
    head = out[0, :4].copy().view("<u4")[0]
    f = out[0, 4:].copy().view("<f4")
    assert head & 0xFFFF == 0 and (head >> 16) == 0b11  # ocean, disrupted (damage 0.9 >= 0.85)
    assert f[0] == pytest.approx(1000.0) and f[1] == 0.0 and f[2] == 0.0  # x km about the system centre
    assert f[3] == pytest.approx(10.0, rel=1e-5)  # radial displacement km
    assert f[4] == pytest.approx(np.pi / 2, rel=1e-6)  # normal change
    assert f[5] == 3.0 and f[6] == pytest.approx(0.9)  # T floored at 3 K, damage
    assert f[7] == pytest.approx(1.225 * 340.0 * 2.0 * 0.1 * 1000.0 * 1e-3, rel=1e-6)  # overpressure kPa
    assert f[8] == pytest.approx(10.0 * 1000.0, rel=1e-5)  # tsunami height m (ocean)
