"""Render-frame packing restated in numpy: the checker of the device's render records.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product never imports it.

* ``pack_records`` restates packRenderRecordsKernel (engine/sph_gravity.cu:1671-1743): 20-byte
  particle records and the optional speed bytes.  Positions use float32 arithmetic as the kernel
  does; the log quantisations are evaluated in float64 and rounded half to even (the kernel's
  lrintf), so a code may differ from the device's float32 logf by one step at a rounding edge.
* ``tracer_records`` restates extractTracersKernel (sph_gravity.cu:1586-1668): 40-byte records.
* ``decode_records`` inverts the quantisation with the decoding of the reference's own test
  (tests/cpp/test_conservation.cpp:320-343).

Parity: pinned by tests/test_oracle_render.py on the decoding round trip (within half a
quantisation step), the codes at the ends of each range and the byte layout.
"""
from __future__ import annotations

import numpy as np

LN20000 = np.log(20000.0)
LN1E7 = np.log(1.0e7)
LN2000 = np.log(2000.0)
VAPOR, PLASMA = 2, 3


def _f32(a):
    return np.asarray(a, np.float32)


def _finite0(a):
    a = np.asarray(a)
    return np.where(np.isfinite(a), a, a.dtype.type(0))


def _quant(norm, scale):
    return np.rint(scale * np.clip(norm, 0.0, 1.0)).astype(np.int64)


def encode_temp(temp_k):
    t = np.maximum(10.0, _finite0(np.asarray(temp_k, np.float64)))
    return _quant(np.log(t * 0.1) / LN20000, 65535.0).astype(np.uint16)


def encode_pressure(pressure_psu):
    p_gpa = _finite0(np.asarray(pressure_psu, np.float64)) * 1000.0
    code = 1 + _quant(np.log(np.maximum(p_gpa, 1e-30) * 1000.0) / LN1E7, 65534.0)
    return np.where(p_gpa > 1.0e-3, code, 0).astype(np.uint16)


def encode_radius(mass, rho):
    vol = np.maximum(np.asarray(mass, np.float64), 0.0) / np.maximum(np.asarray(rho, np.float64), 1.0e-9)
    r_km = np.cbrt(3.0 * vol / (4.0 * np.pi)) * 1000.0
    return _quant(np.log(np.maximum(r_km, 1.0)) / LN2000, 255.0).astype(np.uint8)


def encode_speed(vel):
    v = np.asarray(vel, np.float64)
    s = _finite0(np.sqrt((v * v).sum(axis=1)))
    return _quant(np.log(np.maximum(s, 0.05) * 20.0) / LN2000, 255.0).astype(np.uint8)


def encode_damage(damage):
    return _quant(_finite0(np.asarray(damage, np.float64)), 255.0).astype(np.uint8)


def pack_records(pos, vel, mass, rho, temp_k, pressure, material, body, phase, damage, com):
    """Records (n, 20) uint8 and speed bytes (n,) of particles about ``com`` (PSU, float32)."""
    pos = _f32(pos)
    com = _f32(com)
    n = len(pos)
    xyz = _finite0((pos - com[None, :]) * np.float32(1000.0)).astype(np.float32)
    rec = np.zeros((n, 20), np.uint8)
    rec[:, 0:12] = xyz.view(np.uint8).reshape(n, 12)
    rec[:, 12:14] = encode_temp(temp_k).astype("<u2").view(np.uint8).reshape(n, 2)
    rec[:, 14:16] = encode_pressure(pressure).astype("<u2").view(np.uint8).reshape(n, 2)
    rec[:, 16] = np.asarray(material, np.uint8)
    rec[:, 17] = ((np.asarray(body, np.uint8) & 0x0F) << 4) | (np.asarray(phase, np.uint8) & 0x0F)
    rec[:, 18] = encode_damage(damage)
    rec[:, 19] = encode_radius(mass, rho)
    return rec, encode_speed(vel)


def decode_records(rec):
    """Fields of (n, 20) records, decoded as tests/cpp/test_conservation.cpp:296-343 does."""
    rec = np.ascontiguousarray(rec, np.uint8)
    n = len(rec)
    xyz = rec[:, 0:12].copy().view("<f4").reshape(n, 3)
    t16 = rec[:, 12:14].copy().view("<u2").reshape(n).astype(np.float64)
    p16 = rec[:, 14:16].copy().view("<u2").reshape(n).astype(np.float64)
    return dict(
        xyz_km=xyz,
        temp_k=10.0 * np.exp(t16 / 65535.0 * LN20000),
        pressure_gpa=np.where(p16 == 0, 0.0, 1.0e-3 * np.exp((p16 - 1.0) / 65534.0 * LN1E7)),
        material=rec[:, 16],
        body=rec[:, 17] >> 4,
        phase=rec[:, 17] & 0x0F,
        damage=rec[:, 18] / 255.0,
        radius_km=np.exp(rec[:, 19] / 255.0 * LN2000),
    )


def decode_speed(speed):
    return 0.05 * np.exp(np.asarray(speed, np.float64) / 255.0 * LN2000)


def tracer_records(pos, tracer, rest, ocean, damage, temp_k, phase, div_v, h, living_com, sys_com):
    """40-byte records (count, 40) of the tracers.  ``tracer[i]`` = tracer index of particle i
    or -1; ``rest`` (count, 4) = rest radius and unit normal; float32 arithmetic as the kernel."""
    pos = _f32(pos)
    rest = _f32(rest).reshape(-1, 4)
    count = len(rest)
    out = np.zeros((count, 40), np.uint8)
    idx = np.flatnonzero(np.asarray(tracer) >= 0)
    t = np.asarray(tracer)[idx]
    d = pos[idx] - _f32(living_com)[None, :]
    r = np.sqrt((d * d).sum(axis=1, dtype=np.float32)).astype(np.float32)
    rr = rest[t, 0]
    disp = ((r - rr) * np.float32(1000.0)).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        nrm = d / r[:, None]
    dot = np.clip((nrm * rest[t, 1:4]).sum(axis=1, dtype=np.float32), -1.0, 1.0)
    nd = np.where(r > 1e-9, np.arccos(dot.astype(np.float64)), 0.0).astype(np.float32)
    dam = np.clip(_finite0(_f32(damage)[idx]), 0.0, 1.0).astype(np.float32)
    tk = np.maximum(np.float32(3.0), _finite0(_f32(temp_k)[idx])).astype(np.float32)
    ph = np.asarray(phase, np.uint8)[idx]
    oc = np.asarray(ocean, np.uint8)[t] != 0
    disrupted = (dam >= 0.85) | (ph == VAPOR) | (ph == PLASMA) | (np.abs(r - rr) > 0.5 * np.maximum(rr, np.float32(1e-3)))
    dv = np.maximum(np.float32(0.0), -_finite0(_f32(div_v)[idx])) * np.maximum(_f32(h)[idx], 0.0) * np.float32(1000.0)
    over = np.maximum(np.float32(0.0), np.float32(1.225 * 340.0) * dv * np.float32(1e-3)).astype(np.float32)
    tsu = np.where(oc, np.maximum(np.float32(0.0), _finite0(disp) * np.float32(1000.0)), 0.0).astype(np.float32)
    xyz = _finite0((pos[idx] - _f32(sys_com)[None, :]) * np.float32(1000.0)).astype(np.float32)
    head = (t.astype(np.uint32) | ((oc.astype(np.uint32) | (disrupted.astype(np.uint32) << 1)) << 16)).astype("<u4")
    cols = np.column_stack([xyz, disp, nd, tk, dam, over, tsu]).astype("<f4")
    out[t, 0:4] = head.view(np.uint8).reshape(len(t), 4)
    out[t, 4:40] = np.ascontiguousarray(cols).view(np.uint8).reshape(len(t), 36)
    return out
