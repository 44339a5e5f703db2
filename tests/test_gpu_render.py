"""Render frames on the device (astraea_pack_render, astraea_extract_tracers) against the numpy
restatement of packRenderRecordsKernel / extractTracersKernel (oracle/render.py), on the same
state: the device's own particles, downloaded in the same order.

Bars: positions (float32 km), material, body, phase and damage bytes equal; the log-quantised
temperature, pressure, radius and speed codes within one step of the float64 evaluation and
equal for at least 99 % of the particles; tracer records: index and flags equal, positions and
displacement within 1e-6 relative, normal change within 2e-3 rad (acos near 1 in float32), the
other fields equal to float32 rounding.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import render as R

pytestmark = pytest.mark.gpu

MAX_DT = 5.0 / 1.0e3


def _state(lib):
    from astraea_amd import scenario as S

    a = S.impact_pair(3000, 600, 45.0, 1.5, seed=3)
    b = S.impact_pair(2000, 400, 20.0, 2.0, seed=4)
    st = {k: np.ascontiguousarray(np.concatenate([a[k], b[k]])) for k in a}
    counts = [3600, 2400]
    rng = np.random.default_rng(5)
    st["damage"] = np.where(rng.random(6000) < 0.3, rng.random(6000), 0.0).astype(np.float32)
    ctx = lib.Context(capacity=6000, max_segments=2)
    ctx.set_params(S.plain_params())
    ctx.set_layout(counts)
    ctx.load_state(st)
    ctx.initial_forces()
    ctx.step(MAX_DT, 20, graph=True)
    ctx.synchronize()
    return ctx, counts


def _fields(ctx, names):
    return {k: ctx.download(k) for k in names}


def test_pack_render_matches_oracle(lib):
    ctx, counts = _state(lib)
    com = np.array([[0.1, -0.2, 0.3], [2.0, 1.0, -1.0]])
    rec, spd = ctx.pack_render(com, speed=True)
    rec2 = ctx.pack_render(com)
    assert np.array_equal(rec, rec2)  # the speed bytes change nothing else
    f = _fields(ctx, ["pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "mass", "rho", "temp_k", "pressure",
                      "material", "body", "phase", "damage"])
    seg = np.repeat([0, 1], counts)
    want, want_spd = np.zeros_like(rec), np.zeros_like(spd)
    for s in range(2):
        sel = seg == s
        pos = np.stack([f["pos_x"], f["pos_y"], f["pos_z"]], 1)[sel]
        vel = np.stack([f["vel_x"], f["vel_y"], f["vel_z"]], 1)[sel]
        want[sel], want_spd[sel] = R.pack_records(pos, vel, f["mass"][sel], f["rho"][sel], f["temp_k"][sel],
                                                  f["pressure"][sel], f["material"][sel], f["body"][sel],
                                                  f["phase"][sel], f["damage"][sel], com[s].astype(np.float32))
    exact = [slice(0, 12), 16, 17, 18]
    for c in exact:
        assert np.array_equal(rec[:, c], want[:, c]), c
    g, w = R.decode_records(rec), R.decode_records(want)
    for name, a, b in (("temp", rec[:, 12:14].copy().view("<u2"), want[:, 12:14].copy().view("<u2")),
                       ("pressure", rec[:, 14:16].copy().view("<u2"), want[:, 14:16].copy().view("<u2")),
                       ("radius", rec[:, 19], want[:, 19]), ("speed", spd, want_spd)):
        d = np.abs(a.astype(np.int64) - b.astype(np.int64)).ravel()
        print(name, "max code difference", d.max(), "equal fraction", np.mean(d == 0))
        assert d.max() <= 1 and np.mean(d == 0) >= 0.99, name
    assert g["body"].max() == 1 and g["material"].max() <= 4 and w["phase"].max() <= 5


def test_extract_tracers_matches_oracle(lib):
    ctx, counts = _state(lib)
    n = sum(counts)
    rng = np.random.default_rng(9)
    ids = rng.choice(n, 300, replace=False).astype(np.uint32)
    nrm = rng.normal(size=(300, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    rest = np.column_stack([rng.uniform(0.5, 8.0, 300), nrm]).astype(np.float32)
    ocean = (rng.random(300) < 0.4).astype(np.uint8)
    ctx.set_tracers(ids, rest, ocean)
    assert np.count_nonzero(ctx.download("tracer")) == 300
    living = np.array([[0.0, 0.1, 0.0], [1.0, 0.0, 0.5]])
    sysc = np.array([[0.2, 0.0, 0.0], [0.0, 0.0, 0.0]])
    out = ctx.extract_tracers(living, sysc)
    f = _fields(ctx, ["pos_x", "pos_y", "pos_z", "damage", "temp_k", "phase", "div_v", "h", "original_id"])
    oid = f["original_id"]
    tracer = np.full(n, -1)
    where = {int(p): t for t, p in enumerate(ids)}
    for i in range(n):
        tracer[i] = where.get(int(oid[i]), -1)
    seg = np.repeat([0, 1], counts)
    want = np.zeros_like(out)
    pos = np.stack([f["pos_x"], f["pos_y"], f["pos_z"]], 1)
    for s in range(2):
        sel = seg == s
        tr = np.where(sel, tracer, -1)
        w = R.tracer_records(pos, tr, rest, ocean, f["damage"], f["temp_k"], f["phase"], f["div_v"], f["h"],
                             living[s].astype(np.float32), sysc[s].astype(np.float32))
        hit = np.unique(tr[tr >= 0])
        want[hit] = w[hit]
    assert np.array_equal(out[:, :4], want[:, :4])
    a = out[:, 4:].copy().view("<f4").astype(np.float64)
    b = want[:, 4:].copy().view("<f4").astype(np.float64)
    np.testing.assert_allclose(a[:, [0, 1, 2, 3]], b[:, [0, 1, 2, 3]], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(a[:, 4], b[:, 4], atol=2e-3)
    np.testing.assert_allclose(a[:, 5:], b[:, 5:], rtol=1e-6, atol=1e-6)
    # a new layout clears the tracers
    ctx.set_tracers([], np.zeros((0, 4)), [])
    assert not np.any(ctx.download("tracer"))


def test_render_frame_ac12_and_pcls():
    """tests/cpp/test_conservation.cpp:260-369 (AC12) through the public Simulation API, and the
    PCLS frame of it (server/binary_protocol.cpp) with and without zlib."""
    exe = Path(__file__).resolve().parent / "cpp" / "test_simulation"
    r = subprocess.run([str(exe), "render"], capture_output=True, text=True, timeout=110)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
