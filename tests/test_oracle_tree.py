"""Pins the oracle's integer stages with independent constructions.

- Morton keys: FP32 quantisation (engine/sph_gravity.cu:151-171, 347-356) followed by a
  bit-by-bit interleave written here (not the magic-mask spread the oracle uses).
- Sort permutation: numpy's stable argsort of the keys (CUB SortPairs over bits [0,63) is
  stable, engine/sph_gravity.cu:2028-2046).
- Karras tree: a top-down radix split written here (node covering [f, l] splits after the
  last index whose common prefix with f is longer than the node's), which must give the
  reference's bottom-up arrays (engine/sph_gravity.cu:478-528).
- Node moments: root mass and centre of mass against FP64 sums, boxes contain children.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
from independent import karras as karras_ref  # noqa: E402
from independent import morton as morton_ref_g  # noqa: E402


def interleave(qx, qy, qz):
    k = 0
    for b in range(21):
        k |= ((qx >> b) & 1) << (3 * b + 2)
        k |= ((qy >> b) & 1) << (3 * b + 1)
        k |= ((qz >> b) & 1) << (3 * b)
    return k


def morton_ref(x, y, z, b):
    f32 = np.float32
    out = []
    for v, mn, mx in ((x, b[0], b[3]), (y, b[1], b[4]), (z, b[2], b[5])):
        span = f32(f32(mx) - f32(mn))
        inv = f32(f32(2097151.0) / span) if span > f32(1e-6) else f32(0.0)
        q = f32(f32(f32(v) - f32(mn)) * inv)
        q = min(max(q, f32(0.0)), f32(2097151.0))
        out.append(int(q))
    return interleave(*out)


def delta(keys, i, j):
    n = len(keys)
    if j < 0 or j >= n:
        return -1
    a, b = int(keys[i]), int(keys[j])
    if a != b:
        return 64 - (a ^ b).bit_length()
    x = i ^ j
    return 64 + (32 - x.bit_length())


def karras_top_down(keys):
    """left/right child arrays (reference encoding: leaf j as ~j) by recursive radix splits."""
    n = len(keys)
    left = np.zeros(max(n - 1, 0), np.int32)
    right = np.zeros(max(n - 1, 0), np.int32)
    if n < 2:
        return left, right
    stack = [(0, n - 1, 0)]
    while stack:
        f, l, idx = stack.pop()
        dn = delta(keys, f, l)
        g = f
        for s in range(f, l):
            if delta(keys, f, s) > dn:
                g = s
        left[idx] = ~g if g == f else g
        right[idx] = ~(g + 1) if g + 1 == l else g + 1
        if g != f:
            stack.append((f, g, g))
        if g + 1 != l:
            stack.append((g + 1, l, g + 1))
    return left, right


def test_morton_corner_cases(orc):
    b = [0.0, 0.0, 0.0, 1.0, 1.0, 1.0]
    assert orc.morton_key(0.0, 0.0, 0.0, b) == 0
    assert orc.morton_key(1.0, 1.0, 1.0, b) == 0x7FFFFFFFFFFFFFFF
    assert orc.morton_key(1.0, 0.0, 0.0, b) == 0x4924924924924924
    assert orc.morton_key(0.0, 1.0, 0.0, b) == 0x2492492492492492
    assert orc.morton_key(0.0, 0.0, 1.0, b) == 0x1249249249249249
    # outside the box clamps; a zero-span axis quantises to 0
    assert orc.morton_key(-5.0, 7.0, 0.5, b) == orc.morton_key(0.0, 1.0, 0.5, b)
    flat = [0.0, 3.0, 0.0, 1.0, 3.0, 1.0]
    assert orc.morton_key(0.3, 3.0, 0.7, flat) == morton_ref(0.3, 3.0, 0.7, flat)
    assert (orc.morton_key(0.3, 3.0, 0.7, flat) >> 1) & 0x1249249249249249 == 0


def test_morton_random_vs_bit_interleave(orc):
    rng = np.random.default_rng(11)
    for _ in range(2000):
        lo = rng.normal(0, 10, 3).astype(np.float32)
        hi = (lo + np.abs(rng.normal(0, 10, 3))).astype(np.float32)
        b = [*lo, *hi]
        p = (lo + rng.random(3) * (hi - lo)).astype(np.float32)
        assert orc.morton_key(*p, b) == morton_ref(*p, b)


def golden_cases():
    return json.loads((GOLDEN / "tree_cases.json").read_text())["cases"]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c["name"])
def test_tree_against_golden(orc, case):
    """Committed fixtures (inputs + expected keys/perm/children from the independent
    constructions above, written by tests/golden/make_tree_cases.py)."""
    n = len(case["x"])
    st = {"pos_x": np.array(case["x"], np.float32), "pos_y": np.array(case["y"], np.float32),
          "pos_z": np.array(case["z"], np.float32), "mass": np.array(case["m"], np.float32),
          "h": np.full(n, 0.05, np.float32), "h0": np.full(n, 0.05, np.float32)}
    o = orc.Oracle(orc.default_params(), 64)
    o.set_layout([n])
    o.load_state(st)
    o.sort_and_build()
    t = o.tree(0)
    assert [int(k) for k in t["keys"]] == [int(k) for k in case["keys"]]
    assert all(morton_ref_g(case["x"][case["perm"][k]], case["y"][case["perm"][k]], case["z"][case["perm"][k]], [min(case["x"]), min(case["y"]), min(case["z"]), max(case["x"]), max(case["y"]), max(case["z"])]) == int(case["keys"][k]) for k in range(n))
    assert t["perm"].tolist() == case["perm"]
    assert t["left"].tolist() == case["left"]
    assert t["right"].tolist() == case["right"]


@pytest.mark.parametrize("seed,n,dup", [(1, 300, 0), (2, 257, 64), (3, 100, 99), (4, 2, 0), (5, 513, 128)])
def test_sort_and_karras_vs_independent(orc, seed, n, dup):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 3)).astype(np.float32)
    if dup:
        x[:dup] = x[0]
    st = {"pos_x": x[:, 0].copy(), "pos_y": x[:, 1].copy(), "pos_z": x[:, 2].copy(),
          "mass": rng.uniform(0.5, 2, n).astype(np.float32), "h": np.full(n, 0.1, np.float32),
          "h0": np.full(n, 0.1, np.float32)}
    o = orc.Oracle(orc.default_params(), 64)
    o.set_layout([n])
    o.load_state(st)
    o.sort_and_build()
    t = o.tree(0)
    b = [x[:, 0].min(), x[:, 1].min(), x[:, 2].min(), x[:, 0].max(), x[:, 1].max(), x[:, 2].max()]
    keys = np.array([morton_ref(*x[i], b) for i in range(n)], dtype=np.uint64)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(t["perm"], perm)
    assert np.array_equal(t["keys"], keys[perm])
    left, right = karras_top_down(t["keys"])
    l2, r2 = karras_ref(t["keys"])
    assert left.tolist() == l2 and right.tolist() == r2
    assert np.array_equal(t["left"], left)
    assert np.array_equal(t["right"], right)
    # every node but the root has exactly one parent; leaves partition [0, n)
    if n >= 2:
        children = np.concatenate([t["left"], t["right"]])
        internal = children[children >= 0]
        leaves = ~children[children < 0]
        assert sorted(leaves.tolist()) == list(range(n))
        assert sorted(internal.tolist()) == list(range(1, n - 1))
        assert t["parent_internal"][0] == -1


def test_moments_invariants(orc):
    rng = np.random.default_rng(9)
    n = 2000
    x = rng.normal(0, 1, (n, 3)).astype(np.float32)
    m = rng.uniform(0.5, 2, n).astype(np.float32)
    h = rng.uniform(0.01, 0.1, n).astype(np.float32)
    st = {"pos_x": x[:, 0].copy(), "pos_y": x[:, 1].copy(), "pos_z": x[:, 2].copy(), "mass": m, "h": h, "h0": h}
    o = orc.Oracle(orc.default_params(), 64)
    o.set_layout([n])
    o.load_state(st)
    o.sort_and_build()
    t = o.tree(0)
    xd = x.astype(np.float64)
    md = m.astype(np.float64)
    assert abs(t["node_cm"][0, 3] - md.sum()) <= 1e-5 * md.sum()
    com = (md[:, None] * xd).sum(0) / md.sum()
    assert np.allclose(t["node_cm"][0, :3], com, atol=1e-5)
    assert np.allclose(t["node_min"][0, :3], x.min(0)) and np.allclose(t["node_max"][0, :3], x.max(0))
    # extent = max side; support bound = max over leaves of 2 max(1.08 h, 0.35 h0)
    ext = (t["node_max"][:, :3] - t["node_min"][:, :3]).max(1)
    assert np.allclose(t["node_min"][:, 3], ext)
    hb = 2.0 * np.maximum(1.08 * h.astype(np.float64), 0.35 * h.astype(np.float64))
    assert abs(t["node_max"][0, 3] - hb.max()) <= 1e-6 * hb.max()
    # children boxes inside parents
    for p in range(n - 1):
        for ch in (t["left"][p], t["right"][p]):
            if ch >= 0:
                assert np.all(t["node_min"][ch, :3] >= t["node_min"][p, :3])
                assert np.all(t["node_max"][ch, :3] <= t["node_max"][p, :3])


def test_morton_fp64_box(orc):
    """The mixed form's keys (FP64 positions, FP64 box), pinned by arithmetic: a box 1e4 from the
    origin with unit span, where FP32 loses ten of the 21 bits.  x = 1e4 + 1e-3 quantises to
    floor(1e-3 * 2097151) = 2097 in FP64; FP32 gives 2047 (1e4 + 1e-3 rounds to 1e4 + 2^-10),
    so the FP32 key must differ.  Axes, clamping and the far corner as in the FP32 key."""
    lo = 1.0e4
    b = (lo, lo, lo, lo + 1.0, lo + 1.0, lo + 1.0)
    assert orc.morton_key64(lo + 1e-3, lo, lo, b) == interleave(2097, 0, 0)
    assert orc.morton_key64(lo, lo + 1e-3, lo, b) == interleave(0, 2097, 0)
    assert orc.morton_key64(lo, lo, lo + 0.5, b) == interleave(0, 0, 1048575)  # floor(2097151 / 2)
    assert orc.morton_key64(lo + 1.0, lo + 1.0, lo + 1.0, b) == 0x7FFFFFFFFFFFFFFF
    assert orc.morton_key64(lo - 5.0, lo + 7.0, lo, b) == interleave(0, 2097151, 0)
    assert orc.morton_key(lo + 1e-3, lo, lo, b) == interleave(2047, 0, 0)
    # a zero-span axis quantises to 0
    flat = (lo, 0.0, lo, lo + 1.0, 0.0, lo + 1.0)
    assert orc.morton_key64(lo + 0.25, 0.0, lo, flat) == interleave(524287, 0, 0)
