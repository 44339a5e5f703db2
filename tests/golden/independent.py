"""Independent pure-Python constructions of the integer stages (used to write and check fixtures).

Morton: FP32 quantisation as engine/sph_gravity.cu:335-361, then a bit-by-bit interleave.
Sort: stable sort by key (CUB SortPairs contract, engine/sph_gravity.cu:2028-2046).
Karras: top-down radix split with the reference's tie-break delta(i,j) = 64 + clz32(i^j)
(engine/sph_gravity.cu:173-187), child encoding ~leaf (engine/sph_gravity.cu:508-528).
"""
import numpy as np


def interleave(qx, qy, qz):
    k = 0
    for b in range(21):
        k |= ((qx >> b) & 1) << (3 * b + 2)
        k |= ((qy >> b) & 1) << (3 * b + 1)
        k |= ((qz >> b) & 1) << (3 * b)
    return k


def bounds(x, y, z):
    f = np.float32
    return [f(np.min(x)), f(np.min(y)), f(np.min(z)), f(np.max(x)), f(np.max(y)), f(np.max(z))]


def morton(x, y, z, b):
    f32 = np.float32
    q = []
    for v, mn, mx in ((x, b[0], b[3]), (y, b[1], b[4]), (z, b[2], b[5])):
        span = f32(f32(mx) - f32(mn))
        inv = f32(f32(2097151.0) / span) if span > f32(1e-6) else f32(0.0)
        t = f32(f32(f32(v) - f32(mn)) * inv)
        t = min(max(t, f32(0.0)), f32(2097151.0))
        q.append(int(t))
    return interleave(*q)


def delta(keys, i, j):
    n = len(keys)
    if j < 0 or j >= n:
        return -1
    a, b = int(keys[i]), int(keys[j])
    if a != b:
        return 64 - (a ^ b).bit_length()
    return 64 + (32 - (i ^ j).bit_length())


def karras(keys):
    n = len(keys)
    left = [0] * max(n - 1, 0)
    right = [0] * max(n - 1, 0)
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


def tree_case(name, x, y, z, m):
    x, y, z, m = (np.asarray(a, np.float32) for a in (x, y, z, m))
    b = bounds(x, y, z)
    keys = [morton(x[i], y[i], z[i], b) for i in range(len(x))]
    perm = sorted(range(len(keys)), key=lambda i: keys[i])  # Python's sort is stable
    sk = [keys[i] for i in perm]
    left, right = karras(sk)
    return {"name": name, "x": x.tolist(), "y": y.tolist(), "z": z.tolist(), "m": m.tolist(),
            "keys": [str(k) for k in sk], "perm": perm, "left": left, "right": right}
