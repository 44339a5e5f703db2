"""The allocation-free render entry point (astraea_pack_render_noalloc) against astraea_pack_render.

Both pack the same resident state, so the 20-byte records and the speed bytes must be equal byte
for byte.  Two calls in a row on one GPU, the second with a smaller input (fewer particles and
fewer segments, in another context), then the first context again: the device table the entry
point shares between calls keeps the last call's contents, and no call may read a stale entry.
The caller's device buffers are filled with a marker first, so a byte left unwritten shows.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_DT = 5.0 / 1.0e3


class _Hip:
    """The few HIP runtime calls the test needs to own device buffers (the caller's memory)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self._ok(self.lib.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p.value

    def fill(self, ptr, byte, nbytes):
        self._ok(self.lib.hipMemset(C.c_void_p(ptr), byte, nbytes), "hipMemset")
        self._ok(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def download(self, ptr, nbytes):
        out = np.empty(nbytes, np.uint8)
        self._ok(self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), nbytes, 2), "hipMemcpy")
        return out

    def free(self, ptr):
        self._ok(self.lib.hipFree(C.c_void_p(ptr)), "hipFree")


def _context(lib, parts, steps):
    from astraea_amd import scenario as S

    sts = [S.impact_pair(nt, ni, ang, vr, seed=seed) for nt, ni, ang, vr, seed in parts]
    st = {k: np.ascontiguousarray(np.concatenate([s[k] for s in sts])) for k in sts[0]}
    counts = [len(s["mass"]) for s in sts]
    ctx = lib.Context(int(sum(counts)), max_segments=len(counts))
    ctx.set_params(S.plain_params())
    ctx.set_layout(counts)
    ctx.load_state(st)
    ctx.initial_forces()
    ctx.step(MAX_DT, steps)
    ctx.synchronize()
    return ctx, counts


def _check_equal(hip, ctx, com):
    n = ctx.n
    rec_dev, spd_dev = hip.malloc(20 * n), hip.malloc(n)
    try:
        hip.fill(rec_dev, 0xAB, 20 * n)
        hip.fill(spd_dev, 0xAB, n)
        ctx.pack_render_noalloc(com, rec_dev, spd_dev)
        got_rec = hip.download(rec_dev, 20 * n).reshape(n, 20)
        got_spd = hip.download(spd_dev, n)
    finally:
        hip.free(rec_dev)
        hip.free(spd_dev)
    want_rec, want_spd = ctx.pack_render(com, speed=True)
    assert np.array_equal(got_rec, want_rec)
    assert np.array_equal(got_spd, want_spd)


def test_pack_render_noalloc_matches_pack_render(lib):
    print("no-alloc render: 3 segments then 1 smaller, then the first again", flush=True)
    hip = _Hip()
    a, counts_a = _context(lib, [(3000, 600, 45.0, 1.5, 3), (1200, 300, 20.0, 2.0, 4), (900, 150, 60.0, 1.1, 5)], 6)
    com_a = np.array([[0.01, -0.02, 0.03], [0.5, 0.1, -0.2], [-1.0, 0.0, 0.25]])
    _check_equal(hip, a, com_a)
    b, counts_b = _context(lib, [(1000, 250, 30.0, 1.2, 6)], 4)
    assert sum(counts_b) < sum(counts_a) and len(counts_b) < len(counts_a)
    _check_equal(hip, b, np.array([[0.2, 0.3, -0.1]]))
    _check_equal(hip, a, com_a + 0.125)


def test_pack_render_noalloc_rejects_a_null_output(lib):
    a, _ = _context(lib, [(800, 200, 45.0, 1.5, 7)], 2)
    with pytest.raises(lib.AstraeaError) as e:
        a.pack_render_noalloc([0.0, 0.0, 0.0], 0)
    assert e.value.code == -1
