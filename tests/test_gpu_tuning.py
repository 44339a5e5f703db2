"""Launch variants and the autotuner (astraea_set_tuning, astraea_autotune).

A variant changes how many union records a neighbour pass stages in LDS per round, the sort's digit
width, or how many wavefronts share a block and how deep their stacks are in the gravity walk and
the neighbour build, never the order in which a particle's pairs or records are summed, so every
variant must give bit-identical states; the autotuner's steps are ordinary steps, and its extra
gravity and build launches rewrite what the last step computed, so the tuned run must equal the
same number of untuned steps.
"""
import numpy as np
import pytest

from test_gpu_parity import MAX_DT, perturbed_small_state

pytestmark = pytest.mark.gpu

FIELDS = ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "u", "rho", "h", "damage")


def _context(lib, st, strength):
    from astraea_amd import scenario as S

    n = len(st["mass"])
    ctx = lib.Context(n)
    if strength:
        ctx.set_params(S.plain_params(strength_mask=(1 << S.GRANITE) | (1 << S.IRON)))
    else:
        ctx.set_params(S.plain_params())
    ctx.set_layout([n])
    ctx.load_state(st)
    ctx.initial_forces()
    return ctx


def _state(ctx):
    ctx.synchronize()
    return {k: ctx.download(k) for k in FIELDS}


@pytest.mark.parametrize("strength", [False, True])
def test_variants_are_bit_identical(lib, strength):
    st = perturbed_small_state(strength=strength)
    want = None
    for t in [(256, 128, 128, 8, 4, 1024, 4, 1024), (128, 160, 160, 9, 2, 1024, 8, 512),
              (256, 160, 192, 9, 8, 512, 2, 1024), (128, 128, 256, 8, 4, 512, 4, 512)]:
        ctx = _context(lib, st, strength)
        ctx.set_tuning(*t)
        assert tuple(ctx.tuning().values()) == t
        ctx.step(MAX_DT, 3, graph=True)
        got = _state(ctx)
        if want is None:
            want = got
            continue
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), (t, k)


def test_round_sizes_with_several_rounds_per_bank(lib):
    """Smoothing lengths 1.7 times the state's (about 490 table entries a particle, four times as
    many, so particles spill past the regular region too): a wavefront's union then spans several
    rounds per bank at every round size, including the last, partly filled staging load of a
    160-record round, and every round size gives the same state."""
    st = perturbed_small_state(seed=3)
    st["h"] = (st["h"] * 1.7).astype(np.float32)
    want = None
    for t in [(128, 128, 128, 8, 4, 1024, 4, 1024), (256, 160, 160, 8, 4, 1024, 4, 1024),
              (256, 128, 192, 8, 4, 1024, 4, 1024), (128, 160, 256, 8, 4, 1024, 4, 1024)]:
        ctx = _context(lib, st, False)
        ctx.set_tuning(*t)
        entries = ctx.counters()["nbr_entries"] / len(st["mass"])
        print(t, "entries per particle", entries)
        assert entries > 250.0, entries  # several hundred union members per bank
        ctx.step(MAX_DT, 2, graph=False)
        got = _state(ctx)
        if want is None:
            want = got
            continue
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), (t, k)


def test_autotune_steps_are_ordinary_steps(lib):
    st = perturbed_small_state()
    tuned = _context(lib, st, False)
    r = tuned.autotune(MAX_DT, steps_per_variant=1)
    assert r["chosen"]["pass1_round"] in (128, 256) and r["chosen"]["pass2_round"] in (128, 160)
    assert r["chosen"]["drift_round"] in (128, 160, 192, 256) and r["chosen"]["sort_bits"] in (8, 9)
    assert (r["chosen"]["gravity_waves"], r["chosen"]["gravity_stack"]) in lib.WALK_VARIANTS
    assert (r["chosen"]["build_waves"], r["chosen"]["build_stack"]) in lib.WALK_VARIANTS
    assert tuned.tuning() == r["chosen"]
    assert all(v > 0.0 for d in ("pass1_ms", "pass2_ms", "drift_ms", "sort_ms") for v in r[d].values())
    # a walk variant whose stack filled reads -1; the chosen ones were timed
    assert all(v > 0.0 or v == -1.0 for d in ("gravity_ms", "build_ms") for v in r[d].values())
    g = f"{r['chosen']['gravity_waves']}x{r['chosen']['gravity_stack']}"
    b = f"{r['chosen']['build_waves']}x{r['chosen']['build_stack']}"
    assert r["gravity_ms"][g] > 0.0 and r["build_ms"][b] > 0.0
    plain = _context(lib, st, False)
    plain.step(MAX_DT, 12, graph=False)  # four untimed warm-up steps and eight timed
    a, b = _state(tuned), _state(plain)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


def test_invalid_tuning_is_rejected(lib):
    ctx = _context(lib, perturbed_small_state(), False)
    with pytest.raises(lib.AstraeaError):
        ctx.set_tuning(512, 128, 128)
    with pytest.raises(lib.AstraeaError):
        ctx.set_tuning(256, 128, 128, 11)
    with pytest.raises(lib.AstraeaError):
        ctx.set_tuning(256, 128, 128, 8, gravity_waves=8, gravity_stack=1024)  # 80 KB of LDS per block
    with pytest.raises(lib.AstraeaError):
        ctx.set_tuning(256, 128, 128, 8, build_waves=3)
    assert ctx.tuning() == {"pass1_round": 256, "pass2_round": 160, "drift_round": 160, "sort_bits": 8,
                            "gravity_waves": 4, "gravity_stack": 1024, "build_waves": 4, "build_stack": 1024}
