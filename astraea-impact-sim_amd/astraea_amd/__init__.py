"""ctypes binding of libastraea_hip.so, the MI355X (gfx950) per-step physics library.

The C ABI (include/astraea_hip.h) replaces the reference's CUDA launch layer
(engine/sph_gravity.cuh:61-92) under `astraea::Simulation` (engine/simulation.hpp).
This module mirrors it for Python callers: `Context` owns one device context,
moves per-particle fields as numpy arrays and runs the stage entry points or
whole KDK steps.  `scenario` builds the benchmark bodies on the host.

There is no CPU fallback: if the shared library is missing this module raises
on import.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libastraea_hip.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `make -C astraea-impact-sim_amd/csrc`"
    )
_lib = C.CDLL(str(LIB_PATH))

# ---- constants mirrored from include/astraea_hip.h ----
OK = 0
E_INVALID = -1
PREC_FP32 = 0
PREC_FP64_ACCUM = 1
GROUP_SIZE = 64
SEGMENT_ALIGN = 256
STACK_DEPTH = 192
ENERGY_STRIDE = 18
EN = dict(mass=0, kinetic=1, potential=2, internal=3, mom_x=4, mom_y=5, mom_z=6, ang_x=7, ang_y=8, ang_z=9,
          radiated=10, rad_power_w=11, max_temp=12, max_pressure=13, com_x=14, com_y=15, com_z=16, sim_time=17)
CLASS_STRIDE = 36
CL = dict(body0_mass=0, body0_com=slice(1, 4), body0_vel=slice(4, 7), body1_mass=7, body1_com=slice(8, 11),
          body1_vel=slice(11, 14), seed=slice(14, 20), rem_mass=20, rem_pos=slice(21, 24), rem_vel=slice(24, 27),
          rem_radius=27, bound_planet=28, disk=29, ejecta=30, impactor_bound=31, spin_l=32, spin_i=33,
          max_damage=34, iterations=35)
STAGES = ["dt", "kick1", "drift", "morton", "sort", "permute_tree", "moments", "neighbours", "pass1", "gravity",
          "pass2", "kick2"]

_FIELD_NAMES = [
    "POS_X", "POS_Y", "POS_Z", "VEL_X", "VEL_Y", "VEL_Z", "MASS", "H", "H0", "RHO", "U",
    "SXX", "SYY", "SZZ", "SXY", "SXZ", "SYZ", "DAMAGE", "DAMAGE_CBRT", "EPS_FAIL",
    "MATERIAL", "BODY", "TRACER", "ORIGINAL_ID",
    "POS64_X", "POS64_Y", "POS64_Z", "VEL64_X", "VEL64_Y", "VEL64_Z",
    "ACC_X", "ACC_Y", "ACC_Z", "ACC_GRAV_X", "ACC_GRAV_Y", "ACC_GRAV_Z", "PHI",
    "DRHO_DT", "DU_DT", "PRESSURE", "SOUND_SPEED", "TEMP_K",
    "SHEPARD", "SURF_FACTOR", "DIV_V", "CURL_X", "CURL_Y", "CURL_Z",
    "DS_XX", "DS_YY", "DS_ZZ", "DS_XY", "DS_XZ", "DS_YZ",
    "DDAMAGE_CBRT_DT", "RAD_POWER_W", "PHASE", "NGB_COUNT",
]
FIELD = {name.lower(): k for k, name in enumerate(_FIELD_NAMES)}
_U8 = {"material", "body", "tracer", "phase"}
_U32 = {"original_id", "ngb_count"}
_F64 = {"pos64_x", "pos64_y", "pos64_z", "vel64_x", "vel64_y", "vel64_z"}


def field_dtype(name: str):
    if name in _U8:
        return np.uint8
    if name in _U32:
        return np.uint32
    if name in _F64:
        return np.float64
    return np.float32


class Params(C.Structure):
    """astraea_params (include/astraea_hip.h)."""

    _fields_ = [
        ("theta", C.c_float),
        ("soft_factor", C.c_float),
        ("cfl_factor", C.c_float),
        ("gamma_hhe", C.c_float),
        ("cooling_min_temp_k", C.c_float),
        ("strength_mask", C.c_uint32),
        ("precision", C.c_uint32),
        ("group_size", C.c_uint32),
        ("reserved", C.c_uint32 * 4),
    ]


MAX_LAYERS = 4


class BodySpec(C.Structure):
    """astraea_body_spec (include/astraea_hip.h)."""

    _fields_ = [
        ("radius_m", C.c_double),
        ("num_layers", C.c_uint32),
        ("layer_material", C.c_uint8 * MAX_LAYERS),
        ("layer_mass_frac", C.c_double * MAX_LAYERS),
        ("layer_density_kgm3", C.c_double * MAX_LAYERS),
        ("h_over_spacing", C.c_double),
        ("jitter", C.c_double),
        ("surface_temp_k", C.c_double),
        ("strength_flaws", C.c_uint32),
        ("weibull_m", C.c_double),
        ("weibull_k", C.c_double),
        ("seed", C.c_uint64),
    ]


class PlanetSpec(C.Structure):
    _fields_ = [
        ("mass_kg", C.c_double), ("num_layers", C.c_uint32), ("layer_material", C.c_uint8 * MAX_LAYERS),
        ("layer_mass_frac", C.c_double * MAX_LAYERS), ("surface_temp_k", C.c_double), ("core_temp_k", C.c_double),
        ("surface_pressure_pa", C.c_double), ("gamma_hhe", C.c_double), ("h_factor", C.c_double),
        ("seed", C.c_uint64),
    ]


class PlanetInfo(C.Structure):
    _fields_ = [("mass_kg", C.c_double), ("radius_m", C.c_double), ("central_pressure_pa", C.c_double),
                ("profile_samples", C.c_uint32)]


class Tuning(C.Structure):
    """astraea_tuning: union records staged per round by pass 1, pass 2 and the drift, the radix
    sort's digit width, and the wavefronts per block and stack entries of the gravity walk and of
    the neighbour build."""
    _fields_ = [("pass1_round", C.c_uint32), ("pass2_round", C.c_uint32), ("drift_round", C.c_uint32),
                ("sort_bits", C.c_uint32), ("gravity_waves", C.c_uint32), ("gravity_stack", C.c_uint32),
                ("build_waves", C.c_uint32), ("build_stack", C.c_uint32)]


# the (wavefronts per block, stack entries) variants of the gravity walk and the neighbour build
WALK_VARIANTS = ((2, 1024), (4, 1024), (4, 512), (8, 512))


class BodyOut(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_float)) for n in ("x", "y", "z", "m", "h", "rho", "u", "eps_fail")] + [
        ("mat", C.POINTER(C.c_uint8))
    ]


_P = C.c_void_p
_sig = {
    "astraea_last_error": (C.c_char_p, []),
    "astraea_abi_version": (C.c_int, []),
    "astraea_field_size": (C.c_size_t, [C.c_int]),
    "astraea_default_params": (None, [C.POINTER(Params)]),
    "astraea_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_uint32, C.c_uint32]),
    "astraea_destroy": (C.c_int, [_P]),
    "astraea_stream": (_P, [_P]),
    "astraea_set_params": (C.c_int, [_P, C.POINTER(Params)]),
    "astraea_get_params": (C.c_int, [_P, C.POINTER(Params)]),
    "astraea_set_layout": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "astraea_num_particles": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "astraea_upload_field": (C.c_int, [_P, C.c_int, _P]),
    "astraea_download_field": (C.c_int, [_P, C.c_int, _P]),
    "astraea_clear_stats": (C.c_int, [_P]),
    "astraea_sort_and_build_bvh": (C.c_int, [_P]),
    "astraea_sph_pass1": (C.c_int, [_P]),
    "astraea_evaluate_eos": (C.c_int, [_P]),
    "astraea_gravity": (C.c_int, [_P]),
    "astraea_sph_pass2": (C.c_int, [_P]),
    "astraea_reduce_dt": (C.c_int, [_P, C.c_float, C.POINTER(C.c_float)]),
    "astraea_kick_half": (C.c_int, [_P, C.c_int]),
    "astraea_drift": (C.c_int, [_P]),
    "astraea_evaluate_initial_forces": (C.c_int, [_P]),
    "astraea_step": (C.c_int, [_P, C.c_float, C.c_uint32, C.c_int]),
    "astraea_synchronize": (C.c_int, [_P]),
    "astraea_get_dt": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "astraea_energy": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "astraea_classify": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "astraea_pack_render": (C.c_int, [_P, _P, _P, _P]),
    "astraea_pack_render_noalloc": (C.c_int, [_P, _P, _P, _P]),
    "astraea_set_tracers": (C.c_int, [_P, C.c_uint32, _P, _P, _P]),
    "astraea_extract_tracers": (C.c_int, [_P, _P, _P, _P]),
    "astraea_select_tracers": (C.c_int, [C.c_uint32] + [_P] * 5 + [C.c_uint32] + [_P] * 5),
    "astraea_cube_face_uv_to_normal": (None, [C.c_uint32, C.c_float, C.c_float, _P, _P, _P]),
    "astraea_normal_to_cube_face_uv": (None, [C.c_float, C.c_float, C.c_float, _P, _P, _P]),
    "astraea_relax": (C.c_int, [_P, C.c_uint32, _P, _P, _P]),
    "astraea_counters": (C.c_int, [_P, C.POINTER(C.c_uint64), C.c_int]),
    "astraea_set_neighbour_stride": (C.c_int, [_P, C.c_uint32]),
    "astraea_download_tree": (C.c_int, [_P, C.c_uint32] + [_P] * 9),
    "astraea_enable_stage_timing": (C.c_int, [_P, C.c_int]),
    "astraea_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "astraea_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "astraea_autotune": (C.c_int, [_P, C.c_float, C.c_uint32, C.POINTER(Tuning), C.POINTER(C.c_double)]),
    "astraea_stage_times": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "astraea_default_body_spec": (None, [C.POINTER(BodySpec)]),
    "astraea_default_planet_spec": (None, [C.POINTER(PlanetSpec)]),
    "astraea_build_planet": (C.c_int, [C.POINTER(PlanetSpec), C.c_uint32, C.POINTER(BodyOut), _P,
                                       C.POINTER(PlanetInfo)]),
    "astraea_planet_profile": (C.c_int, [C.POINTER(PlanetSpec), C.c_uint32] + [_P] * 6 + [C.POINTER(PlanetInfo)]),
    "astraea_body_radius_for_mass": (C.c_int, [C.POINTER(BodySpec), C.c_double, C.POINTER(C.c_double)]),
    "astraea_build_body": (C.c_int, [C.POINTER(BodySpec), C.c_uint32, C.POINTER(BodyOut), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "astraea_place_two_bodies": (C.c_int, [C.c_uint32, C.c_uint32] + [_P] * 7 +
                                 [C.c_double] * 5 + [C.POINTER(C.c_double)]),
    "astraea_splitmix64": (C.c_uint64, [C.c_uint64]),
}
for _name, (_res, _args) in _sig.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED_SYMBOLS = sorted(_sig)


class AstraeaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"astraea error {code}: {msg}")
        self.code = code


def _check(rc: int):
    if rc != OK:
        raise AstraeaError(rc, _lib.astraea_last_error().decode())


def default_params(**overrides) -> Params:
    p = Params()
    _lib.astraea_default_params(C.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One device context (astraea_create)."""

    def __init__(self, capacity: int, max_segments: int = 1, device: int = 0):
        h = C.c_void_p()
        _check(_lib.astraea_create(C.byref(h), device, capacity, max_segments))
        self._h = h
        self.counts: list[int] = []

    def close(self):
        if self._h:
            _lib.astraea_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return _lib.astraea_stream(self._h) or 0

    @property
    def n(self) -> int:
        return int(sum(self.counts))

    def set_params(self, p: Params):
        _check(_lib.astraea_set_params(self._h, C.byref(p)))

    def params(self) -> Params:
        p = Params()
        _check(_lib.astraea_get_params(self._h, C.byref(p)))
        return p

    def set_layout(self, counts):
        arr = (C.c_uint32 * len(counts))(*[int(x) for x in counts])
        _check(_lib.astraea_set_layout(self._h, len(counts), arr))
        self.counts = [int(x) for x in counts]

    def upload(self, name: str, values):
        a = np.ascontiguousarray(values, dtype=field_dtype(name))
        if a.shape != (self.n,):
            raise ValueError(f"{name}: expected {self.n} values, got {a.shape}")
        _check(_lib.astraea_upload_field(self._h, FIELD[name], _ptr(a)))

    def download(self, name: str) -> np.ndarray:
        a = np.empty(self.n, dtype=field_dtype(name))
        _check(_lib.astraea_download_field(self._h, FIELD[name], _ptr(a)))
        return a

    def load_state(self, state: dict):
        for k, v in state.items():
            self.upload(k, v)

    # stage entry points (reference launch layer)
    def clear_stats(self):
        _check(_lib.astraea_clear_stats(self._h))

    def sort_and_build(self):
        _check(_lib.astraea_sort_and_build_bvh(self._h))

    def pass1(self):
        _check(_lib.astraea_sph_pass1(self._h))

    def eos(self):
        _check(_lib.astraea_evaluate_eos(self._h))

    def gravity(self):
        _check(_lib.astraea_gravity(self._h))

    def pass2(self):
        _check(_lib.astraea_sph_pass2(self._h))

    def reduce_dt(self, max_dt: float) -> np.ndarray:
        out = (C.c_float * len(self.counts))()
        _check(_lib.astraea_reduce_dt(self._h, max_dt, out))
        return np.array(out[:], dtype=np.float32)

    def kick_half(self, clamp_thermal: bool = False):
        _check(_lib.astraea_kick_half(self._h, int(clamp_thermal)))

    def drift(self):
        _check(_lib.astraea_drift(self._h))

    def initial_forces(self):
        _check(_lib.astraea_evaluate_initial_forces(self._h))

    def step(self, max_dt: float, steps: int = 1, graph: bool = False):
        _check(_lib.astraea_step(self._h, max_dt, steps, int(graph)))

    def synchronize(self):
        _check(_lib.astraea_synchronize(self._h))

    def dt(self) -> np.ndarray:
        out = (C.c_float * len(self.counts))()
        _check(_lib.astraea_get_dt(self._h, out))
        return np.array(out[:], dtype=np.float32)

    def energy(self) -> np.ndarray:
        out = np.zeros(len(self.counts) * ENERGY_STRIDE, dtype=np.float64)
        _check(_lib.astraea_energy(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(len(self.counts), ENERGY_STRIDE)

    def classify(self, target_radius) -> np.ndarray:
        """Largest-remnant classification per segment (ASTRAEA_CL_* columns, see CL);
        target_radius: one PSU radius per segment, or a scalar for all."""
        ns = len(self.counts)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(target_radius, np.float64), (ns,)))
        out = np.zeros(ns * CLASS_STRIDE, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        _check(_lib.astraea_classify(self._h, r.ctypes.data_as(dp), out.ctypes.data_as(dp)))
        return out.reshape(ns, CLASS_STRIDE)

    def relax(self, steps: int, nominal_radius, anchor_temp=None) -> np.ndarray:
        """Damped relaxation, one body per segment; returns v_rms / v_esc per segment.
        anchor_temp: K per particle by original id, or None for the EOS temperatures."""
        ns = len(self.counts)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(nominal_radius, np.float64), (ns,)))
        out = np.zeros(ns, np.float64)
        a = None
        if anchor_temp is not None:
            a = np.ascontiguousarray(anchor_temp, np.float32)
            if len(a) != int(sum(self.counts)):
                raise ValueError("anchor_temp needs one value per particle")
        _check(_lib.astraea_relax(self._h, steps, None if a is None else a.ctypes.data, r.ctypes.data,
                                  out.ctypes.data))
        return out

    def pack_render(self, com_psu, speed: bool = False):
        """Render records (n x 20 bytes) about com_psu (3 per segment) and, if asked, speed bytes."""
        ns, n = len(self.counts), self.n
        com = np.ascontiguousarray(np.broadcast_to(np.asarray(com_psu, np.float64).reshape(-1, 3), (ns, 3)))
        rec = np.zeros((n, 20), np.uint8)
        spd = np.zeros(n, np.uint8) if speed else None
        _check(_lib.astraea_pack_render(self._h, com.ctypes.data, rec.ctypes.data, None if spd is None else spd.ctypes.data))
        return (rec, spd) if speed else rec

    def pack_render_noalloc(self, com_psu, records_dev: int, speed_dev: int | None = None):
        """Allocation-free render records into the caller's device buffers (device addresses)."""
        ns = len(self.counts)
        com = np.ascontiguousarray(np.broadcast_to(np.asarray(com_psu, np.float64).reshape(-1, 3), (ns, 3)))
        _check(_lib.astraea_pack_render_noalloc(self._h, com.ctypes.data, C.c_void_p(records_dev),
                                                None if speed_dev is None else C.c_void_p(speed_dev)))

    def set_tracers(self, particle_id, rest, ocean):
        """Tracer t on the particle of ORIGINAL_ID particle_id[t]; rest: (count, 4) radius + normal."""
        pid = np.ascontiguousarray(particle_id, np.uint32)
        r = np.ascontiguousarray(rest, np.float32).reshape(-1, 4)
        oc = np.ascontiguousarray(ocean, np.uint8)
        assert len(pid) == len(r) == len(oc)
        _check(_lib.astraea_set_tracers(self._h, len(pid), pid.ctypes.data, r.ctypes.data, oc.ctypes.data))
        self._tracers = len(pid)

    def extract_tracers(self, living_com_psu, sys_com_psu):
        """40-byte tracer records (count x 40) about the given centres (3 per segment)."""
        ns = len(self.counts)
        lc = np.ascontiguousarray(np.broadcast_to(np.asarray(living_com_psu, np.float64).reshape(-1, 3), (ns, 3)))
        sc = np.ascontiguousarray(np.broadcast_to(np.asarray(sys_com_psu, np.float64).reshape(-1, 3), (ns, 3)))
        out = np.zeros((getattr(self, "_tracers", 0), 40), np.uint8)
        _check(_lib.astraea_extract_tracers(self._h, lc.ctypes.data, sc.ctypes.data, out.ctypes.data))
        return out

    def counters(self) -> dict:
        out = (C.c_uint64 * 10)()
        _check(_lib.astraea_counters(self._h, out, 10))
        keys = ["interactions", "forced_accepts", "nbr_entries", "nbr_stride", "pass1_pairs", "pass2_pairs",
                "nbr_max", "steps_redone", "spilled_waves", "spill_bytes"]
        return dict(zip(keys, [int(v) for v in out]))

    def set_neighbour_stride(self, stride: int) -> None:
        """Width of the neighbour table in entries per particle; the table must be rebuilt after."""
        _check(_lib.astraea_set_neighbour_stride(self._h, stride))

    def tree(self, seg: int = 0) -> dict:
        n = self.counts[seg]
        ni = max(n - 1, 0)
        out = dict(keys=np.empty(n, np.uint64), perm=np.empty(n, np.uint32), left=np.empty(ni, np.int32),
                   right=np.empty(ni, np.int32), parent_internal=np.empty(ni, np.int32),
                   parent_leaf=np.empty(n, np.int32), node_cm=np.empty((ni, 4), np.float32),
                   node_min=np.empty((ni, 4), np.float32), node_max=np.empty((ni, 4), np.float32))
        order = ["keys", "perm", "left", "right", "parent_internal", "parent_leaf", "node_cm", "node_min", "node_max"]
        _check(_lib.astraea_download_tree(self._h, seg, *[_ptr(out[k]) for k in order]))
        return out

    def enable_stage_timing(self, on: bool = True):
        _check(_lib.astraea_enable_stage_timing(self._h, int(on)))

    def stage_times(self) -> dict:
        ms = (C.c_double * len(STAGES))()
        n = (C.c_uint64 * len(STAGES))()
        _check(_lib.astraea_stage_times(self._h, ms, n))
        return {s: (ms[k], int(n[k])) for k, s in enumerate(STAGES)}

    def tuning(self) -> dict:
        t = Tuning()
        _check(_lib.astraea_get_tuning(self._h, C.byref(t)))
        return {k: int(getattr(t, k)) for k, _ in Tuning._fields_}

    def set_tuning(self, pass1_round: int, pass2_round: int, drift_round: int, sort_bits: int = 8,
                   gravity_waves: int = 4, gravity_stack: int = 1024, build_waves: int = 4, build_stack: int = 1024):
        t = Tuning(pass1_round, pass2_round, drift_round, sort_bits, gravity_waves, gravity_stack, build_waves,
                   build_stack)
        _check(_lib.astraea_set_tuning(self._h, C.byref(t)))

    def autotune(self, max_dt: float, steps_per_variant: int = 1) -> dict:
        """Runs 4 + 8 * steps_per_variant steps (one untimed step per variant, then the timed ones)
        timing every launch variant of the passes, the drift
        and the sort, then times the gravity and neighbour-build variants on the settled state
        (astraea_autotune), and keeps the fastest; returns the choice and the measured mean
        milliseconds (-1 for a walk variant whose stack filled)."""
        t = Tuning()
        ms = (C.c_double * 18)()
        _check(_lib.astraea_autotune(self._h, float(max_dt), int(steps_per_variant), C.byref(t), ms))
        chosen = {k: int(getattr(t, k)) for k, _ in Tuning._fields_}
        walk = [f"{w}x{s}" for w, s in WALK_VARIANTS]
        return dict(chosen=chosen, pass1_ms={128: ms[0], 256: ms[1]}, pass2_ms={128: ms[2], 160: ms[3]},
                    drift_ms={128: ms[4], 160: ms[5], 192: ms[6], 256: ms[7]}, sort_ms={8: ms[8], 9: ms[9]},
                    gravity_ms={v: ms[10 + k] for k, v in enumerate(walk)},
                    build_ms={v: ms[14 + k] for k, v in enumerate(walk)})


def tuning_key(particles: int, segments: int, strength: bool, fp64: bool) -> str:
    """Key of a workload in the tuning table: the particle count to a power of two, the segments
    and the kernel variants in use."""
    return f"gfx950/n{max(int(particles), 1).bit_length()}/s{int(segments)}/{'strength' if strength else 'plain'}" \
           f"/{'fp64' if fp64 else 'fp32'}"


def tune_cached(ctx: "Context", max_dt: float, path, key: str, steps_per_variant: int = 1) -> dict:
    """Applies the tuning stored under `key` in the JSON table at `path`, or runs the autotuner
    (4 + 8 * steps_per_variant steps of the loaded state) and stores its choice there.  A stored
    choice this library no longer offers (a launch variant since replaced) is tuned again."""
    import json

    p = Path(path)
    table = json.loads(p.read_text()) if p.exists() else {}
    if key in table:
        t = table[key]
        try:
            ctx.set_tuning(t["pass1_round"], t["pass2_round"], t["drift_round"], t.get("sort_bits", 8),
                           t.get("gravity_waves", 4), t.get("gravity_stack", 1024), t.get("build_waves", 4),
                           t.get("build_stack", 1024))
            return dict(chosen=dict(t), cached=True)
        except AstraeaError as e:
            if e.code != E_INVALID:
                raise
    r = ctx.autotune(max_dt, steps_per_variant)
    table[key] = r["chosen"]
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    return dict(r, cached=False)


def splitmix64(x: int) -> int:
    return int(_lib.astraea_splitmix64(C.c_uint64(x & 0xFFFFFFFFFFFFFFFF)))


def build_body(n: int, radius_m: float, layers, h_over_spacing=1.3, jitter=0.1, seed=42, flaws=False,
               weibull_m=9.0, weibull_k=4.0e29, surface_temp_k=0.0) -> dict:
    """A layered body on a jittered HCP lattice (astraea_build_body).

    `layers`: [(material, mass_fraction, density_kgm3 or 0), ...] innermost first.
    """
    spec = BodySpec()
    _lib.astraea_default_body_spec(C.byref(spec))
    spec.radius_m = radius_m
    spec.num_layers = len(layers)
    for k, (mat, frac, dens) in enumerate(layers):
        spec.layer_material[k] = mat
        spec.layer_mass_frac[k] = frac
        spec.layer_density_kgm3[k] = dens
    spec.h_over_spacing = h_over_spacing
    spec.jitter = jitter
    spec.seed = seed
    spec.strength_flaws = int(flaws)
    spec.weibull_m = weibull_m
    spec.weibull_k = weibull_k
    spec.surface_temp_k = surface_temp_k
    arrs = {k: np.zeros(n, np.float32) for k in ("x", "y", "z", "m", "h", "rho", "u", "eps_fail")}
    arrs["mat"] = np.zeros(n, np.uint8)
    out = BodyOut(*[arrs[k].ctypes.data_as(C.POINTER(C.c_float)) for k in ("x", "y", "z", "m", "h", "rho", "u", "eps_fail")],
                  arrs["mat"].ctypes.data_as(C.POINTER(C.c_uint8)))
    mass = C.c_double()
    spacing = C.c_double()
    _check(_lib.astraea_build_body(C.byref(spec), n, C.byref(out), C.byref(mass), C.byref(spacing)))
    arrs["mass_kg"] = mass.value
    arrs["spacing_m"] = spacing.value
    return arrs


def planet_spec(mass_kg: float, layers, surface_temp_k: float, core_temp_k: float, surface_pressure_pa: float = 0.0,
                gamma_hhe: float = 1.4, h_factor: float = 1.2, seed: int = 42) -> PlanetSpec:
    """`layers`: [(material, mass_fraction), ...] innermost first."""
    s = PlanetSpec()
    _lib.astraea_default_planet_spec(C.byref(s))
    s.mass_kg = mass_kg
    s.num_layers = len(layers)
    for k, (mat, frac) in enumerate(layers):
        s.layer_material[k] = mat
        s.layer_mass_frac[k] = frac
    s.surface_temp_k, s.core_temp_k, s.surface_pressure_pa = surface_temp_k, core_temp_k, surface_pressure_pa
    s.gamma_hhe, s.h_factor, s.seed = gamma_hhe, h_factor, seed
    return s


def planet_profile(spec: PlanetSpec) -> dict:
    """Solved hydrostatic profile of `spec` (astraea_planet_profile)."""
    info = PlanetInfo()
    _check(_lib.astraea_planet_profile(C.byref(spec), 0, None, None, None, None, None, None, C.byref(info)))
    n = info.profile_samples
    out = {k: np.zeros(n) for k in ("r", "rho", "p", "m", "t")}
    out["mat"] = np.zeros(n, np.uint8)
    _check(_lib.astraea_planet_profile(C.byref(spec), n, *[out[k].ctypes.data for k in ("r", "rho", "p", "m", "t", "mat")],
                                       C.byref(info)))
    out.update(mass_kg=info.mass_kg, radius_m=info.radius_m, central_pressure_pa=info.central_pressure_pa)
    return out


def build_planet(spec: PlanetSpec, n: int) -> dict:
    """A planet of `n` particles in hydrostatic equilibrium (astraea_build_planet); arrays x, y, z,
    m, h, rho, u, eps_fail (PSU), mat, temp_k and the profile's mass, radius, central pressure."""
    arrs = {k: np.zeros(n, np.float32) for k in ("x", "y", "z", "m", "h", "rho", "u", "eps_fail", "temp_k")}
    arrs["mat"] = np.zeros(n, np.uint8)
    out = BodyOut(*[arrs[k].ctypes.data_as(C.POINTER(C.c_float)) for k in ("x", "y", "z", "m", "h", "rho", "u", "eps_fail")],
                  arrs["mat"].ctypes.data_as(C.POINTER(C.c_uint8)))
    info = PlanetInfo()
    _check(_lib.astraea_build_planet(C.byref(spec), n, C.byref(out), arrs["temp_k"].ctypes.data, C.byref(info)))
    arrs.update(mass_kg=info.mass_kg, radius_m=info.radius_m, central_pressure_pa=info.central_pressure_pa)
    return arrs


def select_tracers(x, y, z, h0, material, grid_res: int) -> dict:
    """LivingTerra crust tracers of a body about its centre (astraea_select_tracers)."""
    n = len(x)
    a = [np.ascontiguousarray(v, np.float32) for v in (x, y, z, h0)]
    mat = np.ascontiguousarray(material, np.uint8)
    out = dict(index=np.zeros(n, np.int32), rest_radius=np.zeros(n, np.float32), rest_normal=np.zeros((n, 3), np.float32),
               ocean=np.zeros(n, np.uint8))
    cnt = C.c_uint32()
    _check(_lib.astraea_select_tracers(n, *[v.ctypes.data for v in a], mat.ctypes.data, grid_res,
                                       out["index"].ctypes.data, out["rest_radius"].ctypes.data,
                                       out["rest_normal"].ctypes.data, out["ocean"].ctypes.data, C.byref(cnt)))
    out["count"] = cnt.value
    return out


def cube_face_uv_to_normal(face: int, u: float, v: float):
    n = [C.c_float() for _ in range(3)]
    _lib.astraea_cube_face_uv_to_normal(face, u, v, *[C.byref(q) for q in n])
    return tuple(q.value for q in n)


def normal_to_cube_face_uv(nx: float, ny: float, nz: float):
    f, u, v = C.c_uint32(), C.c_float(), C.c_float()
    _lib.astraea_normal_to_cube_face_uv(nx, ny, nz, C.byref(f), C.byref(u), C.byref(v))
    return f.value, u.value, v.value


def body_radius_for_mass(layers, mass_kg: float) -> float:
    """Outer radius (m) of a body of `layers` ([(material, mass_fraction, density or 0), ...],
    innermost first) holding `mass_kg` (astraea_body_radius_for_mass)."""
    spec = BodySpec()
    _lib.astraea_default_body_spec(C.byref(spec))
    spec.num_layers = len(layers)
    for k, (mat, frac, dens) in enumerate(layers):
        spec.layer_material[k] = mat
        spec.layer_mass_frac[k] = frac
        spec.layer_density_kgm3[k] = dens
    r = C.c_double()
    _check(_lib.astraea_body_radius_for_mass(C.byref(spec), C.c_double(mass_kg), C.byref(r)))
    return r.value


def place_two_bodies(state: dict, n_t: int, n_i: int, r_t_psu: float, r_i_psu: float, v_ratio: float,
                     b: float, sep_radii: float) -> float:
    arrs = [state[k] for k in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "mass")]
    for a in arrs:
        assert a.dtype == np.float32 and a.flags.c_contiguous
    vesc = C.c_double()
    _check(_lib.astraea_place_two_bodies(n_t, n_i, *[_ptr(a) for a in arrs], r_t_psu, r_i_psu, v_ratio, b,
                                         sep_radii, C.byref(vesc)))
    return vesc.value


from . import scenario  # noqa: E402,F401
