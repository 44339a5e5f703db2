"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the timed CPU baseline.  The
product never imports it.  The API mirrors astraea_amd.Context so that a test can
run the same stage sequence on both and compare.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# ASTRAEA_ORACLE_LIB names another build of oracle.cpp: tests/test_oracle_mutants.py loads copies
# with one constant or sign changed by hand to show that the pinning tests catch the change.
LIB_PATH = Path(os.environ.get("ASTRAEA_ORACLE_LIB", str(_HERE / "liboracle.so")))


def build():
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


if not LIB_PATH.exists() and "ASTRAEA_ORACLE_LIB" not in os.environ:
    build()
_lib = C.CDLL(str(LIB_PATH))

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
ENERGY_STRIDE = 18


class Params(C.Structure):
    _fields_ = [
        ("theta", C.c_float), ("soft_factor", C.c_float), ("cfl_factor", C.c_float), ("gamma_hhe", C.c_float),
        ("cooling_min_temp_k", C.c_float), ("strength_mask", C.c_uint32), ("precision", C.c_uint32),
        ("group_size", C.c_uint32), ("reserved", C.c_uint32 * 4),
    ]


_P = C.c_void_p
_sig = {
    "or_create": (_P, []),
    "or_destroy": (None, [_P]),
    "or_set_params": (C.c_int, [_P, C.POINTER(Params), C.c_int]),
    "or_set_layout": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "or_upload_field": (C.c_int, [_P, C.c_int, _P]),
    "or_download_field": (C.c_int, [_P, C.c_int, _P]),
    "or_download_field_f64": (C.c_int, [_P, C.c_int, _P]),
    "or_clear_stats": (C.c_int, [_P]),
    "or_sort_and_build": (C.c_int, [_P]),
    "or_pass1": (C.c_int, [_P]),
    "or_eos": (C.c_int, [_P]),
    "or_gravity": (C.c_int, [_P]),
    "or_pass2": (C.c_int, [_P]),
    "or_reduce_dt": (C.c_int, [_P, C.c_double, C.POINTER(C.c_double)]),
    "or_kick_half": (C.c_int, [_P, C.c_int]),
    "or_drift": (C.c_int, [_P]),
    "or_initial_forces": (C.c_int, [_P]),
    "or_step": (C.c_int, [_P, C.c_double, C.c_uint32]),
    "or_energy": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "or_relax": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, _P, _P]),
    "or_classify_arrays": (None, [C.c_uint32] + [_P] * 10 + [C.c_double, _P]),
    "or_counters": (C.c_int, [_P, C.POINTER(C.c_uint64), C.c_int]),
    "or_neighbour_stats": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "or_set_dt": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "or_get_dt": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "or_download_tree": (C.c_int, [_P, C.c_uint32] + [_P] * 9),
    "or_num_threads": (C.c_int, []),
    "or_cold_energy_psu": (C.c_double, [C.c_int, C.c_double]),
    "or_eos_eval": (None, [C.c_int, C.c_float, C.c_float, C.c_float, C.c_double, C.POINTER(C.c_float),
                           C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint8)]),
    "or_eos_from_rho_temp_si": (None, [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
    "or_invert_density_si": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double]),
    "or_material_param": (C.c_double, [C.c_int, C.c_int]),
    "or_morton_key": (C.c_uint64, [C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)]),
    "or_morton_key64": (C.c_uint64, [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]),
}
for _name, (_res, _args) in _sig.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle call failed: {rc}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def num_threads() -> int:
    return int(_lib.or_num_threads())


def default_params(**kw) -> Params:
    """The reference's compile-time constants (engine/sph_gravity.cu, eos.cu, simulation.cpp)."""
    p = Params()
    p.theta, p.soft_factor, p.cfl_factor, p.gamma_hhe = 0.5, 0.8, 0.25, 1.4
    p.cooling_min_temp_k, p.strength_mask, p.precision = 0.0, 0x17, 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def params_from(p) -> Params:
    q = Params()
    for name, _ in Params._fields_:
        if name != "reserved":
            setattr(q, name, getattr(p, name))
    return q


class Oracle:
    def __init__(self, params, group_size: int = 64):
        self._h = _lib.or_create()
        self.group_size = group_size
        _check(_lib.or_set_params(self._h, C.byref(params_from(params)), group_size))
        self.counts: list[int] = []

    def close(self):
        if self._h:
            _lib.or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        return int(sum(self.counts))

    def set_layout(self, counts):
        arr = (C.c_uint32 * len(counts))(*[int(x) for x in counts])
        _check(_lib.or_set_layout(self._h, len(counts), arr))
        self.counts = [int(x) for x in counts]

    def upload(self, name, values):
        dt = np.uint8 if name in _U8 else np.uint32 if name in _U32 else np.float64 if name in _F64 else np.float32
        a = np.ascontiguousarray(values, dtype=dt)
        assert a.shape == (self.n,), (name, a.shape, self.n)
        _check(_lib.or_upload_field(self._h, FIELD[name], _ptr(a)))

    def load_state(self, state: dict):
        for k, v in state.items():
            self.upload(k, v)

    def download(self, name) -> np.ndarray:
        dt = np.uint8 if name in _U8 else np.uint32 if name in _U32 else np.float64 if name in _F64 else np.float32
        a = np.empty(self.n, dtype=dt)
        _check(_lib.or_download_field(self._h, FIELD[name], _ptr(a)))
        return a

    def download64(self, name) -> np.ndarray:
        a = np.empty(self.n, dtype=np.float64)
        _check(_lib.or_download_field_f64(self._h, FIELD[name], _ptr(a)))
        return a

    def clear_stats(self):
        _check(_lib.or_clear_stats(self._h))

    def sort_and_build(self):
        _check(_lib.or_sort_and_build(self._h))

    def pass1(self):
        _check(_lib.or_pass1(self._h))

    def eos(self):
        _check(_lib.or_eos(self._h))

    def gravity(self):
        _check(_lib.or_gravity(self._h))

    def pass2(self):
        _check(_lib.or_pass2(self._h))

    def reduce_dt(self, max_dt: float):
        out = (C.c_double * len(self.counts))()
        _check(_lib.or_reduce_dt(self._h, max_dt, out))
        return np.array(out[:])

    def set_dt(self, dt):
        arr = (C.c_double * len(self.counts))(*[float(x) for x in dt])
        _check(_lib.or_set_dt(self._h, arr))

    def dt(self):
        out = (C.c_double * len(self.counts))()
        _check(_lib.or_get_dt(self._h, out))
        return np.array(out[:])

    def kick_half(self, clamp_thermal=False):
        _check(_lib.or_kick_half(self._h, int(clamp_thermal)))

    def drift(self):
        _check(_lib.or_drift(self._h))

    def initial_forces(self):
        _check(_lib.or_initial_forces(self._h))

    def step(self, max_dt: float, steps: int = 1):
        _check(_lib.or_step(self._h, max_dt, steps))

    def energy(self) -> np.ndarray:
        out = np.zeros(len(self.counts) * ENERGY_STRIDE, dtype=np.float64)
        _check(_lib.or_energy(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(len(self.counts), ENERGY_STRIDE)

    def relax(self, steps: int, nominal_radius, anchor_temp=None) -> np.ndarray:
        """Damped relaxation, one body per segment; returns v_rms / v_esc per segment."""
        ns = len(self.counts)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(nominal_radius, np.float64), (ns,)))
        out = np.zeros(ns, np.float64)
        total = int(sum(self.counts))
        a = None if anchor_temp is None else np.ascontiguousarray(anchor_temp, np.float64)
        _check(_lib.or_relax(self._h, steps, None if a is None else a.ctypes.data, total, r.ctypes.data,
                             out.ctypes.data))
        return out

    def counters(self) -> dict:
        out = (C.c_uint64 * 6)()
        _check(_lib.or_counters(self._h, out, 6))
        keys = ["interactions", "forced_accepts", "nbr_entries", "nbr_stride", "pass1_pairs", "pass2_pairs"]
        return dict(zip(keys, [int(v) for v in out]))

    def neighbour_stats(self) -> dict:
        """Brute-force size of the neighbour table for the current state: all candidates and the
        largest count of one particle."""
        out = (C.c_uint64 * 2)()
        _check(_lib.or_neighbour_stats(self._h, out))
        return {"nbr_entries": int(out[0]), "nbr_max": int(out[1])}

    def tree(self, seg=0) -> dict:
        n = self.counts[seg]
        ni = max(n - 1, 0)
        out = dict(keys=np.empty(n, np.uint64), perm=np.empty(n, np.uint32), left=np.empty(ni, np.int32),
                   right=np.empty(ni, np.int32), parent_internal=np.empty(ni, np.int32),
                   parent_leaf=np.empty(n, np.int32), node_cm=np.empty((ni, 4), np.float32),
                   node_min=np.empty((ni, 4), np.float32), node_max=np.empty((ni, 4), np.float32))
        order = ["keys", "perm", "left", "right", "parent_internal", "parent_leaf", "node_cm", "node_min", "node_max"]
        _check(_lib.or_download_tree(self._h, seg, *[_ptr(out[k]) for k in order]))
        return out


def eos_eval(mat, rho, u, damage=0.0, gamma=1.4):
    p, cs, t = C.c_float(), C.c_float(), C.c_float()
    ph = C.c_uint8()
    _lib.or_eos_eval(mat, rho, u, damage, gamma, C.byref(p), C.byref(cs), C.byref(t), C.byref(ph))
    return p.value, cs.value, t.value, ph.value


CLASS_STRIDE = 36


def classify_arrays(x, y, z, vx, vy, vz, m, phi, body, damage, target_radius: float) -> np.ndarray:
    """Largest-remnant classification of one body pair's arrays in input order
    (engine/simulation.cpp:1000-1227); returns CLASS_STRIDE doubles in ASTRAEA_CL_* order."""
    n = len(m)
    d = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, vx, vy, vz, m)]
    ph = np.ascontiguousarray(phi, np.float32)
    b = np.ascontiguousarray(body, np.uint8)
    dm = np.ascontiguousarray(damage, np.float32)
    assert all(len(a) == n for a in d + [ph, b, dm])
    out = np.zeros(CLASS_STRIDE, np.float64)
    _lib.or_classify_arrays(n, *[a.ctypes.data for a in d], ph.ctypes.data, b.ctypes.data, dm.ctypes.data,
                            float(target_radius), out.ctypes.data)
    return out


def cold_energy_psu(mat, rho_psu):
    return float(_lib.or_cold_energy_psu(mat, rho_psu))


def eos_from_rho_temp_si(mat, rho, temp):
    u, p = C.c_double(), C.c_double()
    _lib.or_eos_from_rho_temp_si(mat, rho, temp, C.byref(u), C.byref(p))
    return u.value, p.value


def invert_density_si(mat, p, temp, guess):
    return float(_lib.or_invert_density_si(mat, p, temp, guess))


PARAM_NAMES = ["rho0", "a", "b", "A", "B", "E0", "Eiv", "Ecv", "alpha", "beta", "cv", "tMelt", "tRef", "mu", "Y0",
               "epsFailMin", "kH2", "gamma", "bulk"]


def material(mat) -> dict:
    return {k: float(_lib.or_material_param(mat, i)) for i, k in enumerate(PARAM_NAMES)}


def morton_key(x, y, z, bounds6) -> int:
    b = (C.c_float * 6)(*[float(v) for v in bounds6])
    return int(_lib.or_morton_key(x, y, z, b))


def morton_key64(x, y, z, bounds6) -> int:
    """The Morton key of the mixed form (FP64 positions and box)."""
    b = (C.c_double * 6)(*[float(v) for v in bounds6])
    return int(_lib.or_morton_key64(float(x), float(y), float(z), b))
