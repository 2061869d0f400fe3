"""FP64 CPU oracle for cuco's gravity path -- TEST INFRASTRUCTURE ONLY.

This package is the parity checker for the MI355X engine in ``cuco_amd``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it, and only to check results or to time the
CPU baseline.  The product (``cuco_amd``) never imports it.

The C restatement lives in ``cuco_oracle.c`` (each function cites the
reference lines it follows); this module is a thin ctypes/numpy binding.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
NTAB = 1000
DRIFT_TABLE_LENGTH = 1000


class OrcParams(C.Structure):
    _fields_ = [
        ("M", C.c_int32),
        ("nthreads", C.c_int32),
        ("L", C.c_double),
        ("G", C.c_double),
        ("asmth", C.c_double),
        ("rcut", C.c_double),
        ("rcut2", C.c_double),
        ("asmthfac", C.c_double),
        ("to_grid_fac", C.c_double),
        ("boxsize", C.c_double),
        ("boxhalf", C.c_double),
        ("h", C.c_double),
        ("h_inv", C.c_double),
        ("h3_inv", C.c_double),
        ("table", C.c_float * NTAB),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp, fp, up, i64p = (C.POINTER(C.c_double), C.POINTER(C.c_float),
                            C.POINTER(C.c_uint32), C.POINTER(C.c_int64))
        i32p = C.POINTER(C.c_int32)
        P = C.POINTER(OrcParams)
        sigs = {
            "orc_shortrange_table": (None, [fp]),
            "orc_set_params": (None, [P, C.c_int32, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double]),
            "orc_calc_keys": (None, [P, dp, C.c_int64, up, up]),
            "orc_sort_pairs": (None, [up, up, C.c_int64]),
            "orc_cell_table": (None, [up, C.c_int64, C.c_int64, up, up]),
            "orc_pp_force": (C.c_int, [P, dp, dp, C.c_int64, i64p, C.c_int64,
                                       dp, C.POINTER(C.c_uint64)]),
            "orc_pp_force_bruteforce": (C.c_int, [P, dp, dp, C.c_int64, dp]),
            "orc_pp_last_timing": (None, [dp, dp]),
            "orc_pm_deposit": (None, [P, dp, dp, C.c_int64, dp]),
            "orc_pm_potential": (C.c_int, [P, dp, dp]),
            "orc_pm_interpolate": (C.c_int, [P, dp, dp, C.c_int64, dp]),
            "orc_pm_force": (C.c_int, [P, dp, dp, C.c_int64, dp]),
            "orc_fft1d": (None, [C.c_int64, dp, C.c_int]),
            "orc_fft3d_r2c": (C.c_int, [C.c_int32, dp, dp]),
            "orc_fft3d_c2r": (C.c_int, [C.c_int32, dp, dp]),
            "orc_drift_tables": (None, [C.c_double] * 5 + [dp, dp]),
            "orc_table_factor": (C.c_double, [dp] + [C.c_double] * 4),
            "orc_kdk_run": (C.c_int, [P, dp, dp, dp, C.c_int64, C.c_int32, dp,
                                      dp, dp]),
            "orc_timeline_run": (C.c_int, [P, dp, dp, dp, dp, dp, dp, C.c_int64,
                                           i32p, i32p, i32p, dp, dp, dp,
                                           C.c_int32, C.c_int32, i32p]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _pos(pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    assert pos.ndim == 2 and pos.shape[1] == 3
    return pos


def params(M, L, G, asmth, rcut, soft_h, nthreads=0) -> OrcParams:
    p = OrcParams()
    p.nthreads = nthreads
    lib().orc_set_params(C.byref(p), M, L, G, asmth, rcut, soft_h)
    return p


def shortrange_table() -> np.ndarray:
    t = np.zeros(NTAB, dtype=np.float32)
    lib().orc_shortrange_table(t.ctypes.data_as(C.POINTER(C.c_float)))
    return t


def calc_keys(p, pos):
    pos = _pos(pos)
    n = pos.shape[0]
    keys = np.zeros(n, np.uint32)
    idx = np.zeros(n, np.uint32)
    lib().orc_calc_keys(C.byref(p), _d(pos), n, _u(keys), _u(idx))
    return keys, idx


def sort_pairs(keys, vals):
    keys = np.ascontiguousarray(keys, np.uint32).copy()
    vals = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_sort_pairs(_u(keys), _u(vals), keys.shape[0])
    return keys, vals


def cell_table(sorted_keys, ncells):
    sorted_keys = np.ascontiguousarray(sorted_keys, np.uint32)
    start = np.zeros(ncells, np.uint32)
    end = np.zeros(ncells, np.uint32)
    lib().orc_cell_table(_u(sorted_keys), sorted_keys.shape[0], ncells,
                         _u(start), _u(end))
    return start, end


def pp_force(p, pos, mass, targets=None):
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, np.float64)
    n = pos.shape[0]
    cnt = C.c_uint64(0)
    if targets is None:
        acc = np.zeros((n, 3))
        rc = lib().orc_pp_force(C.byref(p), _d(pos), _d(mass), n, None, 0,
                                _d(acc), C.byref(cnt))
    else:
        t = np.ascontiguousarray(targets, np.int64)
        acc = np.zeros((t.shape[0], 3))
        rc = lib().orc_pp_force(C.byref(p), _d(pos), _d(mass), n,
                                t.ctypes.data_as(C.POINTER(C.c_int64)),
                                t.shape[0], _d(acc), C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f"orc_pp_force failed: {rc}")
    return acc, int(cnt.value)


def pp_last_timing():
    """(setup_s, eval_s) of the last pp_force call."""
    a, b = C.c_double(), C.c_double()
    lib().orc_pp_last_timing(C.byref(a), C.byref(b))
    return a.value, b.value


def pp_force_bruteforce(p, pos, mass):
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, np.float64)
    acc = np.zeros_like(pos)
    lib().orc_pp_force_bruteforce(C.byref(p), _d(pos), _d(mass), pos.shape[0],
                                  _d(acc))
    return acc


def pm_deposit(p, pos, mass):
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, np.float64)
    rho = np.zeros((p.M,) * 3)
    lib().orc_pm_deposit(C.byref(p), _d(pos), _d(mass), pos.shape[0], _d(rho))
    return rho


def pm_potential(p, rho):
    rho = np.ascontiguousarray(rho, np.float64)
    phi = np.zeros_like(rho)
    if lib().orc_pm_potential(C.byref(p), _d(rho), _d(phi)) != 0:
        raise RuntimeError("orc_pm_potential failed")
    return phi


def pm_interpolate(p, phi, pos):
    pos = _pos(pos)
    phi = np.ascontiguousarray(phi, np.float64)
    acc = np.zeros_like(pos)
    if lib().orc_pm_interpolate(C.byref(p), _d(phi), _d(pos), pos.shape[0],
                                _d(acc)) != 0:
        raise RuntimeError("orc_pm_interpolate failed")
    return acc


def pm_force(p, pos, mass):
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, np.float64)
    acc = np.zeros_like(pos)
    if lib().orc_pm_force(C.byref(p), _d(pos), _d(mass), pos.shape[0],
                          _d(acc)) != 0:
        raise RuntimeError("orc_pm_force failed")
    return acc


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, np.complex128).copy()
    lib().orc_fft1d(x.shape[0], x.ctypes.data_as(C.POINTER(C.c_double)), sign)
    return x


def fft3d_r2c(a):
    a = np.ascontiguousarray(a, np.float64)
    M = a.shape[0]
    out = np.zeros((M, M, M // 2 + 1), np.complex128)
    lib().orc_fft3d_r2c(M, _d(a), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def fft3d_c2r(c):
    c = np.ascontiguousarray(c, np.complex128)
    M = c.shape[0]
    out = np.zeros((M, M, M))
    lib().orc_fft3d_c2r(M, c.ctypes.data_as(C.POINTER(C.c_double)), _d(out))
    return out


def drift_tables(a_begin, a_max, omega0, omega_lambda, hubble):
    d = np.zeros(DRIFT_TABLE_LENGTH)
    k = np.zeros(DRIFT_TABLE_LENGTH)
    lib().orc_drift_tables(a_begin, a_max, omega0, omega_lambda, hubble, _d(d),
                           _d(k))
    return d, k


def table_factor(table, a_begin, a_max, loga0, loga1):
    table = np.ascontiguousarray(table, np.float64)
    return lib().orc_table_factor(_d(table), a_begin, a_max, loga0, loga1)


def kdk_run(p, pos, vel, mass, factors):
    pos = _pos(pos).copy()
    vel = _pos(vel).copy()
    mass = np.ascontiguousarray(mass, np.float64)
    f = np.ascontiguousarray(factors, np.float64).reshape(-1, 3)
    n = pos.shape[0]
    app = np.zeros_like(pos)
    apm = np.zeros_like(pos)
    rc = lib().orc_kdk_run(C.byref(p), _d(pos), _d(vel), _d(mass), n,
                           f.shape[0], _d(f), _d(app), _d(apm))
    if rc != 0:
        raise RuntimeError(f"orc_kdk_run failed: {rc}")
    return pos, vel, app, apm


TIMELINE_FIELDS = ("time_begin", "time_max", "omega0", "omega_lambda", "hubble",
                   "err_tol_int_accuracy", "max_size_timestep",
                   "min_size_timestep", "softening_halo",
                   "softening_halo_max_phys")


class Timeline:
    """Integer-timeline state of the oracle run (arrays updated in place)."""

    def __init__(self, p, tl: dict, drift, kick, pos, vel, mass):
        self.p = p
        self.tl = np.array([tl[k] for k in TIMELINE_FIELDS], np.float64)
        self.drift = np.ascontiguousarray(drift, np.float64)
        self.kick = np.ascontiguousarray(kick, np.float64)
        self.pos = _pos(pos).copy()
        self.vel = _pos(vel).copy()
        self.mass = np.ascontiguousarray(mass, np.float64)
        n = self.pos.shape[0]
        self.ti_beg = np.zeros(n, np.int32)
        self.ti_end = np.zeros(n, np.int32)
        self.state = np.zeros(4, np.int32)
        self.time = C.c_double(tl["time_begin"])
        self.app = np.zeros((n, 3))
        self.apm = np.zeros((n, 3))
        self.status = np.zeros(3, np.int32)

    def run(self, max_steps, ti_stop=-1):
        i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        rc = lib().orc_timeline_run(
            C.byref(self.p), _d(self.tl), _d(self.drift), _d(self.kick),
            _d(self.pos), _d(self.vel), _d(self.mass), self.pos.shape[0],
            i32(self.ti_beg), i32(self.ti_end), i32(self.state), C.byref(self.time),
            _d(self.app), _d(self.apm), max_steps, ti_stop, i32(self.status))
        if rc != 0:
            raise RuntimeError(f"orc_timeline_run failed: {rc}")
        return {"ti_current": int(self.state[0]), "pm_ti_begstep": int(self.state[1]),
                "pm_ti_endstep": int(self.state[2]), "steps": int(self.state[3]),
                "time": self.time.value, "finished": int(self.status[0]),
                "at_stop": int(self.status[1]), "flags": int(self.status[2])}
