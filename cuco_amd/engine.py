"""ctypes binding of ``libcuco_engine.so`` (C ABI in ``include/cuco_engine.h``).

The names mirror the reference host interface of the gravity path:

* ``Engine(...)``            -- linklist_init + pm_init_periodic
                                (gravlist.cpp:48-115, pm_periodic.c:25-43)
* ``Engine.set_softening``   -- set_softenings (gravlist.cpp:28-45)
* ``Engine.gravity_linklist``-- gravity_linklist_cuda (gravlist.cpp:148-193)
* ``Engine.pmforce_periodic``-- pmforce_periodic (pm_periodic.c:67-314)
* ``Engine.compute_accelerations`` -- accel.c:8-23
* ``Engine.move_particles``  -- move_particles + do_box_wrapping (predict.c)
* ``Engine.kick``            -- the kicks of advance_and_find_timesteps
                                (timestep.c:51-59, 86-92)

Errors surface as :class:`CucoError` carrying the library's status and
message; nothing falls back to a CPU path.  If the shared library is absent
the import of this module fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcuco_engine.so")

FP64 = 0
FP32 = 1
NTAB = 1000

STAGES = ["keys", "sort", "cells", "pp", "deposit", "fft_fwd", "green_x",
          "fft_inv", "interp"]

EXPORTED = [
    "cuco_status_string", "cuco_last_error", "cuco_create", "cuco_destroy",
    "cuco_set_softening", "cuco_upload_particles", "cuco_download_particles",
    "cuco_download_accel", "cuco_pp_force", "cuco_pm_force",
    "cuco_compute_accelerations", "cuco_drift", "cuco_kick", "cuco_kdk_steps",
    "cuco_pmforce_periodic", "cuco_gravity_linklist", "cuco_device_alloc",
    "cuco_device_free", "cuco_copy_to_device", "cuco_copy_from_device",
    "cuco_radix_sort", "cuco_fft3d_r2c", "cuco_fft3d_c2r", "cuco_get_sorted",
    "cuco_get_cell_table", "cuco_get_shortrange_table", "cuco_get_pm_grid",
    "cuco_count_pairs", "cuco_set_timing", "cuco_get_stage_times",
    "cuco_stage_name", "cuco_synchronize", "cuco_init_drift_table",
    "cuco_table_factor", "cuco_kdk_factors", "cuco_copy_device_to_device",
    "cuco_get_tuning", "cuco_set_tuning", "cuco_autotune", "cuco_set_params",
    "cuco_ensemble_create", "cuco_ensemble_size", "cuco_write_snapshot",
    "cuco_read_snapshot", "cuco_read_output_list", "cuco_find_next_output_time",
    "cuco_download_records", "cuco_upload_records", "cuco_timeline_init",
    "cuco_timeline_run", "cuco_get_timesteps", "cuco_init_records",
    "cuco_radix_sort_noalloc", "cuco_device_drift_table", "cuco_checkpoint_bytes",
    "cuco_checkpoint_info", "cuco_checkpoint_write", "cuco_checkpoint_read",
]


class Tuning(C.Structure):
    _fields_ = [("pp_block_shift", C.c_int32), ("pp_workgroups_per_cu", C.c_int32),
                ("pm_overlap", C.c_int32), ("pp_symmetric", C.c_int32),
                ("pp_cell_level", C.c_int32), ("fft_split", C.c_int32),
                ("pp_wide_tiles", C.c_int32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class CucoError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: status {status}: {message}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("pm_grid", C.c_int32),
        ("reserved0", C.c_int32),
        ("box_size", C.c_double),
        ("G", C.c_double),
        ("asmth", C.c_double),
        ("rcut", C.c_double),
        ("soft_h", C.c_double),
    ]


class Particle(C.Structure):
    """PARTICLE record of allvars.h:35-45 (60 bytes)."""
    _fields_ = [
        ("Pos", C.c_float * 3),
        ("Mass", C.c_float),
        ("Vel", C.c_float * 3),
        ("GravAccel", C.c_float * 3),
        ("GravPM", C.c_float * 3),
        ("Ti_endstep", C.c_int32),
        ("Ti_begstep", C.c_int32),
    ]


class Header(C.Structure):
    """HEADER record of allvars.h:47-57 (32 bytes)."""
    _fields_ = [
        ("NumPart", C.c_int32), ("Mass", C.c_float), ("Time", C.c_float),
        ("Redshift", C.c_float), ("BoxSize", C.c_float), ("Omega0", C.c_float),
        ("OmegaLambda", C.c_float), ("HubbleParam", C.c_float),
    ]


class Timeline(C.Structure):
    """Integrator parameters of begrun.c:33-56."""
    _fields_ = [(f, C.c_double) for f in (
        "time_begin", "time_max", "omega0", "omega_lambda", "hubble",
        "err_tol_int_accuracy", "max_size_timestep", "min_size_timestep",
        "softening_halo", "softening_halo_max_phys")]


class TimelineState(C.Structure):
    _fields_ = [("ti_current", C.c_int32), ("pm_ti_begstep", C.c_int32),
                ("pm_ti_endstep", C.c_int32), ("steps", C.c_int32),
                ("time", C.c_double), ("timebase_interval", C.c_double),
                ("active", C.c_int64), ("finished", C.c_int32),
                ("at_stop", C.c_int32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class CheckpointHeader(C.Structure):
    """Header of a checkpoint file (cuco_checkpoint_header, 168 bytes)."""
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_int32), ("nboxes", C.c_int32),
                ("n", C.c_int64), ("timeline", C.c_int32), ("pp_precision", C.c_int32),
                ("pm_precision", C.c_int32), ("reserved", C.c_int32),
                ("params", Timeline), ("state", TimelineState)]


PARTICLE_DTYPE = np.dtype([
    ("Pos", "<f4", 3), ("Mass", "<f4"), ("Vel", "<f4", 3),
    ("GravAccel", "<f4", 3), ("GravPM", "<f4", 3),
    ("Ti_endstep", "<i4"), ("Ti_begstep", "<i4"),
])

_lib = None


def lib():
    """Load the engine library (fails loudly when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m cuco_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, dp, up = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint32)
    i32, i64 = C.c_int32, C.c_int64
    sigs = {
        "cuco_status_string": (C.c_char_p, [C.c_int]),
        "cuco_last_error": (C.c_char_p, []),
        "cuco_create": (C.c_int, [C.POINTER(Params), i64, i32, C.POINTER(vp)]),
        "cuco_destroy": (C.c_int, [vp]),
        "cuco_set_softening": (C.c_int, [vp, C.c_double]),
        "cuco_set_params": (C.c_int, [vp, i32, C.POINTER(Params)]),
        "cuco_ensemble_create": (C.c_int, [C.POINTER(Params), C.POINTER(i64), i32,
                                           i32, C.POINTER(vp)]),
        "cuco_ensemble_size": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i64)]),
        "cuco_write_snapshot": (C.c_int, [C.c_char_p, C.POINTER(Header), vp]),
        "cuco_read_snapshot": (C.c_int, [C.c_char_p, C.POINTER(Header), vp, i64]),
        "cuco_read_output_list": (C.c_int, [C.c_char_p, dp, i32, C.POINTER(i32)]),
        "cuco_find_next_output_time": (i32, [dp, i32, C.c_double, C.c_double,
                                             C.c_double, i32]),
        "cuco_download_records": (C.c_int, [vp, vp]),
        "cuco_timeline_init": (C.c_int, [vp, C.POINTER(Timeline)]),
        "cuco_timeline_run": (C.c_int, [vp, i32, i32, i32, C.POINTER(TimelineState)]),
        "cuco_get_timesteps": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32)]),
        "cuco_upload_records": (C.c_int, [vp, vp, i64]),
        "cuco_init_records": (C.c_int, [vp, i64, C.c_double]),
        "cuco_checkpoint_bytes": (i64, [C.POINTER(CheckpointHeader)]),
        "cuco_checkpoint_info": (C.c_int, [C.c_char_p, C.POINTER(CheckpointHeader)]),
        "cuco_checkpoint_write": (C.c_int, [vp, C.c_char_p]),
        "cuco_checkpoint_read": (C.c_int, [vp, C.c_char_p]),
        "cuco_upload_particles": (C.c_int, [vp, i64, dp, dp, dp, C.c_double]),
        "cuco_download_particles": (C.c_int, [vp, dp, dp]),
        "cuco_download_accel": (C.c_int, [vp, dp, dp]),
        "cuco_pp_force": (C.c_int, [vp, i32]),
        "cuco_pm_force": (C.c_int, [vp, i32]),
        "cuco_compute_accelerations": (C.c_int, [vp, i32, i32]),
        "cuco_drift": (C.c_int, [vp, C.c_double]),
        "cuco_kick": (C.c_int, [vp, C.c_double, C.c_double]),
        "cuco_kdk_steps": (C.c_int, [vp, i32, i32, dp, i32]),
        "cuco_pmforce_periodic": (C.c_int, [vp, vp, i64, i32]),
        "cuco_gravity_linklist": (C.c_int, [vp, vp, i64, i32]),
        "cuco_device_alloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
        "cuco_device_free": (C.c_int, [vp]),
        "cuco_copy_to_device": (C.c_int, [vp, vp, C.c_size_t]),
        "cuco_copy_from_device": (C.c_int, [vp, vp, C.c_size_t]),
        "cuco_copy_device_to_device": (C.c_int, [vp, vp, C.c_size_t]),
        "cuco_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "cuco_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "cuco_autotune": (C.c_int, [vp, C.POINTER(Tuning), dp]),
        "cuco_radix_sort": (C.c_int, [vp, vp, C.c_uint32, C.c_uint32]),
        "cuco_radix_sort_noalloc": (C.c_int, [vp, vp, C.c_uint32, C.c_uint32]),
        "cuco_fft3d_r2c": (C.c_int, [i32, i32, vp, vp]),
        "cuco_fft3d_c2r": (C.c_int, [i32, i32, vp, vp]),
        "cuco_get_sorted": (C.c_int, [vp, up, up]),
        "cuco_get_cell_table": (C.c_int, [vp, up, up]),
        "cuco_get_shortrange_table": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "cuco_get_pm_grid": (C.c_int, [vp, i32, dp]),
        "cuco_count_pairs": (C.c_int, [vp, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]),
        "cuco_set_timing": (C.c_int, [vp, i32]),
        "cuco_get_stage_times": (C.c_int, [vp, dp, i32, C.POINTER(i32), i32]),
        "cuco_stage_name": (C.c_char_p, [i32]),
        "cuco_synchronize": (C.c_int, [vp]),
        "cuco_init_drift_table": (C.c_int, [C.c_double] * 5 + [dp, dp]),
        "cuco_device_drift_table": (C.c_int, [i32] + [C.c_double] * 5 + [dp, dp]),
        "cuco_table_factor": (C.c_double, [dp] + [C.c_double] * 4),
        "cuco_kdk_factors": (C.c_int, [dp, dp] + [C.c_double] * 4 + [i32, dp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int, where: str) -> None:
    if rc != 0:
        msg = lib().cuco_last_error().decode(errors="replace")
        raise CucoError(rc, where, msg)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _xyz(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 3 or (n is not None and a.shape[0] != n):
        raise ValueError("expected an (n, 3) array")
    return a


DRIFT_TABLE_LENGTH = 1000


def init_drift_table(a_begin, a_max, omega0, omega_lambda, hubble):
    """init_drift_table (driftfac.c:14-42): returns (drift, kick) tables."""
    d = np.zeros(DRIFT_TABLE_LENGTH)
    k = np.zeros(DRIFT_TABLE_LENGTH)
    _check(lib().cuco_init_drift_table(a_begin, a_max, omega0, omega_lambda,
                                       hubble, _dp(d), _dp(k)),
           "cuco_init_drift_table")
    return d, k


def device_drift_table(a_begin, a_max, omega0, omega_lambda, hubble, device=0):
    """init_drift_table (driftfac.c:14-42) built on the device (the tables
    cuco_timeline_init uses): returns (drift, kick)."""
    d = np.zeros(DRIFT_TABLE_LENGTH)
    k = np.zeros(DRIFT_TABLE_LENGTH)
    _check(lib().cuco_device_drift_table(device, a_begin, a_max, omega0, omega_lambda,
                                         hubble, _dp(d), _dp(k)), "cuco_device_drift_table")
    return d, k


def table_factor(table, a_begin, a_max, loga0, loga1) -> float:
    """get_drift_factor / get_gravkick_factor (driftfac.c:66-126) in ln a."""
    t = np.ascontiguousarray(table, np.float64)
    return lib().cuco_table_factor(_dp(t), a_begin, a_max, loga0, loga1)


def kdk_factors(drift, kick, a_begin, a_max, a_start, a_end, nsteps):
    """(nsteps, 3) array of (kick to mid, drift, kick from mid) per step."""
    f = np.zeros((nsteps, 3))
    d = np.ascontiguousarray(drift, np.float64)
    k = np.ascontiguousarray(kick, np.float64)
    _check(lib().cuco_kdk_factors(_dp(d), _dp(k), a_begin, a_max, a_start,
                                  a_end, nsteps, _dp(f)), "cuco_kdk_factors")
    return f


TIMEBASE = 1 << 28  # allvars.h:11


def write_snapshot(path: str, header: Header, P: np.ndarray) -> None:
    """savepositions (io.c:9-40): HEADER + PARTICLE[NumPart]."""
    P = np.ascontiguousarray(P, PARTICLE_DTYPE)
    h = Header(*[getattr(header, f) for f, _ in Header._fields_])
    h.NumPart = P.shape[0]
    _check(lib().cuco_write_snapshot(os.fsencode(path), C.byref(h),
                                     P.ctypes.data_as(C.c_void_p)),
           "cuco_write_snapshot")


def read_snapshot(path: str, header_only: bool = False):
    """read_ic (read_ic.c:9-45): returns (Header, PARTICLE array or None)."""
    h = Header()
    _check(lib().cuco_read_snapshot(os.fsencode(path), C.byref(h), None, 0),
           "cuco_read_snapshot")
    if header_only:
        return h, None
    P = np.zeros(h.NumPart, PARTICLE_DTYPE)
    _check(lib().cuco_read_snapshot(os.fsencode(path), C.byref(h),
                                    P.ctypes.data_as(C.c_void_p), P.shape[0]),
           "cuco_read_snapshot")
    return h, P


def checkpoint_info(path: str) -> dict:
    """The checked header of a checkpoint file (host only)."""
    h = CheckpointHeader()
    _check(lib().cuco_checkpoint_info(os.fsencode(path), C.byref(h)), "cuco_checkpoint_info")
    out = {f: getattr(h, f) for f, _ in CheckpointHeader._fields_
           if f not in ("params", "state")}
    out["params"] = {f: getattr(h.params, f) for f, _ in Timeline._fields_}
    out["state"] = h.state.as_dict()
    out["bytes"] = int(lib().cuco_checkpoint_bytes(C.byref(h)))
    return out


def init_records(P: np.ndarray, time_begin: float) -> None:
    """The particle part of init() (init.c:27-36), in place: Vel *= a^(3/2)
    at a = time_begin; GravAccel, GravPM and the Ti fields set to 0."""
    if P.dtype != PARTICLE_DTYPE or not P.flags.c_contiguous:
        raise ValueError("expected a contiguous PARTICLE array")
    _check(lib().cuco_init_records(P.ctypes.data_as(C.c_void_p), P.shape[0],
                                   time_begin), "cuco_init_records")


def read_output_list(path: str, max_times: int = 500) -> np.ndarray:
    """read_outputlist (begrun.c:80-102); MAXLEN_OUTPUTLIST is 500."""
    t = np.zeros(max_times)
    n = C.c_int32(0)
    _check(lib().cuco_read_output_list(os.fsencode(path), _dp(t), max_times,
                                       C.byref(n)), "cuco_read_output_list")
    return t[:n.value].copy()


def find_next_output_time(times, time_begin, time_max, timebase_interval,
                          ti_curr) -> int:
    """find_next_outputtime (run.c:75-104)."""
    t = np.ascontiguousarray(times, np.float64)
    return int(lib().cuco_find_next_output_time(_dp(t), t.shape[0], time_begin,
                                                time_max, timebase_interval,
                                                ti_curr))


class DeviceBuffer:
    """A raw device allocation (allocateArray/freeArray, gravlist.cu:33-43)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _check(lib().cuco_device_alloc(C.byref(p), nbytes), "cuco_device_alloc")
        self.ptr = p
        self.nbytes = nbytes

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().cuco_copy_to_device(self.ptr, arr.ctypes.data_as(C.c_void_p),
                                         arr.nbytes), "cuco_copy_to_device")

    def download(self, arr: np.ndarray) -> np.ndarray:
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        _check(lib().cuco_copy_from_device(arr.ctypes.data_as(C.c_void_p),
                                           self.ptr, arr.nbytes),
               "cuco_copy_from_device")
        return arr

    def copy_from(self, other: "DeviceBuffer", nbytes: int) -> None:
        _check(lib().cuco_copy_device_to_device(self.ptr, other.ptr, nbytes),
               "cuco_copy_device_to_device")

    def free(self) -> None:
        if self.ptr:
            _check(lib().cuco_device_free(self.ptr), "cuco_device_free")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def radix_sort(keys: np.ndarray, values: np.ndarray | None, key_bits: int):
    """RadixSort::sort on the device (radixsort.h:170-187); returns copies."""
    keys = np.ascontiguousarray(keys, np.uint32)
    n = keys.shape[0]
    kb = DeviceBuffer(max(4, keys.nbytes))
    kb.upload(keys)
    vb = None
    if values is not None:
        values = np.ascontiguousarray(values, np.uint32)
        vb = DeviceBuffer(max(4, values.nbytes))
        vb.upload(values)
    _check(lib().cuco_radix_sort(kb.ptr, vb.ptr if vb else None, n, key_bits),
           "cuco_radix_sort")
    ko = kb.download(np.empty(n, np.uint32))
    vo = vb.download(np.empty(n, np.uint32)) if vb else None
    return ko, vo


def fft3d_r2c(a: np.ndarray, precision: int = FP64) -> np.ndarray:
    """Device r2c (replaces cufftExecR2C, pm_periodic.c:140)."""
    M = a.shape[0]
    rt = np.float64 if precision == FP64 else np.float32
    ct = np.complex128 if precision == FP64 else np.complex64
    a = np.ascontiguousarray(a, rt)
    out = np.empty((M, M, M // 2 + 1), ct)
    bi, bo = DeviceBuffer(a.nbytes), DeviceBuffer(out.nbytes)
    bi.upload(a)
    _check(lib().cuco_fft3d_r2c(M, precision, bi.ptr, bo.ptr), "cuco_fft3d_r2c")
    return bo.download(out)


def fft3d_c2r(c: np.ndarray, precision: int = FP64) -> np.ndarray:
    """Device c2r (replaces cufftExecC2R, pm_periodic.c:201); unnormalized."""
    M = c.shape[0]
    rt = np.float64 if precision == FP64 else np.float32
    ct = np.complex128 if precision == FP64 else np.complex64
    c = np.ascontiguousarray(c, ct)
    out = np.empty((M, M, M), rt)
    bi, bo = DeviceBuffer(c.nbytes), DeviceBuffer(out.nbytes)
    bi.upload(c)
    _check(lib().cuco_fft3d_c2r(M, precision, bi.ptr, bo.ptr), "cuco_fft3d_c2r")
    return bo.download(out)


class Engine:
    """One engine handle: device-resident particles and force buffers."""

    def __init__(self, pm_grid: int, box_size: float, G: float, asmth: float,
                 rcut: float, soft_h: float, max_particles: int, device: int = 0):
        self.params = Params(pm_grid, 0, box_size, G, asmth, rcut, soft_h)
        h = C.c_void_p()
        _check(lib().cuco_create(C.byref(self.params), int(max_particles),
                                 device, C.byref(h)), "cuco_create")
        self.h = h
        self.M = pm_grid
        self.n = 0

    # -------------------------------------------------------- lifecycle
    def close(self) -> None:
        if getattr(self, "h", None):
            lib().cuco_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_softening(self, soft_h: float) -> None:
        _check(lib().cuco_set_softening(self.h, soft_h), "cuco_set_softening")

    def set_params(self, G: float, asmth: float, rcut: float, soft_h: float,
                   box: int = -1) -> None:
        """setSimParam: new G, split scale, cut and softening (box -1: all)."""
        p = Params(self._grid_of(box), 0, self._size_of(box), G, asmth, rcut, soft_h)
        _check(lib().cuco_set_params(self.h, box, C.byref(p)), "cuco_set_params")

    def _grid_of(self, box: int) -> int:
        return self.M

    def _size_of(self, box: int) -> float:
        return self.params.box_size

    # -------------------------------------------------------- state
    def upload(self, pos, vel=None, mass=None, uniform_mass: float = 0.0) -> None:
        pos = _xyz(pos)
        n = pos.shape[0]
        vel = _xyz(vel, n) if vel is not None else None
        if mass is not None:
            mass = np.ascontiguousarray(mass, np.float64)
            if mass.shape != (n,):
                raise ValueError("mass must have shape (n,)")
        _check(lib().cuco_upload_particles(self.h, n, _dp(pos), _dp(vel),
                                           _dp(mass), uniform_mass),
               "cuco_upload_particles")
        self.n = n

    def download_particles(self):
        pos = np.empty((self.n, 3))
        vel = np.empty((self.n, 3))
        _check(lib().cuco_download_particles(self.h, _dp(pos), _dp(vel)),
               "cuco_download_particles")
        return pos, vel

    def accelerations(self, pp: bool = True, pm: bool = True):
        app = np.empty((self.n, 3)) if pp else None
        apm = np.empty((self.n, 3)) if pm else None
        _check(lib().cuco_download_accel(self.h, _dp(app), _dp(apm)),
               "cuco_download_accel")
        if pp and pm:
            return app, apm
        return app if pp else apm

    # -------------------------------------------------------- forces
    def gravity_linklist(self, precision: int = FP64) -> None:
        _check(lib().cuco_pp_force(self.h, precision), "cuco_pp_force")

    def pmforce_periodic(self, precision: int = FP64) -> None:
        _check(lib().cuco_pm_force(self.h, precision), "cuco_pm_force")

    def compute_accelerations(self, precision: int = FP64, do_pm: bool = True):
        _check(lib().cuco_compute_accelerations(self.h, precision, int(do_pm)),
               "cuco_compute_accelerations")

    # -------------------------------------------------------- integrator
    def move_particles(self, drift_factor: float) -> None:
        _check(lib().cuco_drift(self.h, drift_factor), "cuco_drift")

    def kick(self, kick_pp: float, kick_pm: float) -> None:
        _check(lib().cuco_kick(self.h, kick_pp, kick_pm), "cuco_kick")

    def kdk_steps(self, factors, precision: int = FP64, use_graph: bool = False):
        f = np.ascontiguousarray(factors, np.float64).reshape(-1, 3)
        _check(lib().cuco_kdk_steps(self.h, precision, f.shape[0], _dp(f),
                                    int(use_graph)), "cuco_kdk_steps")

    def download_records(self) -> np.ndarray:
        """The particles as PARTICLE records (savepositions' P[])."""
        P = np.zeros(self.n, PARTICLE_DTYPE)
        _check(lib().cuco_download_records(self.h, P.ctypes.data_as(C.c_void_p)),
               "cuco_download_records")
        return P

    def upload_records(self, P: np.ndarray) -> None:
        P = np.ascontiguousarray(P, PARTICLE_DTYPE)
        _check(lib().cuco_upload_records(self.h, P.ctypes.data_as(C.c_void_p),
                                         P.shape[0]), "cuco_upload_records")
        self.n = P.shape[0]

    # -------------------------------------------------------- integer timeline
    def timeline_init(self, **kw) -> None:
        """init + init_drift_table; keywords are the Timeline fields."""
        t = Timeline(**kw)
        _check(lib().cuco_timeline_init(self.h, C.byref(t)), "cuco_timeline_init")

    def timeline_run(self, precision: int = FP64, max_steps: int = 1,
                     ti_stop: int = -1) -> dict:
        """Up to max_steps passes of run()'s body; returns the state."""
        s = TimelineState()
        _check(lib().cuco_timeline_run(self.h, precision, max_steps, ti_stop,
                                       C.byref(s)), "cuco_timeline_run")
        return s.as_dict()

    def checkpoint_write(self, path: str) -> None:
        """Particles, velocities, last accelerations and integer timeline."""
        _check(lib().cuco_checkpoint_write(self.h, os.fsencode(path)), "cuco_checkpoint_write")

    def checkpoint_read(self, path: str) -> None:
        """Restores a checkpoint written by a handle with the same parameters."""
        _check(lib().cuco_checkpoint_read(self.h, os.fsencode(path)), "cuco_checkpoint_read")
        self.n = int(checkpoint_info(path)["n"])

    def timesteps(self):
        b = np.empty(self.n, np.int32)
        e = np.empty(self.n, np.int32)
        _check(lib().cuco_get_timesteps(self.h, b.ctypes.data_as(C.POINTER(C.c_int32)),
                                        e.ctypes.data_as(C.POINTER(C.c_int32))),
               "cuco_get_timesteps")
        return b, e

    # -------------------------------------------------------- compat shims
    def pmforce_particles(self, P: np.ndarray, precision: int = FP64) -> None:
        """pmforce_periodic() on a PARTICLE array (writes GravPM in place)."""
        assert P.dtype == PARTICLE_DTYPE and P.flags.c_contiguous
        _check(lib().cuco_pmforce_periodic(self.h, P.ctypes.data_as(C.c_void_p),
                                           P.shape[0], precision),
               "cuco_pmforce_periodic")
        self.n = P.shape[0]

    def gravity_particles(self, P: np.ndarray, precision: int = FP64) -> None:
        """gravity_linklist_cuda() on a PARTICLE array (writes GravAccel)."""
        assert P.dtype == PARTICLE_DTYPE and P.flags.c_contiguous
        _check(lib().cuco_gravity_linklist(self.h, P.ctypes.data_as(C.c_void_p),
                                           P.shape[0], precision),
               "cuco_gravity_linklist")
        self.n = P.shape[0]

    # -------------------------------------------------------- inspection
    def sorted_order(self):
        keys = np.empty(self.n, np.uint32)
        idx = np.empty(self.n, np.uint32)
        _check(lib().cuco_get_sorted(self.h, keys.ctypes.data_as(C.POINTER(C.c_uint32)),
                                     idx.ctypes.data_as(C.POINTER(C.c_uint32))),
               "cuco_get_sorted")
        return keys, idx

    def cell_table(self):
        nc = self.M ** 3
        s = np.empty(nc, np.uint32)
        e = np.empty(nc, np.uint32)
        _check(lib().cuco_get_cell_table(self.h, s.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         e.ctypes.data_as(C.POINTER(C.c_uint32))),
               "cuco_get_cell_table")
        return s, e

    def shortrange_table(self):
        t = np.empty(NTAB, np.float32)
        _check(lib().cuco_get_shortrange_table(self.h, t.ctypes.data_as(C.POINTER(C.c_float))),
               "cuco_get_shortrange_table")
        return t

    def pm_grid(self, which: int) -> np.ndarray:
        out = np.empty((self.M,) * 3)
        _check(lib().cuco_get_pm_grid(self.h, which, _dp(out)), "cuco_get_pm_grid")
        return out

    def count_pairs(self):
        a = C.c_uint64(0)
        b = C.c_uint64(0)
        _check(lib().cuco_count_pairs(self.h, C.byref(a), C.byref(b)),
               "cuco_count_pairs")
        return int(a.value), int(b.value)

    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().cuco_get_tuning(self.h, C.byref(t)), "cuco_get_tuning")
        return t.as_dict()

    def set_tuning(self, pp_block_shift: int, pp_workgroups_per_cu: int,
                   pm_overlap: int = 0, pp_symmetric: int = 1, pp_cell_level: int = 0,
                   fft_split: int = 0, pp_wide_tiles: int = 0) -> None:
        t = Tuning(pp_block_shift, pp_workgroups_per_cu, int(pm_overlap),
                   int(pp_symmetric), int(pp_cell_level), int(fft_split),
                   int(pp_wide_tiles))
        _check(lib().cuco_set_tuning(self.h, C.byref(t)), "cuco_set_tuning")

    def autotune(self):
        """Picks the FP32 tuning (pair kernel, FFT split, PM overlap) by
        timing the candidates on the current particles."""
        t = Tuning()
        ms = C.c_double(0)
        _check(lib().cuco_autotune(self.h, C.byref(t), C.byref(ms)), "cuco_autotune")
        return t.as_dict(), ms.value

    def set_timing(self, on: bool) -> None:
        _check(lib().cuco_set_timing(self.h, int(on)), "cuco_set_timing")

    def stage_times(self, reset: bool = True):
        ms = np.zeros(len(STAGES))
        ev = C.c_int32(0)
        _check(lib().cuco_get_stage_times(self.h, _dp(ms), len(STAGES),
                                          C.byref(ev), int(reset)),
               "cuco_get_stage_times")
        return dict(zip(STAGES, ms.tolist())), int(ev.value)

    def synchronize(self) -> None:
        _check(lib().cuco_synchronize(self.h), "cuco_synchronize")


class Ensemble(Engine):
    """Independent periodic boxes advanced together (cuco_ensemble_create).

    ``boxes`` is a list of dicts with the keys of :class:`Engine`'s
    parameters (pm_grid, box_size, G, asmth, rcut, soft_h); ``n_per_box``
    the particle count of each box.  Particle arrays hold box 0's particles,
    then box 1's, and so on.
    """

    def __init__(self, boxes, n_per_box, device: int = 0):
        nb = len(boxes)
        if nb == 0 or len(n_per_box) != nb:
            raise ValueError("need one particle count per box")
        arr = (Params * nb)()
        for k, b in enumerate(boxes):
            arr[k] = Params(int(b["pm_grid"]), 0, float(b["box_size"]), float(b["G"]),
                            float(b["asmth"]), float(b["rcut"]), float(b["soft_h"]))
        counts = (C.c_int64 * nb)(*[int(x) for x in n_per_box])
        h = C.c_void_p()
        _check(lib().cuco_ensemble_create(arr, counts, nb, device, C.byref(h)),
               "cuco_ensemble_create")
        self.h = h
        self.box_params = arr
        self.params = arr[0]
        self.grids = [int(b["pm_grid"]) for b in boxes]
        self.M = self.grids[0]
        self.counts = [int(x) for x in n_per_box]
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        self.n = 0

    def _grid_of(self, box: int) -> int:
        return self.grids[max(box, 0)]

    def _size_of(self, box: int) -> float:
        return self.box_params[max(box, 0)].box_size

    def set_params(self, G, asmth, rcut, soft_h, box: int = -1) -> None:
        boxes = range(len(self.grids)) if box < 0 else [box]
        for b in boxes:
            p = Params(self.grids[b], 0, self.box_params[b].box_size, G, asmth, rcut, soft_h)
            _check(lib().cuco_set_params(self.h, b, C.byref(p)), "cuco_set_params")

    def size(self):
        nb = C.c_int32(0)
        _check(lib().cuco_ensemble_size(self.h, C.byref(nb), None), "cuco_ensemble_size")
        counts = (C.c_int64 * nb.value)()
        _check(lib().cuco_ensemble_size(self.h, C.byref(nb), counts), "cuco_ensemble_size")
        return nb.value, list(counts)

    def box_slice(self, b: int) -> slice:
        return slice(int(self.offsets[b]), int(self.offsets[b + 1]))

    def cell_table(self):
        nc = sum(m ** 3 for m in self.grids)
        s = np.empty(nc, np.uint32)
        e = np.empty(nc, np.uint32)
        _check(lib().cuco_get_cell_table(self.h, s.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         e.ctypes.data_as(C.POINTER(C.c_uint32))),
               "cuco_get_cell_table")
        return s, e

    def pm_grid(self, which: int):
        """List of each box's mesh."""
        out = np.empty(sum(m ** 3 for m in self.grids))
        _check(lib().cuco_get_pm_grid(self.h, which, _dp(out)), "cuco_get_pm_grid")
        res, o = [], 0
        for m in self.grids:
            res.append(out[o:o + m ** 3].reshape(m, m, m))
            o += m ** 3
        return res
