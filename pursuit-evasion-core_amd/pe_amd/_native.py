"""ctypes binding of libpe_amd.so (declarations in include/pe_amd.h).

The library is the only compute path: if it is missing or no gfx950 device is
visible, calls raise.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpe_amd.so"

PE_OK = 0
PE_ERR_INVALID = -1
PE_ERR_MEMORY = -2
PE_ERR_NONFINITE = -3
PE_ERR_HIP = -4
PE_ERR_NODEVICE = -5

_LIB = None


class NativeLibraryMissing(ImportError):
    pass


def lib():
    """Load libpe_amd.so (once) and declare its C ABI."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run "
                "`python pursuit-evasion-core_amd/build.py`"
            )
        L = ctypes.CDLL(str(LIB_PATH))
        _declare(L)
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    """Map a PE_ERR_* return code to the exception the reference raises."""
    if rc == PE_OK:
        return
    message = lib().pe_last_error().decode(errors="replace")
    if rc == PE_ERR_INVALID:
        raise ValueError(message)
    if rc == PE_ERR_MEMORY:
        raise MemoryError(message)
    raise RuntimeError(message or f"pe_amd error {rc}")


def ptr(a):
    """Pointer to a C-contiguous numpy array (or NULL for None)."""
    if a is None:
        return None
    if not a.flags.c_contiguous:
        raise ValueError("C-contiguous array required")
    return a.ctypes.data_as(ctypes.c_void_p)


class HJI5DConfig(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_int32),
        ("headings", ctypes.c_int32),
        ("p_speeds", ctypes.c_int32),
        ("e_speeds", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("extent", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("radius", ctypes.c_double),
        ("p_min", ctypes.c_double),
        ("p_max", ctypes.c_double),
        ("e_min", ctypes.c_double),
        ("e_max", ctypes.c_double),
        ("p_acceleration", ctypes.c_double),
        ("e_acceleration", ctypes.c_double),
        ("p_curvature", ctypes.c_double),
        ("e_curvature", ctypes.c_double),
        ("exit_radius", ctypes.c_double),
        ("capture", ctypes.c_int32),
        ("vp_chunk", ctypes.c_int32),
        ("p_length", ctypes.c_double),
        ("p_width", ctypes.c_double),
        ("e_length", ctypes.c_double),
        ("e_width", ctypes.c_double),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


PE_HJI5D_LAYOUT_PLANE = 1
PE_HJI5D_GENERIC = 2
PE_HJI5D_NO_PREFETCH = 4
PE_HJI5D_PSI_OUTER = 8
PE_HJI5D_PSI_SINGLE = 16


class HJI5DStats(ctypes.Structure):
    _fields_ = [
        ("states", ctypes.c_int64),
        ("internal_steps", ctypes.c_int64),
        ("time", ctypes.c_double),
        ("cfl_dt_limit", ctypes.c_double),
        ("maximum_increase", ctypes.c_double),
        ("last_residual", ctypes.c_double),
        ("update_seconds", ctypes.c_double),
        ("stage_ms", ctypes.c_double * 2),
        ("stage_launches", ctypes.c_int64 * 2),
    ]


class SLCar(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("actions", ctypes.c_int32),
        ("states", ctypes.c_void_p),
        ("successors", ctypes.c_void_p),
        ("weights", ctypes.c_void_p),
        ("allowed", ctypes.c_void_p),
        ("position", ctypes.c_void_p),
        ("positions", ctypes.c_int32),
        ("clearance", ctypes.c_void_p),
    ]


class CarAxes(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 4), ("axis", ctypes.c_void_p * 4)]


class SLGame(ctypes.Structure):
    _fields_ = [
        ("p", SLCar),
        ("e", SLCar),
        ("inside", ctypes.c_void_p),
        ("near", ctypes.c_void_p),
        ("exit_value", ctypes.c_float),
        ("p_length", ctypes.c_float),
        ("p_width", ctypes.c_float),
        ("e_length", ctypes.c_float),
        ("e_width", ctypes.c_float),
    ]


class SLStats(ctypes.Structure):
    _fields_ = [
        ("sweeps", ctypes.c_int64),
        ("maps", ctypes.c_int32),
        ("lds_e_side", ctypes.c_int32),
        ("pairs", ctypes.c_int64),
        ("active_pairs", ctypes.c_int64),
        ("update_seconds", ctypes.c_double),
        ("last_residual", ctypes.c_double),
        ("p_ms", ctypes.c_double),
        ("e_ms", ctypes.c_double),
        ("p_launches", ctypes.c_int64),
        ("e_launches", ctypes.c_int64),
    ]


class MLPConfig(ctypes.Structure):
    _fields_ = [
        ("extra", ctypes.c_int32),
        ("rff", ctypes.c_int32),
        ("hidden", ctypes.c_int32),
        ("layers", ctypes.c_int32),
        ("seed", ctypes.c_uint32),
        ("rff_scale", ctypes.c_float),
        ("state_shift", ctypes.c_float * 8),
        ("state_scale", ctypes.c_float * 8),
        ("learning_rate", ctypes.c_float),
        ("weight_decay", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("under_weight", ctypes.c_float),
    ]


PE_SL_EARLY_EXIT = 1
PE_SL_ORDERED = 2
PE_SL_GRAPH = 4
PE_SL_GLOBAL_GATHER = 8
PE_SL_WIDE_IDS = 16
PE_SL_ONE_PASS = 32
PE_SL_U_MAJOR = 64
PE_SL_FREEZE = 128
PE_SL_STENCIL_P_SIDE = 256
PE_SL_SEPARATE_P_SIDE = 512

class World8DConfig(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_int32),
        ("headings", ctypes.c_int32),
        ("p_speeds", ctypes.c_int32),
        ("e_speeds", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("linear_speed_boundary", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("extent", ctypes.c_double),
        ("arena", ctypes.c_double),
        ("radius", ctypes.c_double),
        ("p_min", ctypes.c_double),
        ("p_max", ctypes.c_double),
        ("e_min", ctypes.c_double),
        ("e_max", ctypes.c_double),
        ("p_accel", ctypes.c_double),
        ("e_accel", ctypes.c_double),
        ("p_curvature", ctypes.c_double),
        ("e_curvature", ctypes.c_double),
        ("cfl", ctypes.c_double),
    ]


class World8DStats(ctypes.Structure):
    _fields_ = [
        ("states", ctypes.c_int64),
        ("internal_steps", ctypes.c_int64),
        ("time", ctypes.c_double),
        ("max_internal_dt", ctypes.c_double),
        ("update_seconds", ctypes.c_double),
        ("probes", ctypes.c_int64),
        ("kernel", ctypes.c_int32),
        ("fast_division", ctypes.c_int32),
    ]


PE_WORLD8D_GENERIC = 1
PE_WORLD8D_SLAB_256 = 2
PE_WORLD8D_SLAB_1024 = 4
PE_RTABLE_NOALLOC_MAX = 65536


P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
F64 = ctypes.c_double

# name -> (restype, argtypes); the exported symbols of include/pe_amd.h.
SIGNATURES = {
    "pe_last_error": (ctypes.c_char_p, []),
    "pe_abi_version": (ctypes.c_int, []),
    "pe_abi_struct_size": (ctypes.c_int, [ctypes.c_char_p]),
    "pe_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, P, P]),
    "pe_copy_bandwidth": (ctypes.c_int, [I64, I32, P]),
    "pe_hji5d_create": (ctypes.c_int, [ctypes.POINTER(HJI5DConfig), P]),
    "pe_hji5d_advance_to": (ctypes.c_int, [P, F64, I64]),
    "pe_hji5d_step": (ctypes.c_int, [P, F64, I64]),
    "pe_hji5d_run_to_residual": (ctypes.c_int, [P, F64, I64, I64, P]),
    "pe_hji5d_read": (ctypes.c_int, [P, P, P]),
    "pe_hji5d_resolved": (ctypes.c_int, [P, P]),
    "pe_hji5d_trace": (ctypes.c_int, [P, P, I64, P]),
    "pe_hji5d_get_stats": (ctypes.c_int, [P, ctypes.POINTER(HJI5DStats)]),
    "pe_hji5d_set_profiling": (ctypes.c_int, [P, I32]),
    "pe_hji5d_destroy": (None, [P]),
    "pe_geom_clearance": (ctypes.c_int, [P, I64, P, I32, F64, P, P]),
    "pe_geom_backups": (ctypes.c_int, [P, I64, P, I32, F64, P, P]),
    "pe_geom_actions": (ctypes.c_int, [P, I64, P, I32, F64, P, I32, F64, P, P, P]),
    "pe_geom_backups_multi": (ctypes.c_int, [P, I64, P, I32, P, P, P, P, P]),
    "pe_geom_actions_multi": (ctypes.c_int, [P, I64, P, I32, P, I32, F64, P, P, P, P, P, P]),
    "pe_geom_connections": (ctypes.c_int, [P, P, I64, P, I32, F64, P, P]),
    "pe_grid_stencil": (ctypes.c_int, [I32, P, P, P, P, P, ctypes.c_uint32, F64, P, I64, P, P,
                                       P]),
    "pe_car_stencil": (ctypes.c_int, [ctypes.POINTER(CarAxes), P, P, I64, P, I32, F64, P, P, P,
                                      P, P, P]),
    "pe_sl_create": (ctypes.c_int, [ctypes.POINTER(SLGame), I32, P]),
    "pe_sl_create_ex": (ctypes.c_int, [ctypes.POINTER(SLGame), I32, I32, P]),
    "pe_sl_initialize_f64": (ctypes.c_int, [P, P]),
    "pe_sl_read_f64": (ctypes.c_int, [P, I32, P]),
    "pe_sl_initialize": (ctypes.c_int, [P, P]),
    "pe_sl_sweep": (ctypes.c_int, [P, I32, I32]),
    "pe_sl_run": (ctypes.c_int, [P, F64, I32, I32, I32, P, P]),
    "pe_sl_read": (ctypes.c_int, [P, I32, P]),
    "pe_sl_query": (ctypes.c_int, [P, I32, I64, P, P, P, P, P]),
    "pe_sl_trace": (ctypes.c_int, [P, P, I64, P]),
    "pe_sl_get_stats": (ctypes.c_int, [P, ctypes.POINTER(SLStats)]),
    "pe_sl_set_profiling": (ctypes.c_int, [P, I32]),
    "pe_sl_set_lds_threads": (ctypes.c_int, [P, I32]),
    "pe_sl_destroy": (None, [P]),
    "pe_sl_warm_start": (ctypes.c_int, [P, P, ctypes.c_float]),
    "pe_sl_order_from_mlp": (ctypes.c_int, [P, P]),
    "pe_sl_pair_inputs": (ctypes.c_int, [P, I32, I64, I64, P]),
    "pe_history_create": (ctypes.c_int, [P, I64, I64, P]),
    "pe_history_query": (ctypes.c_int, [P, I64, I32, P, P, P, ctypes.c_float, P]),
    "pe_history_destroy": (None, [P]),
    "pe_mlp_create": (ctypes.c_int, [ctypes.POINTER(MLPConfig), P]),
    "pe_mlp_param_count": (ctypes.c_int, [P, P]),
    "pe_mlp_get_params": (ctypes.c_int, [P, P, I64]),
    "pe_mlp_set_params": (ctypes.c_int, [P, P, I64]),
    "pe_mlp_forward": (ctypes.c_int, [P, P, I64, P]),
    "pe_mlp_features": (ctypes.c_int, [P, P, I64, P, P]),
    "pe_mlp_gradients": (ctypes.c_int, [P, P, P, I64, P, P]),
    "pe_mlp_set_dataset": (ctypes.c_int, [P, P, P, I64]),
    "pe_mlp_train": (ctypes.c_int, [P, I32, I32, ctypes.c_uint32, P]),
    "pe_mlp_gemm_check": (ctypes.c_int, [P, P, I32, I32, I32, I32, P]),
    "pe_mlp_destroy": (None, [P]),
    "pe_world8d_create": (ctypes.c_int, [ctypes.POINTER(World8DConfig), P]),
    "pe_world8d_set_field": (ctypes.c_int, [P, P, P, F64]),
    "pe_world8d_advance_to": (ctypes.c_int, [P, F64, I64]),
    "pe_world8d_read": (ctypes.c_int, [P, P, P]),
    "pe_world8d_count_nonfinite": (ctypes.c_int, [P, P]),
    "pe_world8d_set_probes": (ctypes.c_int, [P, P, P, I64]),
    "pe_world8d_probe": (ctypes.c_int, [P, P]),
    "pe_world8d_get_stats": (ctypes.c_int, [P, ctypes.POINTER(World8DStats)]),
    "pe_world8d_destroy": (None, [P]),
    "pe_world8d_check_division": (ctypes.c_int, [P, I32, I32, I32, P]),
    "pe_world8d_set_clearance": (ctypes.c_int, [P, P, P]),
    "pe_world8d_set_step_limit": (ctypes.c_int, [P, F64]),
    "pe_rtable_create": (ctypes.c_int, [P, I32, P, P, P, P, P, P]),
    "pe_rtable_field": (ctypes.c_int, [P, P, P, I64, P]),
    "pe_rtable_value": (ctypes.c_int, [P, P, I64, F64, P, P]),
    "pe_rtable_field_noalloc": (ctypes.c_int, [P, P, P, I64, P]),
    "pe_rtable_destroy": (None, [P]),
}


STRUCTS = {
    "pe_car_axes": CarAxes,
    "pe_hji5d_config": HJI5DConfig,
    "pe_hji5d_stats": HJI5DStats,
    "pe_sl_car": SLCar,
    "pe_sl_game": SLGame,
    "pe_sl_stats": SLStats,
    "pe_mlp_config": MLPConfig,
    "pe_world8d_config": World8DConfig,
    "pe_world8d_stats": World8DStats,
}


def _declare(L):
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def device_info():
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int32()
    hbm = ctypes.c_int64()
    check(lib().pe_device_info(name, 256, ctypes.byref(cus), ctypes.byref(hbm)))
    return dict(name=name.value.decode(), compute_units=cus.value, hbm_bytes=hbm.value)


def copy_bandwidth(nbytes, repeats=10):
    gbps = ctypes.c_double()
    check(lib().pe_copy_bandwidth(int(nbytes), int(repeats), ctypes.byref(gbps)))
    return gbps.value


def grid_stencil(states, n, lo, hi, step, period=None, periodic=0, slack=1e-10):
    """pe_grid_stencil: multilinear stencils of states [m, d] on a uniform
    grid; returns node index [m, 2^d] int64, weights [m, 2^d] float64 and
    validity [m] bool."""
    d = len(n)
    q = np.ascontiguousarray(np.asarray(states, np.float64).reshape(-1, d))
    dims = np.ascontiguousarray(n, np.int32)
    lo_, hi_, st = (np.ascontiguousarray(a, np.float64) for a in (lo, hi, step))
    per = None if period is None else np.ascontiguousarray(period, np.float64)
    idx = np.empty((len(q), 1 << d), np.int64)
    w = np.empty((len(q), 1 << d), np.float64)
    valid = np.empty(len(q), np.uint8)
    check(lib().pe_grid_stencil(d, ptr(dims), ptr(lo_), ptr(hi_), ptr(st), ptr(per),
                                int(periodic), float(slack), ptr(q), len(q), ptr(idx), ptr(w),
                                ptr(valid)))
    return idx, w, valid.astype(bool)


def car_stencil(axes, dense_to_compact, states, geometry=None):
    """pe_car_stencil: ids [m,16] int32, weights [m,16] float32, valid [m]
    bool, mass [m] float64 and the count of rows whose vertices were all
    illegal, for successor states [m,4].  geometry = (obstacles [nc,3],
    half-width, params[7]) enables the nearest-supported fallback."""
    ax = [np.ascontiguousarray(a, np.float64) for a in axes]
    spec = CarAxes()
    for k, a in enumerate(ax):
        spec.n[k] = len(a)
        spec.axis[k] = a.ctypes.data
    d2c = np.ascontiguousarray(dense_to_compact, np.int32).reshape(-1)
    if d2c.size != int(np.prod([len(a) for a in ax])):
        raise ValueError("dense_to_compact does not match the axes")
    q = np.ascontiguousarray(states, np.float64).reshape(-1, 4)
    m = len(q)
    ids = np.empty((m, 16), np.int32)
    w = np.empty((m, 16), np.float32)
    valid = np.empty(m, np.uint8)
    mass = np.empty(m, np.float64)
    fallbacks = ctypes.c_int64()
    if geometry is None:
        obs, half, par = None, 1.0, None
    else:
        obs = np.ascontiguousarray(geometry[0], np.float64).reshape(-1, 3)
        half = float(geometry[1])
        par = np.ascontiguousarray(geometry[2], np.float64)
    check(lib().pe_car_stencil(ctypes.byref(spec), ptr(d2c), ptr(q), m, ptr(obs),
                               0 if obs is None else len(obs), half, ptr(par), ptr(ids), ptr(w),
                               ptr(valid), ptr(mass), ctypes.byref(fallbacks)))
    return ids, w, valid.astype(bool), mass, int(fallbacks.value)


class HJI5D:
    """Owning wrapper of a pe_hji5d handle (device-resident 5D solve)."""

    def __init__(self, cfg: HJI5DConfig):
        self._h = ctypes.c_void_p()
        check(lib().pe_hji5d_create(ctypes.byref(cfg), ctypes.byref(self._h)))
        self.cfg = cfg
        self.n = (
            cfg.positions * cfg.positions * cfg.headings * cfg.p_speeds * cfg.e_speeds
        )

    def close(self):
        if self._h:
            lib().pe_hji5d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def advance_to(self, goal: float, max_steps: int = 0):
        check(lib().pe_hji5d_advance_to(self._h, float(goal), int(max_steps)))

    def step(self, dt: float, count: int):
        check(lib().pe_hji5d_step(self._h, float(dt), int(count)))

    def run_to_residual(self, tol, max_steps, check_every=16) -> bool:
        flag = ctypes.c_int32()
        check(
            lib().pe_hji5d_run_to_residual(
                self._h, float(tol), int(max_steps), int(check_every), ctypes.byref(flag)
            )
        )
        return bool(flag.value)

    def read(self, field=True, arrival=True):
        f = np.empty(self.n, np.float32) if field is True else field
        a = np.empty(self.n, np.float32) if arrival is True else arrival
        f = None if f is False else f
        a = None if a is False else a
        check(lib().pe_hji5d_read(self._h, ptr(f), ptr(a)))
        return f, a

    def resolved(self) -> int:
        c = ctypes.c_int64()
        check(lib().pe_hji5d_resolved(self._h, ctypes.byref(c)))
        return c.value

    def trace(self):
        count = ctypes.c_int64()
        check(lib().pe_hji5d_trace(self._h, None, 0, ctypes.byref(count)))
        out = np.empty(count.value, np.float64)
        check(lib().pe_hji5d_trace(self._h, ptr(out), count.value, ctypes.byref(count)))
        return out

    def stats(self) -> dict:
        s = HJI5DStats()
        check(lib().pe_hji5d_get_stats(self._h, ctypes.byref(s)))
        return dict(
            states=s.states,
            internal_steps=s.internal_steps,
            time=s.time,
            cfl_dt_limit=s.cfl_dt_limit,
            maximum_increase=s.maximum_increase,
            last_residual=s.last_residual,
            update_seconds=s.update_seconds,
            stage_ms=list(s.stage_ms),
            stage_launches=list(s.stage_launches),
        )

    def set_profiling(self, on: bool):
        check(lib().pe_hji5d_set_profiling(self._h, int(bool(on))))


class SL:
    """Owning wrapper of a pe_sl handle: the 8D backup of one or more maps.

    `games` is a list of dicts with the numpy tables of each map (built by
    pe_amd.local_game.grid_game.game_tables); they are copied to the device
    at construction."""

    def __init__(self, games, precision=0):
        self._keep = []
        self.precision = precision
        arr = (SLGame * len(games))()
        for k, g in enumerate(games):
            for side in ("p", "e"):
                car = g[side]
                c = getattr(arr[k], side)
                c.n = len(car["states"])
                c.actions = car["allowed"].shape[1]
                for name, dtype in (("states", np.float64), ("successors", np.int32),
                                    ("weights", np.float32), ("allowed", np.uint8),
                                    ("position", np.int32)):
                    a = np.ascontiguousarray(car[name], dtype)
                    self._keep.append(a)
                    setattr(c, name, a.ctypes.data)
                c.positions = int(car["positions"])
                if car.get("clearance") is not None:
                    a = np.ascontiguousarray(car["clearance"], np.float64)
                    self._keep.append(a)
                    c.clearance = a.ctypes.data
            for name in ("inside", "near"):
                a = np.ascontiguousarray(g[name], np.uint8)
                self._keep.append(a)
                setattr(arr[k], name, a.ctypes.data)
            arr[k].exit_value = g["exit_value"]
            arr[k].p_length, arr[k].p_width = g["p_body"]
            arr[k].e_length, arr[k].e_width = g["e_body"]
        self._h = ctypes.c_void_p()
        check(lib().pe_sl_create_ex(arr, len(games), int(precision), ctypes.byref(self._h)))
        self._keep = []
        self.shapes = [(len(g["p"]["states"]), len(g["e"]["states"])) for g in games]

    def close(self):
        if self._h:
            lib().pe_sl_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def initialize(self, v0=None):
        dtype = np.float64 if (v0 is not None and np.asarray(
            v0[0] if isinstance(v0, (list, tuple)) else v0).dtype == np.float64) else np.float32
        a = None if v0 is None else np.ascontiguousarray(
            np.concatenate([np.ravel(v) for v in v0]) if isinstance(v0, (list, tuple)) else v0,
            dtype)
        fn = lib().pe_sl_initialize_f64 if dtype == np.float64 else lib().pe_sl_initialize
        check(fn(self._h, ptr(a)))

    def sweep(self, count=1, flags=PE_SL_EARLY_EXIT):
        check(lib().pe_sl_sweep(self._h, int(count), int(flags)))

    def run(self, tol, max_sweeps, check_every=1, flags=PE_SL_EARLY_EXIT):
        done, conv = ctypes.c_int32(), ctypes.c_int32()
        check(lib().pe_sl_run(self._h, float(tol), int(max_sweeps), int(check_every),
                              int(flags), ctypes.byref(done), ctypes.byref(conv)))
        return done.value, bool(conv.value)

    def read(self, index=0, dtype=None):
        """Field of map `index` [p.n, e.n]; float32, or float64 for FP64
        handles (or when dtype=np.float64 is asked for)."""
        np_, ne = self.shapes[index]
        dtype = dtype or (np.float64 if self.precision == 1 else np.float32)
        out = np.empty((np_, ne), dtype)
        fn = lib().pe_sl_read_f64 if out.dtype == np.float64 else lib().pe_sl_read
        check(fn(self._h, int(index), ptr(out)))
        return out

    def query(self, p_ids, p_weights, e_ids, e_weights, index=0):
        arrays = [np.ascontiguousarray(p_ids, np.int32), np.ascontiguousarray(p_weights, np.float32),
                  np.ascontiguousarray(e_ids, np.int32), np.ascontiguousarray(e_weights, np.float32)]
        n = len(arrays[0])
        if any(a.size != n * 16 for a in arrays):
            raise ValueError("query stencils [n, 16] required for both cars")
        out = np.empty(n, np.float32)
        check(lib().pe_sl_query(self._h, int(index), n, *[ptr(a) for a in arrays], ptr(out)))
        return out

    def trace(self):
        count = ctypes.c_int64()
        check(lib().pe_sl_trace(self._h, None, 0, ctypes.byref(count)))
        out = np.empty(count.value, np.float64)
        check(lib().pe_sl_trace(self._h, ptr(out), count.value, ctypes.byref(count)))
        return out.reshape(-1, len(self.shapes))

    def stats(self):
        s = SLStats()
        check(lib().pe_sl_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in SLStats._fields_}

    def set_profiling(self, on):
        check(lib().pe_sl_set_profiling(self._h, int(bool(on))))

    def set_lds_threads(self, threads):
        check(lib().pe_sl_set_lds_threads(self._h, int(threads)))

    def pair_inputs(self, first, count, index=0):
        out = np.empty((int(count), 11), np.float32)
        check(lib().pe_sl_pair_inputs(self._h, int(index), int(first), int(count), ptr(out)))
        return out

    def warm_start(self, mlp, delta):
        check(lib().pe_sl_warm_start(self._h, mlp._h, float(delta)))

    def order_from(self, mlp):
        """Stored command orders of every inside pair from the network's
        field (pe_sl_order_from_mlp); values are unchanged."""
        check(lib().pe_sl_order_from_mlp(self._h, mlp._h))


class History:
    """Owning wrapper of a pe_history handle: a packed float16 field history
    [frames, states] resident on the device, queried through product
    stencils."""

    def __init__(self, values):
        v = np.ascontiguousarray(values, np.float16)
        if v.ndim != 2:
            raise ValueError("history [frames, states] required")
        self.frames, self.states = v.shape
        self._h = ctypes.c_void_p()
        check(lib().pe_history_create(ptr(v.view(np.uint16)), self.frames, self.states,
                                      ctypes.byref(self._h)))

    def query(self, ids, weights, kept, exit_value):
        """[frames, n] sums over the terms of each query row."""
        ids = np.ascontiguousarray(ids, np.int64)
        n = ids.shape[0]
        terms = int(np.prod(ids.shape[1:])) if ids.ndim > 1 else 1
        w = np.ascontiguousarray(weights, np.float32).reshape(n, terms)
        k = np.ascontiguousarray(kept, np.uint8).reshape(n, terms)
        out = np.empty((self.frames, n), np.float32)
        check(lib().pe_history_query(self._h, n, terms, ptr(ids), ptr(w), ptr(k),
                                     float(exit_value), ptr(out)))
        return out

    def close(self):
        if self._h:
            lib().pe_history_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class World8D:
    """Owning wrapper of a pe_world8d handle: the dense 8D square game,
    fields resident on the device."""

    def __init__(self, cfg: World8DConfig):
        self._h = ctypes.c_void_p()
        check(lib().pe_world8d_create(ctypes.byref(cfg), ctypes.byref(self._h)))
        self.cfg = cfg
        n, h = cfg.positions, cfg.headings
        self.n = n ** 4 * h * h * cfg.p_speeds * cfg.e_speeds
        self.probe_count = 0

    def set_field(self, field, arrival=None, time=0.0):
        f = np.ascontiguousarray(field, np.float32).reshape(-1)
        a = None if arrival is None else np.ascontiguousarray(arrival, np.float32).reshape(-1)
        if f.size != self.n or (a is not None and a.size != self.n):
            raise ValueError("field size mismatch")
        check(lib().pe_world8d_set_field(self._h, ptr(f), ptr(a), float(time)))

    def advance_to(self, goal: float, max_steps: int = 0):
        check(lib().pe_world8d_advance_to(self._h, float(goal), int(max_steps)))

    def read(self, field=True, arrival=True):
        f = np.empty(self.n, np.float32) if field is True else field
        a = np.empty(self.n, np.float32) if arrival is True else arrival
        f = None if f is False else f
        a = None if a is False else a
        check(lib().pe_world8d_read(self._h, ptr(f), ptr(a)))
        return f, a

    def count_nonfinite(self) -> int:
        c = ctypes.c_int64()
        check(lib().pe_world8d_count_nonfinite(self._h, ctypes.byref(c)))
        return c.value

    def set_clearance(self, p_clearance, e_clearance):
        """Footprint-map mode: clearance grids [n, n, nh] of each car."""
        pc = np.ascontiguousarray(p_clearance, np.float32).reshape(-1)
        ec = np.ascontiguousarray(e_clearance, np.float32).reshape(-1)
        c = self.cfg
        if pc.size != c.positions ** 2 * c.headings or ec.size != pc.size:
            raise ValueError("clearance grids must be [positions, positions, headings]")
        check(lib().pe_world8d_set_clearance(self._h, ptr(pc), ptr(ec)))

    def set_step_limit(self, max_dt):
        check(lib().pe_world8d_set_step_limit(self._h, float(max_dt)))

    def set_probes(self, indices, weights):
        ii = np.ascontiguousarray(indices, np.uint64).reshape(-1, 256)
        ww = np.ascontiguousarray(weights, np.float64).reshape(-1, 256)
        if ii.shape != ww.shape:
            raise ValueError("probe indices and weights must match")
        check(lib().pe_world8d_set_probes(self._h, ptr(ii), ptr(ww), len(ii)))
        self.probe_count = len(ii)

    def probe(self):
        out = np.empty(self.probe_count, np.float64)
        check(lib().pe_world8d_probe(self._h, ptr(out)))
        return out

    def stats(self) -> dict:
        s = World8DStats()
        check(lib().pe_world8d_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in World8DStats._fields_}

    def close(self):
        if self._h:
            lib().pe_world8d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def world8d_check_division(divisors, lo=-90, hi=70):
    """Mismatch count of the world kernels' reciprocal division against the
    IEEE division over every numerator with exponent in [lo, hi), per
    divisor (pe_world8d_check_division)."""
    d = np.ascontiguousarray(divisors, np.float32)
    out = np.zeros(len(d), np.int64)
    check(lib().pe_world8d_check_division(ptr(d), len(d), int(lo), int(hi), ptr(out)))
    return out


class RelTable:
    """Owning wrapper of a pe_rtable handle: a 5D field history resident on
    the device, queried in batches."""

    def __init__(self, fields, axes, times):
        f = np.ascontiguousarray(fields, np.float32)
        frames = f.shape[0]
        shape = np.array([len(a) for a in axes], np.int32)
        lo = np.array([a[0] for a in axes], np.float64)
        hi = np.array([a[-1] for a in axes], np.float64)
        step = np.array([a[1] - a[0] for a in axes], np.float64)
        t = np.ascontiguousarray(times, np.float64)
        if f.size != int(np.prod(shape)) * frames or len(t) != frames:
            raise ValueError("History shape differs from axes and times")
        self._h = ctypes.c_void_p()
        check(lib().pe_rtable_create(ptr(f.reshape(-1)), frames, ptr(shape), ptr(lo), ptr(hi),
                                     ptr(step), ptr(t), ctypes.byref(self._h)))

    def field(self, states, budgets):
        s = np.ascontiguousarray(states, np.float64).reshape(-1, 5)
        b = np.ascontiguousarray(budgets, np.float64).reshape(-1)
        if len(b) != len(s):
            raise ValueError("one budget per state required")
        out = np.empty(len(s), np.float64)
        check(lib().pe_rtable_field(self._h, ptr(s), ptr(b), len(s), ptr(out)))
        return out

    def field_noalloc(self, states, budgets):
        """field() through pe_rtable_field_noalloc (no device allocation;
        at most PE_RTABLE_NOALLOC_MAX queries per call)."""
        s = np.ascontiguousarray(states, np.float64).reshape(-1, 5)
        b = np.ascontiguousarray(budgets, np.float64).reshape(-1)
        if len(b) != len(s):
            raise ValueError("one budget per state required")
        out = np.empty(len(s), np.float64)
        check(lib().pe_rtable_field_noalloc(self._h, ptr(s), ptr(b), len(s), ptr(out)))
        return out

    def value(self, states, capture_radius, speed_limits):
        s = np.ascontiguousarray(states, np.float64).reshape(-1, 5)
        lim = np.ascontiguousarray(speed_limits, np.float64)
        out = np.empty(len(s), np.float64)
        check(lib().pe_rtable_value(self._h, ptr(s), len(s), float(capture_radius), ptr(lim),
                                    ptr(out)))
        return out

    def close(self):
        if self._h:
            lib().pe_rtable_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
