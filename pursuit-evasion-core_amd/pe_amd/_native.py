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
        ("reserved", ctypes.c_int32),
        ("p_length", ctypes.c_double),
        ("p_width", ctypes.c_double),
        ("e_length", ctypes.c_double),
        ("e_width", ctypes.c_double),
    ]


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


P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
F64 = ctypes.c_double

# name -> (restype, argtypes); the exported symbols of include/pe_amd.h.
SIGNATURES = {
    "pe_last_error": (ctypes.c_char_p, []),
    "pe_abi_version": (ctypes.c_int, []),
    "pe_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, P, P]),
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
