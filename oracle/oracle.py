"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the CPU restatement of the reference
algorithms for this path.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker: the product
(pursuit-evasion-core_amd/) never loads it.
"""

from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path
import subprocess

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_DIR = HERE / "_ref"
_LIB = None


def build(force: bool = False) -> Path:
    """Compile the oracle with make (gcc + OpenMP)."""
    cmd = ["make", "-C", str(HERE)]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def build_reference() -> bool:
    """Compile the reference's own CPU sources into oracle/_ref when present."""
    if not Path("/root/reference/src").is_dir():
        return False
    subprocess.run(
        ["make", "-f", str(HERE / "ref.mk")], check=True, stdout=subprocess.DEVNULL
    )
    return True


class HJI5DCfg(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("nh", ctypes.c_int),
        ("nvp", ctypes.c_int),
        ("nve", ctypes.c_int),
        ("order", ctypes.c_int),
        ("threads", ctypes.c_int),
        ("store_f64", ctypes.c_int),
        ("extent", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("radius", ctypes.c_double),
        ("vpmin", ctypes.c_double),
        ("vpmax", ctypes.c_double),
        ("vemin", ctypes.c_double),
        ("vemax", ctypes.c_double),
        ("ap", ctypes.c_double),
        ("ae", ctypes.c_double),
        ("kp", ctypes.c_double),
        ("ke", ctypes.c_double),
    ]


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        _LIB = ctypes.CDLL(str(LIB_PATH))
        _declare(_LIB)
    return _LIB


P = ctypes.c_void_p


def _declare(L):
    L.pe_oracle_hji5d_solve.argtypes = [
        ctypes.POINTER(HJI5DCfg),
        ctypes.c_double,
        ctypes.c_double,
        P,
        P,
        P,
        P,
        P,
        ctypes.c_int64,
    ]
    L.pe_oracle_hji5d_solve.restype = ctypes.c_int
    L.pe_oracle_hji5d_cfl_limit.argtypes = [ctypes.POINTER(HJI5DCfg)]
    L.pe_oracle_hji5d_cfl_limit.restype = ctypes.c_double
    L.pe_oracle_hji5d_steps.argtypes = [
        ctypes.POINTER(HJI5DCfg),
        P,
        P,
        ctypes.c_double,
        ctypes.c_double,
        ctypes.c_int,
    ]
    L.pe_oracle_hji5d_steps.restype = ctypes.c_double
    L.pe_oracle_hji5d_init.argtypes = [ctypes.POINTER(HJI5DCfg), P, P]
    L.pe_oracle_hji5d_init.restype = None
    _declare_8d(L)


def _declare_8d(L):
    """Declarations of the 8D contact-game oracle (filled in by sl8d.c)."""
    if not hasattr(L, "pe_oracle_sl_sweep"):
        return
    L.pe_oracle_sl_sweep.restype = ctypes.c_double
    L.pe_oracle_sl_sweep.argtypes = [
        ctypes.c_int64,  # np
        ctypes.c_int64,  # ne
        ctypes.c_int,  # npa
        ctypes.c_int,  # nea
        P,  # pids
        P,  # pw
        P,  # pallowed
        P,  # eids
        P,  # ew
        P,  # eallowed
        P,  # inside (u8 [np, ne])
        P,  # V in (f64)
        P,  # V out (f64)
        ctypes.c_int,  # threads
    ]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def hji5d_cfg(config, grid, *, order=2, cfl=0.8, threads=None, store_f64=False):
    """Options from a ReversibleConfig-like object and a RelativeGrid."""
    p, e = config.pursuer, config.evader
    return HJI5DCfg(
        grid.positions,
        grid.headings,
        grid.p_speeds,
        grid.e_speeds,
        order,
        threads or os.cpu_count() or 1,
        int(store_f64),
        grid.extent,
        cfl,
        config.capture_radius,
        p.speed_min,
        p.speed_max,
        e.speed_min,
        e.speed_max,
        p.acceleration,
        e.acceleration,
        p.max_curvature,
        e.max_curvature,
    )


def hji5d_solve(cfg: HJI5DCfg, horizon, output_dt, *, history=True, trace_cap=0):
    """Whole 5D solve; returns a dict like the reference's output files."""
    n = cfg.nx * cfg.nx * cfg.nh * cfg.nve * cfg.nvp
    dtype = np.float64 if cfg.store_f64 else np.float32
    outputs = math.ceil(horizon / output_dt - 1e-10)
    frames = np.empty((outputs + 1, n), dtype) if history else None
    final = np.empty(n, dtype)
    arrival = np.empty(n, np.float32)
    stats = np.zeros(6)
    trace = np.zeros(max(trace_cap, 1))
    rc = lib().pe_oracle_hji5d_solve(
        ctypes.byref(cfg),
        horizon,
        output_dt,
        _ptr(frames),
        _ptr(final),
        _ptr(arrival),
        _ptr(stats),
        _ptr(trace) if trace_cap else None,
        trace_cap,
    )
    if rc == -1:
        raise ValueError("Invalid solver options")
    if rc == -2:
        raise RuntimeError("Nonfinite field update")
    steps = int(stats[1])
    return dict(
        fields=frames,
        final=final,
        arrival=arrival,
        cfl_dt_limit=stats[0],
        internal_steps=steps,
        frames=int(stats[2]),
        maximum_increase=stats[3],
        time=stats[4],
        last_change=stats[5],
        trace=trace[: min(steps, trace_cap)] if trace_cap else None,
    )


def hji5d_cfl_limit(cfg: HJI5DCfg) -> float:
    return lib().pe_oracle_hji5d_cfl_limit(ctypes.byref(cfg))


def hji5d_init(cfg: HJI5DCfg):
    n = cfg.nx * cfg.nx * cfg.nh * cfg.nve * cfg.nvp
    field = np.empty(n, np.float32)
    arrival = np.empty(n, np.float32)
    lib().pe_oracle_hji5d_init(ctypes.byref(cfg), _ptr(field), _ptr(arrival))
    return field, arrival


def hji5d_time_steps(cfg: HJI5DCfg, field, arrival, dt, count, t=0.0) -> float:
    """Run `count` fixed-dt RK steps in place; returns wall seconds."""
    assert field.dtype == np.float32 and field.flags.c_contiguous
    return lib().pe_oracle_hji5d_steps(
        ctypes.byref(cfg), _ptr(field), _ptr(arrival), dt, t, count
    )
