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
        ("exit_radius", ctypes.c_double),
    ]


class World8DCfg(ctypes.Structure):
    """pe_oracle_world8d_cfg: native_world.cpp's argv options."""
    _fields_ = [(k, ctypes.c_int) for k in (
        "positions", "headings", "p_speeds", "e_speeds", "order", "linear_speed_boundary",
        "threads")] + [(k, ctypes.c_float) for k in (
        "extent", "arena", "radius", "p_min", "p_max", "e_min", "e_max", "p_accel", "e_accel",
        "p_curvature", "e_curvature")] + [("cfl", ctypes.c_double)]


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
    W = ctypes.POINTER(World8DCfg)
    L.pe_oracle_world8d_cfl_limit.argtypes = [W]
    L.pe_oracle_world8d_cfl_limit.restype = ctypes.c_double
    L.pe_oracle_world8d_states.argtypes = [W]
    L.pe_oracle_world8d_states.restype = ctypes.c_int64
    L.pe_oracle_world8d_init.argtypes = [W, P, P]
    L.pe_oracle_world8d_init.restype = None
    L.pe_oracle_world8d_solve.argtypes = [W, P, P, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, P, P]
    L.pe_oracle_world8d_solve.restype = ctypes.c_int64
    _declare_8d(L)


def _declare_8d(L):
    """Declarations of the 8D contact-game oracle (sl8d.c, geometry.c)."""
    I64, I32, D = ctypes.c_int64, ctypes.c_int, ctypes.c_double
    L.pe_oracle_batch_clearance.argtypes = [P, I64, P, I32, D, P, P, I32]
    L.pe_oracle_batch_backups.argtypes = [P, I64, P, I32, D, P, P, I32]
    L.pe_oracle_batch_actions.argtypes = [P, I64, P, I32, D, P, I32, D, P, P, P, I32]
    L.pe_oracle_batch_connections.argtypes = [P, P, I64, P, I32, D, P, P, I32]
    for name in ("clearance", "backups", "actions", "connections"):
        getattr(L, "pe_oracle_batch_" + name).restype = None
    L.pe_oracle_contact.argtypes = [P, P, D, D, D, D]
    L.pe_oracle_contact.restype = D
    L.pe_oracle_step.argtypes = [P, D, D, D, P, P]
    L.pe_oracle_step.restype = None
    L.pe_oracle_sl_init.argtypes = [I64, I64, P, P, P, D, D, D, D, D, P, I32]
    L.pe_oracle_sl_init.restype = None
    L.pe_oracle_car_stencil.restype = ctypes.c_int64
    L.pe_oracle_car_stencil.argtypes = [P, P, P, P, P, P, P, I64, P, I32, D, P, P, P, P, P, P]
    L.pe_oracle_grid_stencil.restype = None
    L.pe_oracle_grid_stencil.argtypes = [I32, P, P, P, P, P, ctypes.c_uint32, D, P, I64, P, P, P]
    L.pe_oracle_nearest.restype = None
    L.pe_oracle_nearest.argtypes = [P, I32, P, I64, I32, D, D, D, P, P]
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


def hji5d_cfg(config, grid, *, order=2, cfl=0.8, threads=None, store_f64=False,
              exit_radius=0.0):
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
        exit_radius,
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


# ----------------------------------------------------- dense 8D world game

def world8d_cfl_limit(cfg: World8DCfg) -> float:
    return lib().pe_oracle_world8d_cfl_limit(ctypes.byref(cfg))


def world8d_init(cfg: World8DCfg):
    n = lib().pe_oracle_world8d_states(ctypes.byref(cfg))
    field = np.empty(n, np.float32)
    arrival = np.empty(n, np.float32)
    lib().pe_oracle_world8d_init(ctypes.byref(cfg), _ptr(field), _ptr(arrival))
    return field, arrival


def world8d_solve(cfg: World8DCfg, horizon, output_dt, start=0.0, field=None, arrival=None,
                  history=True):
    """(frames [k+1, N] including the initial field when history, final,
    arrival, steps, cfl dt limit)."""
    if field is None:
        field, arrival = world8d_init(cfg)
    field = np.ascontiguousarray(field, np.float32).copy()
    arrival = np.ascontiguousarray(arrival, np.float32).copy()
    count = int(np.ceil((horizon - start) / output_dt - 1e-10))
    frames = np.empty((count, field.size), np.float32) if history else None
    first = field.copy()
    stats = np.zeros(1, np.float64)
    steps = lib().pe_oracle_world8d_solve(ctypes.byref(cfg), _ptr(field), _ptr(arrival), start,
                                          horizon, output_dt, _ptr(frames), _ptr(stats))
    if history:
        frames = np.concatenate([first[None], frames]) if start == 0 else frames
    return frames, field, arrival, int(steps), float(stats[0])


# ------------------------------------------------------------------ 8D game

def model_params(limits, body, dt):
    """[low, high, a, k, L, W, dt] (local_game/native.py:141-153)."""
    return np.array([limits.speed_min, limits.speed_max, limits.acceleration,
                     limits.max_curvature, body.length, body.width, dt], np.float64)


def _states(s):
    s = np.ascontiguousarray(np.asarray(s, np.float64).reshape(-1, 4))
    return s


def _obs(obstacles):
    return np.ascontiguousarray(np.asarray(obstacles, np.float64).reshape(-1, 3))


def geom_clearance(states, obstacles, half, params, threads=None):
    s, o = _states(states), _obs(obstacles)
    out = np.empty(len(s))
    lib().pe_oracle_batch_clearance(_ptr(s), len(s), _ptr(o), len(o), half,
                                    _ptr(params), _ptr(out), threads or os.cpu_count())
    return out


def geom_backups(states, obstacles, half, params, threads=None):
    s, o = _states(states), _obs(obstacles)
    out = np.empty(len(s), np.uint8)
    lib().pe_oracle_batch_backups(_ptr(s), len(s), _ptr(o), len(o), half,
                                  _ptr(params), _ptr(out), threads or os.cpu_count())
    return out.astype(bool)


def geom_actions(states, controls, dt, obstacles, half, params, threads=None):
    s, o = _states(states), _obs(obstacles)
    u = np.ascontiguousarray(np.asarray(controls, np.float64).reshape(-1, 2))
    ok = np.empty((len(s), len(u)), np.uint8)
    nxt = np.empty((len(s), len(u), 4))
    lib().pe_oracle_batch_actions(_ptr(s), len(s), _ptr(u), len(u), dt, _ptr(o), len(o),
                                  half, _ptr(params), _ptr(ok), _ptr(nxt),
                                  threads or os.cpu_count())
    return ok.astype(bool), nxt


def geom_connections(start, target, obstacles, half, params, threads=None):
    a, b, o = _states(start), _states(target), _obs(obstacles)
    out = np.empty(len(a), np.uint8)
    lib().pe_oracle_batch_connections(_ptr(a), _ptr(b), len(a), _ptr(o), len(o), half,
                                      _ptr(params), _ptr(out), threads or os.cpu_count())
    return out.astype(bool)


def sl_init(pstates, estates, inside, exit_value, pbody, ebody, threads=None):
    ps, es = _states(pstates), _states(estates)
    ins = np.ascontiguousarray(inside, np.uint8)
    V = np.empty((len(ps), len(es)))
    lib().pe_oracle_sl_init(len(ps), len(es), _ptr(ps), _ptr(es), _ptr(ins), exit_value,
                            pbody.length, pbody.width, ebody.length, ebody.width, _ptr(V),
                            threads or os.cpu_count())
    return V


def sl_sweep(V, p, e, inside, threads=None):
    """One FP64 Jacobi sweep.  p / e: objects with successors [n,A,16] int32,
    weights [n,A,16] float32 and allowed [n,A] bool.  Returns (out, change)."""
    V = np.ascontiguousarray(V, np.float64)
    out = np.empty_like(V)
    arrs = []
    for g in (p, e):
        arrs += [np.ascontiguousarray(g.successors, np.int32),
                 np.ascontiguousarray(g.weights, np.float32),
                 np.ascontiguousarray(g.allowed, np.uint8)]
    ins = np.ascontiguousarray(inside, np.uint8)
    change = lib().pe_oracle_sl_sweep(
        V.shape[0], V.shape[1], p.successors.shape[1], e.successors.shape[1],
        *[_ptr(a) for a in arrs], _ptr(ins), _ptr(V), _ptr(out), threads or os.cpu_count())
    return out, change


# ------------------------------------------------------------ car tables

def car_stencil(axes, dense_to_compact, nxt, obstacles, half, params):
    """Successor stencils of one car grid (stencil.c): ids [m,16] int32,
    weights [m,16] float32, valid [m] bool, mass [m], raw vertex weights
    [m,16] before renormalisation, and the number of fallback rows."""
    n = np.array([len(a) for a in axes], np.int32)
    ax = [np.ascontiguousarray(a, np.float64) for a in axes]
    d2c = np.ascontiguousarray(dense_to_compact, np.int32).reshape(-1)
    q = _states(nxt)
    o = _obs(obstacles)
    m = len(q)
    ids = np.empty((m, 16), np.int32)
    w = np.empty((m, 16), np.float32)
    valid = np.empty(m, np.uint8)
    mass = np.empty(m)
    raw = np.empty((m, 16))
    fb = lib().pe_oracle_car_stencil(_ptr(n), *[_ptr(a) for a in ax], _ptr(d2c), _ptr(q), m,
                                     _ptr(o), len(o), half, _ptr(params), _ptr(ids), _ptr(w),
                                     _ptr(valid), _ptr(mass), _ptr(raw))
    return ids, w, valid.astype(bool), mass, raw, int(fb)


def nearest(features, queries, k, dx, dh, dv):
    """k nearest feature rows of each query state (ties by index):
    ids [m,k], distances [m,k]."""
    f = np.ascontiguousarray(features, np.float64).reshape(-1, 5)
    q = _states(queries)
    ids = np.empty((len(q), k), np.int32)
    dist = np.empty((len(q), k))
    lib().pe_oracle_nearest(_ptr(f), len(f), _ptr(q), len(q), k, dx, dh, dv, _ptr(ids),
                            _ptr(dist))
    return ids, dist


def grid_stencil(states, n, lo, hi, step, period=None, periodic=0, slack=1e-10):
    """Multilinear stencil on a uniform grid (stencil.c): index [m, 2^d]
    int64, weights [m, 2^d] float64, valid [m] bool."""
    d = len(n)
    q = np.ascontiguousarray(np.asarray(states, np.float64).reshape(-1, d))
    arr = [np.ascontiguousarray(a, t) for a, t in ((n, np.int32), (lo, np.float64),
                                                   (hi, np.float64), (step, np.float64))]
    per = np.ascontiguousarray(period if period is not None else np.zeros(d), np.float64)
    idx = np.empty((len(q), 1 << d), np.int64)
    w = np.empty((len(q), 1 << d), np.float64)
    valid = np.empty(len(q), np.uint8)
    lib().pe_oracle_grid_stencil(d, *[_ptr(a) for a in arr], _ptr(per), int(periodic), slack,
                                 _ptr(q), len(q), _ptr(idx), _ptr(w), _ptr(valid))
    return idx, w, valid.astype(bool)
