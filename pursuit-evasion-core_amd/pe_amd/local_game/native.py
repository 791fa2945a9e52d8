"""Per-vehicle geometry engine on the GPU.

Same interface as the reference's GeometryEngine (local_game/native.py:135-253):
clearance, backup_safe, connections, matched-free batched actions, with a
model [speed_min, speed_max, acceleration, max_curvature, length, width,
control_dt] per car.  The computation runs in libpe_amd.so (pe_geom_*,
FP64 HIP kernels) instead of an OpenMP library; `threads` is accepted for
interface compatibility only.
"""

from __future__ import annotations

import numpy as np

from .. import _native


class GeometryEngine:
    def __init__(self, arena, limits, role, threads=6, control_dt=0.05):
        self.arena = arena
        self.limits = limits
        self.role = role
        self.threads = threads
        body = getattr(arena, role)
        self.params = np.array(
            [limits.speed_min, limits.speed_max, limits.acceleration,
             limits.max_curvature, body.length, body.width, control_dt],
            np.float64,
        )
        self.obstacles = np.ascontiguousarray(arena.circles, np.float64).reshape(-1, 3)

    @staticmethod
    def _states(s):
        s = np.asarray(s, float)
        if s.shape[-1] != 4 or not np.isfinite(s).all():
            raise ValueError("Finite vehicle states[...,4] required")
        return np.ascontiguousarray(s.reshape(-1, 4))

    def clearance(self, s):
        x = self._states(s)
        out = np.empty(len(x))
        _native.check(_native.lib().pe_geom_clearance(
            _native.ptr(x), len(x), _native.ptr(self.obstacles), len(self.obstacles),
            float(self.arena.half_width), _native.ptr(self.params), _native.ptr(out)))
        return out.reshape(np.shape(s)[:-1])

    def backup_safe(self, s):
        x = self._states(s)
        out = np.empty(len(x), np.uint8)
        _native.check(_native.lib().pe_geom_backups(
            _native.ptr(x), len(x), _native.ptr(self.obstacles), len(self.obstacles),
            float(self.arena.half_width), _native.ptr(self.params), _native.ptr(out)))
        return out.astype(bool).reshape(np.shape(s)[:-1])

    def connections(self, start, target):
        a, b = self._states(start), self._states(target)
        if a.shape != b.shape:
            raise ValueError("Matched configuration arrays required")
        out = np.empty(len(a), np.uint8)
        _native.check(_native.lib().pe_geom_connections(
            _native.ptr(a), _native.ptr(b), len(a), _native.ptr(self.obstacles),
            len(self.obstacles), float(self.arena.half_width), _native.ptr(self.params),
            _native.ptr(out)))
        return out.astype(bool)

    def actions(self, s, controls, dt):
        x = self._states(s)
        u = np.ascontiguousarray(controls, float).reshape(-1, 2)
        ok = np.empty((len(x), len(u)), np.uint8)
        nxt = np.empty((len(x), len(u), 4))
        _native.check(_native.lib().pe_geom_actions(
            _native.ptr(x), len(x), _native.ptr(u), len(u), float(dt),
            _native.ptr(self.obstacles), len(self.obstacles), float(self.arena.half_width),
            _native.ptr(self.params), _native.ptr(ok), _native.ptr(nxt)))
        return ok.astype(bool), nxt


def _many(engines, state_lists):
    """Concatenated states, their map indices, the circles of every map with
    offsets, and the half-widths, for the multi-map queries (one car model
    for all maps)."""
    params = engines[0].params
    if any(not np.array_equal(e.params, params) for e in engines):
        raise ValueError("One vehicle model is required for a multi-map query")
    xs = [GeometryEngine._states(s) for s in state_lists]
    x = np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros((0, 4)))
    map_of = np.concatenate([np.full(len(v), b, np.int32) for b, v in enumerate(xs)])
    obs = np.ascontiguousarray(np.concatenate([e.obstacles for e in engines]), np.float64)
    off = np.concatenate([[0], np.cumsum([len(e.obstacles) for e in engines])]).astype(np.int64)
    halves = np.array([e.arena.half_width for e in engines], np.float64)
    return x, np.ascontiguousarray(map_of), obs, off, halves, params, [len(v) for v in xs]


def backup_safe_many(engines, state_lists):
    """GeometryEngine.backup_safe of several maps in one launch
    (pe_geom_backups_multi); returns one mask per map."""
    x, map_of, obs, off, halves, params, sizes = _many(engines, state_lists)
    out = np.empty(len(x), np.uint8)
    _native.check(_native.lib().pe_geom_backups_multi(
        _native.ptr(x), len(x), _native.ptr(map_of), len(engines), _native.ptr(obs),
        _native.ptr(off), _native.ptr(halves), _native.ptr(params), _native.ptr(out)))
    return np.split(out.astype(bool), np.cumsum(sizes)[:-1])


def actions_many(engines, state_lists, controls, dt):
    """GeometryEngine.actions of several maps in one launch
    (pe_geom_actions_multi); returns (ok, next) per map."""
    x, map_of, obs, off, halves, params, sizes = _many(engines, state_lists)
    u = np.ascontiguousarray(controls, float).reshape(-1, 2)
    ok = np.empty((len(x), len(u)), np.uint8)
    nxt = np.empty((len(x), len(u), 4))
    _native.check(_native.lib().pe_geom_actions_multi(
        _native.ptr(x), len(x), _native.ptr(map_of), len(engines), _native.ptr(u), len(u),
        float(dt), _native.ptr(obs), _native.ptr(off), _native.ptr(halves),
        _native.ptr(params), _native.ptr(ok), _native.ptr(nxt)))
    cut = np.cumsum(sizes)[:-1]
    return list(zip(np.split(ok.astype(bool), cut), np.split(nxt, cut)))
