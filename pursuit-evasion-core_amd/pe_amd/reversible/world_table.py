"""Queries of a saved dense 8D square-game table (reversible/world_table.py).

Reads the files `world_dp.solve` writes (memory-mapped float32 fields) and
answers the reference reader's queries: multilinear field values at a state
and budget, first-crossing values, containment.  States outside the
numerical domain stay explicit (NaN).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from .model import ReversibleConfig
from .world_dp import PERIODIC_AXES, TUBE_PROJECTION, WorldGrid


def interpolation_stencil(states, axes, periodic_axes):
    """Node indices [m, 2^d], weights [m, 2^d], validity [m] and batch shape of
    the multilinear stencils of `states` [..., d] (world_table.py:13-45).

    Periodic axes wrap; other axes clamp the coordinate into the grid and
    mark states more than 1e-10 outside it invalid; the last cell of an axis
    is [n-2, n-1]."""
    s = np.asarray(states, dtype=float)
    d = len(axes)
    if s.shape[-1] != d:
        raise ValueError(f"Expected final state dimension {d}.")
    batch = s.shape[:-1]
    flat = s.reshape(-1, d)
    finite = np.isfinite(flat)
    valid = finite.all(1)
    q_all = np.where(finite, flat, 0.0)
    sizes = [len(a) for a in axes]
    strides = [int(np.prod(sizes[k + 1:], dtype=np.int64)) for k in range(d)]
    # Corner c: bit (d - 1 - k) of c selects the upper node on axis k (the
    # order of itertools.product([0, 1], repeat=d)).
    upper = ((np.arange(2 ** d)[:, None] >> (d - 1 - np.arange(d))[None, :]) & 1).astype(bool)
    index = np.zeros((len(flat), 2 ** d), dtype=np.int64)
    weight = np.ones((len(flat), 2 ** d), dtype=float)
    for k, axis in enumerate(axes):
        n, a0 = sizes[k], float(axis[0])
        h = float(axis[1] - axis[0])
        q = q_all[:, k]
        if k in periodic_axes:
            f = ((q - a0) % (h * n)) / h
            cell = np.floor(f)
            low = cell.astype(np.int64) % n
            frac = f - cell
            high = (low + 1) % n
        else:
            valid &= (q >= a0 - 1e-10) & (q <= float(axis[-1]) + 1e-10)
            f = (np.clip(q, a0, float(axis[-1])) - a0) / h
            low = np.minimum(np.floor(f).astype(np.int64), n - 2)
            frac = np.clip(f - low, 0, 1)
            high = low + 1
        up = upper[None, :, k]
        index += np.where(up, high[:, None], low[:, None]) * strides[k]
        weight *= np.where(up, frac[:, None], 1 - frac[:, None])
    return index, weight, valid, batch


class WorldTable:
    """Reader of a completed dense 8D solve, continuations included
    (world_table.py:48-220)."""

    state_dimension = 8

    def __init__(self, folder, *, allow_legacy_speed_boundary=False):
        self.folder = Path(folder)
        self.meta = json.loads((self.folder / "meta.json").read_text())
        if self.meta.get("status") != "complete" or self.meta.get("backend") not in (
                "native_world8d", "optimized_dp_world8d"):
            raise ValueError("A completed reversible world8D solve is required.")
        self.speed_boundary = self.meta.get("speed_boundary", "legacy_signed")
        if self.speed_boundary != "linear" and not allow_legacy_speed_boundary:
            raise ValueError(
                "This preliminary table used signed-distance extrapolation on speed axes. "
                "Use a corrected linear-speed-boundary table for control; explicit "
                "allow_legacy_speed_boundary=True is available for diagnostics only.")
        if self.meta.get("tube_projection") != TUBE_PROJECTION:
            raise ValueError("Unsupported numerical field convention.")
        self.config = ReversibleConfig.from_dict(self.meta["config"])
        g = self.meta["grid"]
        self.grid = WorldGrid(**{k: g[k] for k in ("positions", "headings", "p_speeds",
                                                    "e_speeds", "extent")})
        self.axes = self.grid.axes(self.config)
        self.local_times = np.asarray(self.meta["times"], dtype=float)
        self.times = self.local_times.copy()
        self.horizon = float(self.meta["horizon"])
        self.has_history = bool(self.meta["history_saved"])
        self.fields = None
        # Logical frame t lives in segment _owners[t], row _rows[t].
        self._segments, self._owners, self._rows = [], np.zeros(0, int), np.zeros(0, int)
        if self.has_history:
            self.fields = self._map("fields.f32", (len(self.local_times), self.grid.states))
            self._segments = [self.fields]
            self._owners = np.zeros(len(self.local_times), dtype=int)
            self._rows = np.arange(len(self.local_times))
        if "parent_table" in self.meta and self.has_history:
            self.parent = WorldTable((self.folder / self.meta["parent_table"]).resolve(),
                                     allow_legacy_speed_boundary=allow_legacy_speed_boundary)
            if (not self.parent.has_history or self.parent.config != self.config
                    or self.parent.grid != self.grid):
                raise ValueError("The retained parent history/grid/model is required.")
            if self.local_times[0] <= self.parent.times[-1] or not np.isclose(
                    self.meta["start_time"], self.parent.horizon):
                raise ValueError("Continuation history must start after the parent boundary.")
            owner = len(self.parent._segments)
            self._segments = self.parent._segments + [self.fields]
            self._owners = np.concatenate(
                (self.parent._owners, np.full(len(self.local_times), owner, int)))
            self._rows = np.concatenate((self.parent._rows, np.arange(len(self.local_times))))
            self.times = np.concatenate((self.parent.times, self.local_times))
        if self.meta.get("final_source", "final.f32") == "history_last":
            if self.fields is None:
                raise ValueError("The final field is stored in the retained history.")
            self.final = self.fields[-1]
        else:
            self.final = self._map("final.f32", (self.grid.states,))

    def _map(self, name, shape):
        path = self.folder / name
        if path.stat().st_size != int(np.prod(shape, dtype=np.int64)) * 4:
            raise ValueError(f"Field size mismatch: {name}")
        return np.memmap(path, dtype=np.float32, mode="r", shape=shape)

    def _sample(self, frames, index, weight):
        """Interpolated values of each query at its logical frame."""
        t = np.broadcast_to(np.asarray(frames, dtype=int), (len(index),))
        owner, row = self._owners[t], self._rows[t]
        out = np.empty(len(index), dtype=float)
        for o in np.unique(owner):
            m = owner == o
            out[m] = (self._segments[o][row[m, None], index[m]] * weight[m]).sum(1)
        return out

    def field(self, states, budget=None):
        """Signed field at `states` for the final field or a budget in
        [0, horizon] (linear in time between saved frames)."""
        ii, ww, valid, batch = interpolation_stencil(states, self.axes, PERIODIC_AXES)
        if budget is None:
            return np.where(valid, (self.final[ii] * ww).sum(1), np.nan).reshape(batch)
        tau = np.broadcast_to(np.asarray(budget, dtype=float), batch).reshape(-1)
        valid &= np.isfinite(tau) & (tau >= 0) & (tau <= self.horizon + 1e-10)
        if not self.has_history:
            if np.any(valid & ~np.isclose(tau, self.horizon, atol=1e-10, rtol=0)):
                raise ValueError(
                    "Only the final field is retained; arbitrary budgets require history.")
            return np.where(valid, (self.final[ii] * ww).sum(1), np.nan).reshape(batch)
        k = np.clip(np.searchsorted(self.times, tau, side="right") - 1, 0,
                    len(self.times) - 2)
        w = (tau - self.times[k]) / (self.times[k + 1] - self.times[k])
        lo, hi = self._sample(k, ii, ww), self._sample(k + 1, ii, ww)
        return np.where(valid, lo + (hi - lo) * w, np.nan).reshape(batch)

    def value(self, states):
        """First crossing of zero of the interpolated signed history (the
        numerical time to P success, E faults included); 0 on the target, NaN
        where unresolved or P is outside the square."""
        if not self.has_history:
            raise ValueError("Time-to-success queries require retained signed histories.")
        ii, ww, valid, batch = interpolation_stencil(states, self.axes, PERIODIC_AXES)
        lo = np.zeros(len(ii), int)
        hi = np.full(len(ii), len(self.times) - 1, int)
        valid &= self._sample(hi, ii, ww) <= 0
        # Bisection for the first frame at or below zero.
        while np.any(hi - lo > 1):
            mid = (lo + hi) // 2
            inside = self._sample(mid, ii, ww) <= 0
            hi = np.where(inside, mid, hi)
            lo = np.where(inside, lo, mid)
        a, b = self._sample(lo, ii, ww), self._sample(hi, ii, ww)
        f = np.divide(a, a - b, out=np.zeros_like(a), where=abs(a - b) > 1e-14)
        values = self.times[lo] + np.clip(f, 0, 1) * (self.times[hi] - self.times[lo])
        s = np.asarray(states, dtype=float).reshape(-1, 8)
        c = self.config
        finite = np.isfinite(s).all(1)
        speeds_ok = ((s[:, 3] >= c.pursuer.speed_min) & (s[:, 3] <= c.pursuer.speed_max)
                     & (s[:, 7] >= c.evader.speed_min) & (s[:, 7] <= c.evader.speed_max))
        p_inside = np.max(abs(s[:, :2]), axis=1) < c.arena_half_width
        captured = np.linalg.norm(s[:, :2] - s[:, 4:6], axis=1) <= c.capture_radius
        e_fault = np.max(abs(s[:, 4:6]), axis=1) >= c.arena_half_width
        values = np.where(valid & p_inside, values, np.nan)
        values = np.where(finite & speeds_ok & p_inside & (captured | e_fault), 0.0, values)
        return values.reshape(batch)

    def contains(self, states, budget=None):
        f = self.field(states, budget)
        return np.isfinite(f) & (f <= 0)
