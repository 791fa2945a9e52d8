"""Reader of a saved 5D relative-game history, for controllers.

Drop-in for the reference's reversible/relative_table.py: the class name,
constructor arguments (`folder`, `config`, `max_budget`), attributes
(`times`, `axes`, `shape`, `spacings`, `fields`, `physics`, `horizon`),
query methods and error messages are its interface.  The work is done on
the GPU: the history is copied to the device on the first query and every
query batch is one launch (pe_rtable_field / pe_rtable_value); `stencil` is
pe_grid_stencil; `gradient` evaluates all ten shifted states of a batch in a
single field query.

Storage order of the grid is [vP, vE, psi, y, x]; queries use the public
order [x, y, psi, vP, vE].
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from .. import _native
from .free_dp import physics_signature
from .grid_stencil import axis_spec

# storage axis of each public coordinate x, y, psi, vP, vE
_STORAGE_OF = (4, 3, 2, 0, 1)
# public column of each storage axis vP, vE, psi, y, x
_PUBLIC_OF = (3, 4, 2, 1, 0)
HEADING = 2


def _checked_meta(folder):
    meta = json.loads((folder / "meta.json").read_text())
    rules = (
        (meta.get("schema") == "reversible_relative5d_v1", "Not a reversible relative5D table"),
        (meta.get("status") == "complete" and meta.get("history_saved"),
         "A completed solve with signed history is required"),
        (meta.get("tube_projection") == "min over every internal update",
         "Unsupported budget convention"),
    )
    for ok, message in rules:
        if not ok:
            raise ValueError(message)
    return meta


def _checked_axes(meta):
    axes = tuple(np.asarray(a, dtype=float) for a in meta["axes"])
    shape = tuple(len(a) for a in axes)
    if len(axes) != 5 or min(shape) < 2:
        raise ValueError("Expected five nondegenerate coordinate axes")
    steps = np.array([a[1] - a[0] for a in axes])
    uniform = all(np.isfinite(a).all() and h > 0 and np.allclose(np.diff(a), h)
                  for a, h in zip(axes, steps))
    if not uniform:
        raise ValueError("Axes must be finite, increasing and uniformly spaced")
    heading = axes[HEADING]
    if not (np.isclose(heading[0], -np.pi) and np.isclose(steps[HEADING] * len(heading),
                                                          2 * np.pi)):
        raise ValueError("Heading axis must be periodic without a repeated endpoint")
    if "shape" in meta and tuple(meta["shape"]) != shape:
        raise ValueError("Metadata shape differs from coordinate axes")
    return axes, shape, steps


class RelativeTable:
    state_dimension = 5

    def __init__(self, folder, config=None, *, max_budget=None):
        self.folder = Path(folder).resolve()
        self.meta = _checked_meta(self.folder)
        self.axes, self.shape, self.spacings = _checked_axes(self.meta)
        saved = np.asarray(self.meta["times"], dtype=float)
        if len(saved) < 2 or (np.diff(saved) <= 0).any():
            raise ValueError("Invalid output times")
        frames = len(saved)
        if max_budget is not None:
            at = np.flatnonzero(np.abs(saved - max_budget) <= 1e-10)
            if len(at) == 0 or at[0] == 0:
                raise ValueError("max_budget must be a positive retained output budget")
            frames = int(at[0]) + 1
        self.times = saved[:frames].copy()
        history = self.folder / "fields.f32"
        if history.stat().st_size != 4 * len(saved) * int(np.prod(self.shape)):
            raise ValueError("History byte length differs from metadata")
        self.fields = np.memmap(history, mode="r", dtype=np.float32,
                                shape=(len(saved),) + self.shape)[:frames]
        self.physics = self.meta["physics"]
        self.horizon = float(self.times[-1])
        self._device = None
        if config is not None:
            self.assert_compatible(config)

    def assert_compatible(self, config):
        if physics_signature(config) != self.physics:
            raise ValueError("The free-space table does not match the vehicle/capture model")

    def _gpu(self):
        """The retained history on the device (uploaded once)."""
        if self._device is None:
            frames = np.ascontiguousarray(self.fields).reshape(len(self.times), -1)
            self._device = _native.RelTable(frames, self.axes, self.times)
        return self._device

    @staticmethod
    def _rows(states):
        s = np.asarray(states, dtype=float)
        if s.ndim < 1 or s.shape[-1] != 5:
            raise ValueError("Relative states must end in [x,y,psi,vP,vE]")
        return s.reshape(-1, 5), s.shape[:-1]

    def stencil(self, states):
        """32 node indices and weights per state (storage order, vertex c
        taking the upper node of storage axis k when bit 4 - k is set),
        validity and the batch shape."""
        rows, batch = self._rows(states)
        n, lo, hi, step, _, mask = axis_spec(self.axes, (HEADING,))
        period = [0.0, 0.0, 2 * np.pi, 0.0, 0.0]
        index, weight, valid = _native.grid_stencil(rows[:, list(_PUBLIC_OF)], n, lo, hi, step,
                                                    period, mask, 1e-10)
        return index, weight, valid, batch

    def field(self, states, budget):
        """Signed field at states and budgets: multilinear in space, linear in
        the budget between saved frames, NaN outside either."""
        rows, batch = self._rows(states)
        tau = np.broadcast_to(np.asarray(budget, dtype=float), batch).reshape(-1)
        if len(rows) == 0:
            return np.full(batch, np.nan)
        return self._gpu().field(rows, tau).reshape(batch)

    def value(self, states):
        """First budget at which the interpolated history reaches zero (0 for
        captured physical states, NaN where it never does)."""
        if self.meta.get("capture_rule") == "rectangle_contact":
            raise ValueError(
                "Use local_game.free_value.ContactFreeValue for rectangular time queries")
        rows, batch = self._rows(states)
        if len(rows) == 0:
            return np.full(batch, np.nan)
        p, e = self.physics["pursuer"], self.physics["evader"]
        limits = (p["speed_min"], p["speed_max"], e["speed_min"], e["speed_max"])
        radius = self.physics["capture_radius"]
        return self._gpu().value(rows, radius, limits).reshape(batch)

    def gradient(self, states, budget):
        """Central difference quotient of the field over one grid spacing per
        public coordinate (the shifted states clipped to the non-periodic
        axes; NaN where they collapse), shape [..., 5]."""
        s = np.asarray(states, dtype=float)
        batch = s.shape[:-1]
        rows = s.reshape(-1, 5)
        m = len(rows)
        tau = np.broadcast_to(np.asarray(budget, dtype=float), batch).reshape(-1)
        # [5 coordinates, 2 sides, m states] shifted copies
        shifted = np.broadcast_to(rows, (5, 2, m, 5)).copy()
        for j, k in enumerate(_STORAGE_OF):
            h = self.spacings[k]
            shifted[j, 0, :, j] -= h
            shifted[j, 1, :, j] += h
            if j != HEADING:
                np.clip(shifted[j, :, :, j], self.axes[k][0], self.axes[k][-1],
                        out=shifted[j, :, :, j])
        values = self.field(shifted.reshape(-1, 5), np.tile(tau, 10)).reshape(5, 2, m)
        width = shifted[np.arange(5), 1, :, np.arange(5)] - shifted[np.arange(5), 0, :,
                                                                     np.arange(5)]
        rise = values[:, 1] - values[:, 0]
        quotient = np.full((5, m), np.nan)
        np.divide(rise, width, out=quotient, where=width > 0)
        return quotient.T.reshape(batch + (5,))

    def hamiltonian_controls(self, states, budget):
        """Commands (acceleration, curvature) of the pursuer minimising and
        the evader maximising the field's rate of change, from the sign of
        the gradient terms they multiply; accelerations that would push a
        speed past its limit are cut to zero."""
        s = np.asarray(states, dtype=float)
        g = self.gradient(s, budget)
        x, y, vp, ve = s[..., 0], s[..., 1], s[..., 3], s[..., 4]
        gx, gy, gpsi, gvp, gve = (g[..., k] for k in range(5))
        p, e = self.physics["pursuer"], self.physics["evader"]

        def limited(acc, v, lim):
            acc = np.where(v <= lim["speed_min"] + 1e-10, np.maximum(acc, 0), acc)
            return np.where(v >= lim["speed_max"] - 1e-10, np.minimum(acc, 0), acc)

        # pursuer turning rotates the relative frame: d psi and d(x, y) terms
        p_turn = vp * (y * gx - x * gy - gpsi)
        e_turn = ve * gpsi
        p_acc = limited(-p["acceleration"] * np.sign(gvp), vp, p)
        e_acc = limited(e["acceleration"] * np.sign(gve), ve, e)
        pursuer = np.stack([p_acc, -p["max_curvature"] * np.sign(p_turn)], axis=-1)
        evader = np.stack([e_acc, e["max_curvature"] * np.sign(e_turn)], axis=-1)
        return pursuer, evader
