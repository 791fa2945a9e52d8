"""Queries of a saved 5D relative-game history: the drop-in for
reversible/relative_table.py.

Same constructor checks, `max_budget` view, `stencil`, `field`, `value`,
`gradient` and `hamiltonian_controls` (relative_table.py:11-215).  `field`
and `value` run on the GPU (pe_rtable: the history is copied to device
memory on the first query and stays there); `stencil` returns the host
arrays the reference returns, and `gradient` / `hamiltonian_controls` are
the reference's host formulas over `field`.
"""

from __future__ import annotations

from itertools import product
import json
from pathlib import Path

import numpy as np

from .. import _native
from .free_dp import physics_signature


class RelativeTable:
    state_dimension = 5

    def __init__(self, folder, config=None, *, max_budget=None):
        self.folder = Path(folder).resolve()
        self.meta = json.loads((self.folder / "meta.json").read_text())
        if self.meta.get("schema") != "reversible_relative5d_v1":
            raise ValueError("Not a reversible relative5D table")
        if self.meta["status"] != "complete" or not self.meta["history_saved"]:
            raise ValueError("A completed solve with signed history is required")
        if self.meta["tube_projection"] != "min over every internal update":
            raise ValueError("Unsupported budget convention")
        source_times = np.asarray(self.meta["times"], dtype=float)
        self.times = source_times.copy()
        self.axes = tuple(np.asarray(a, dtype=float) for a in self.meta["axes"])
        self.shape = tuple(len(a) for a in self.axes)
        if len(self.axes) != 5 or min(self.shape) < 2:
            raise ValueError("Expected five nondegenerate coordinate axes")
        self.spacings = np.array([a[1] - a[0] for a in self.axes])
        for a, h in zip(self.axes, self.spacings):
            if not np.isfinite(a).all() or not np.allclose(np.diff(a), h) or h <= 0:
                raise ValueError("Axes must be finite, increasing and uniformly spaced")
        if not np.isclose(self.axes[2][0], -np.pi) or not np.isclose(
                self.spacings[2] * self.shape[2], 2 * np.pi):
            raise ValueError("Heading axis must be periodic without a repeated endpoint")
        if "shape" in self.meta and tuple(self.meta["shape"]) != self.shape:
            raise ValueError("Metadata shape differs from coordinate axes")
        if len(self.times) < 2 or np.any(np.diff(self.times) <= 0):
            raise ValueError("Invalid output times")
        if max_budget is not None:
            hit = np.flatnonzero(np.isclose(source_times, max_budget, rtol=0, atol=1e-10))
            if not len(hit) or hit[0] == 0:
                raise ValueError("max_budget must be a positive retained output budget")
            self.times = source_times[: hit[0] + 1].copy()
        expected = int(np.prod(self.shape)) * len(source_times) * 4
        if (self.folder / "fields.f32").stat().st_size != expected:
            raise ValueError("History byte length differs from metadata")
        self.fields = np.memmap(self.folder / "fields.f32", mode="r", dtype=np.float32,
                                shape=(len(source_times), *self.shape))[: len(self.times)]
        self._strides = np.cumprod((1,) + self.shape[:0:-1])[::-1]
        self._bits = np.array(list(product([0, 1], repeat=5)), dtype=np.int64)
        self.physics = self.meta["physics"]
        self.horizon = float(self.times[-1])
        self._device = None
        if config is not None:
            self.assert_compatible(config)

    def assert_compatible(self, config):
        if physics_signature(config) != self.physics:
            raise ValueError("The free-space table does not match the vehicle/capture model")

    def _gpu(self):
        """The history on the device (uploaded once)."""
        if self._device is None:
            flat = np.ascontiguousarray(self.fields.reshape(len(self.times), -1))
            self._device = _native.RelTable(flat, self.axes, self.times)
        return self._device

    def stencil(self, states):
        """32-vertex indices and weights, validity and batch shape
        (relative_table.py:76-109)."""
        states = np.asarray(states, dtype=float)
        if states.ndim < 1 or states.shape[-1] != 5:
            raise ValueError("Relative states must end in [x,y,psi,vP,vE]")
        shape = states.shape[:-1]
        q_all = states.reshape(-1, 5)[:, [3, 4, 2, 1, 0]]
        finite = np.isfinite(q_all)
        valid = finite.all(1)
        q_all = np.where(finite, q_all, 0)
        ii = np.zeros((len(q_all), 32), dtype=np.int64)
        ww = np.ones((len(q_all), 32))
        for k, (axis, step, n) in enumerate(zip(self.axes, self.spacings, self.shape)):
            q = q_all[:, k]
            if k == 2:
                f = ((q + np.pi) % (2 * np.pi)) / step
                low = np.floor(f).astype(np.int64) % n
                frac = f - np.floor(f)
            else:
                valid &= (q >= axis[0] - 1e-10) & (q <= axis[-1] + 1e-10)
                f = (np.clip(q, axis[0], axis[-1]) - axis[0]) / step
                low = np.minimum(np.floor(f).astype(np.int64), n - 2)
                frac = f - low
            node = low[:, None] + self._bits[None, :, k]
            if k == 2:
                node %= n
            ii += node * self._strides[k]
            ww *= np.where(self._bits[None, :, k], frac[:, None], 1 - frac[:, None])
        return ii, ww, valid, shape

    def _batch(self, states):
        s = np.asarray(states, dtype=float)
        if s.ndim < 1 or s.shape[-1] != 5:
            raise ValueError("Relative states must end in [x,y,psi,vP,vE]")
        return s.reshape(-1, 5), s.shape[:-1]

    def field(self, states, budget):
        """Signed field at states and budgets (relative_table.py:111-131)."""
        flat, shape = self._batch(states)
        tau = np.broadcast_to(np.asarray(budget, dtype=float), shape).reshape(-1)
        if not len(flat):
            return np.full(shape, np.nan)
        return self._gpu().field(flat, tau).reshape(shape)

    def value(self, states):
        """First crossing of the interpolated signed history, never an
        interpolated arrival (relative_table.py:133-161)."""
        if self.meta.get("capture_rule") == "rectangle_contact":
            raise ValueError(
                "Use local_game.free_value.ContactFreeValue for rectangular time queries")
        flat, shape = self._batch(states)
        if not len(flat):
            return np.full(shape, np.nan)
        p, e = self.physics["pursuer"], self.physics["evader"]
        limits = [p["speed_min"], p["speed_max"], e["speed_min"], e["speed_max"]]
        return self._gpu().value(flat, self.physics["capture_radius"], limits).reshape(shape)

    def gradient(self, states, budget):
        """Central differences of the signed field over one spacing, clipped
        at the non-periodic ends, in public [x, y, psi, vP, vE] order
        (relative_table.py:163-189)."""
        s = np.asarray(states, dtype=float)
        out = []
        for j, k in enumerate([4, 3, 2, 0, 1]):
            h = self.spacings[k]
            left, right = s.copy(), s.copy()
            left[..., j] -= h
            right[..., j] += h
            if j != 2:
                left[..., j] = np.clip(left[..., j], self.axes[k][0], self.axes[k][-1])
                right[..., j] = np.clip(right[..., j], self.axes[k][0], self.axes[k][-1])
            width = right[..., j] - left[..., j]
            diff = self.field(right, budget) - self.field(left, budget)
            out.append(np.divide(diff, width, out=np.full_like(diff, np.nan), where=width > 0))
        return np.stack(out, axis=-1)

    def hamiltonian_controls(self, states, budget):
        """Pointwise P-min / E-max controls of the interpolated field gradient
        (relative_table.py:191-215)."""
        s = np.asarray(states, dtype=float)
        g = self.gradient(s, budget)
        p, e = self.physics["pursuer"], self.physics["evader"]
        cp = s[..., 3] * (s[..., 1] * g[..., 0] - s[..., 0] * g[..., 1] - g[..., 2])
        ce = s[..., 4] * g[..., 2]
        ap = -p["acceleration"] * np.sign(g[..., 3])
        ae = e["acceleration"] * np.sign(g[..., 4])
        ap = np.where(s[..., 3] <= p["speed_min"] + 1e-10, np.maximum(ap, 0), ap)
        ap = np.where(s[..., 3] >= p["speed_max"] - 1e-10, np.minimum(ap, 0), ap)
        ae = np.where(s[..., 4] <= e["speed_min"] + 1e-10, np.maximum(ae, 0), ae)
        ae = np.where(s[..., 4] >= e["speed_max"] - 1e-10, np.minimum(ae, 0), ae)
        return (np.stack([ap, -p["max_curvature"] * np.sign(cp)], axis=-1),
                np.stack([ae, e["max_curvature"] * np.sign(ce)], axis=-1))
