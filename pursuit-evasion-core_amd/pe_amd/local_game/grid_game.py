"""Compact eight-state semi-Lagrangian contact game: vehicle grids, pair
domain and the GPU backup.

Host-side mirror of the reference's local_game/grid_game.py:

  VehicleGrid       legal 4D single-car states of a regular grid with, per
                    command, the 16-vertex multilinear stencil of the
                    held-command successor (:22-155)
  make_grid         legality (braking backup), commands, successors (:158-196)
  PairDomain        packed pair layout used by the output tables (:336-386)

and the glue to the GPU backup in libpe_amd.so (pe_sl_*): `game_tables`
packs two VehicleGrids into the C-ABI tables, `GameSolver` owns the
device-resident value field.  The geometry (legality, successors,
connections) runs on the GPU too (local_game/native.py).

Value layout on the device: dense [p.n, e.n] over the product of the two
cars' legal states; pairs beyond the exit radius hold the exit value.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import product

import numpy as np

from .. import _native
from .native import GeometryEngine

# Vertex offsets of the 4D multilinear stencil, axis 0 most significant.
BITS = np.asarray(list(product([0, 1], repeat=4)), np.int64)


@dataclass
class VehicleGrid:
    axes: tuple
    states: np.ndarray
    dense_to_compact: np.ndarray
    controls: np.ndarray
    successors: np.ndarray
    weights: np.ndarray
    allowed: np.ndarray
    retained_mass: np.ndarray

    def bind_geometry(self, arena, limits, role, dt=0.05):
        self._engine = GeometryEngine(arena, limits, role, threads=1, control_dt=dt)
        self._tree = None
        return self

    @property
    def shape(self):
        return tuple(len(a) for a in self.axes)

    def position_index(self):
        """Index of each legal state's (x, y) cell, x-major."""
        dense = np.flatnonzero(self.dense_to_compact >= 0)
        nx, ny, nh, nv = self.shape
        return (dense // (nh * nv)).astype(np.int32)

    def positions(self):
        """(x, y) of every cell, index = ix * ny + iy."""
        x, y = np.meshgrid(self.axes[0], self.axes[1], indexing="ij")
        return np.stack([x.ravel(), y.ravel()], -1)

    def _nearest_supported(self, states):
        """Fallback stencil for successors whose 16 vertices are all illegal:
        up to 16 connected legal states among the 64 nearest in scaled
        (x, y, cos, sin, v) features, inverse-square-distance weights
        (grid_game.py:38-79)."""
        from scipy.spatial import cKDTree

        dx = self.axes[0][1] - self.axes[0][0]
        dh = 2 * np.sin(np.pi / len(self.axes[2]))
        dv = np.min(np.diff(self.axes[3]))

        def features(s):
            return np.stack([s[:, 0] / dx, s[:, 1] / dx, np.cos(s[:, 2]) / dh,
                             np.sin(s[:, 2]) / dh, s[:, 3] / dv], -1)

        if self._tree is None:
            self._tree = cKDTree(features(self.states))
        distance, index = self._tree.query(features(states), k=min(64, len(self.states)))
        distance = distance.reshape(len(states), -1)
        index = index.reshape(len(states), -1)
        target = self.states[index]
        linked = self._engine.connections(
            np.broadcast_to(states[:, None], target.shape).reshape(-1, 4),
            target.reshape(-1, 4),
        ).reshape(index.shape)
        ids = np.zeros((len(states), 16), np.int32)
        wts = np.zeros((len(states), 16), np.float32)
        ok = np.zeros(len(states), bool)
        for row in range(len(states)):
            chosen = np.flatnonzero(linked[row])[:16]
            if len(chosen) == 0:
                continue
            w = 1.0 / np.maximum(distance[row, chosen], 1e-8) ** 2
            ids[row, : len(chosen)] = index[row, chosen]
            wts[row, : len(chosen)] = w / w.sum()
            ok[row] = True
        return ids, wts, ok

    def stencil(self, states):
        """16-vertex multilinear stencil of states[..., 4] on the grid.

        Vertices on illegal configurations get weight zero and the rest are
        renormalised; heading is periodic, the other axes clip within 1e-9.
        Returns compact ids [m,16] int32, weights [m,16] float32, validity,
        the input batch shape and the retained mass (grid_game.py:81-123)."""
        s = np.asarray(states, float)
        shape = s.shape[:-1]
        flat = s.reshape(-1, 4)
        lens = [len(a) for a in self.axes]
        strides = [lens[1] * lens[2] * lens[3], lens[2] * lens[3], lens[3], 1]
        dense = np.zeros((len(flat), 16), np.int64)
        w = np.ones((len(flat), 16))
        valid = np.isfinite(flat).all(1)
        for k, axis in enumerate(self.axes):
            x = flat[:, k]
            n = len(axis)
            if k == 2:
                cell = ((x + np.pi) % (2 * np.pi)) / (2 * np.pi / n)
                j = np.floor(cell).astype(np.int64) % n
                frac = cell - np.floor(cell)
            else:
                valid &= (x >= axis[0] - 1e-9) & (x <= axis[-1] + 1e-9)
                z = np.clip(x, axis[0], axis[-1])
                j = np.clip(np.searchsorted(axis, z, side="right") - 1, 0, n - 2)
                frac = (z - axis[j]) / (axis[j + 1] - axis[j])
            coord = j[:, None] + BITS[None, :, k]
            if k == 2:
                coord %= n
            dense += coord * strides[k]
            w *= np.where(BITS[None, :, k], frac[:, None], 1 - frac[:, None])
        ids = self.dense_to_compact[dense]
        w = np.where(ids >= 0, w, 0.0)
        mass = w.sum(1)
        missing = valid & (mass <= 1e-9)
        valid &= mass > 1e-9
        w = np.divide(w, mass[:, None], out=np.zeros_like(w), where=mass[:, None] > 1e-9)
        ids = np.maximum(ids, 0).astype(np.int32)
        w = w.astype(np.float32)
        if missing.any() and hasattr(self, "_engine"):
            unique, inverse = np.unique(flat[missing], axis=0, return_inverse=True)
            ni, nw, nv = self._nearest_supported(unique)
            inverse = inverse.ravel()
            ids[missing] = ni[inverse]
            w[missing] = nw[inverse]
            valid[missing] = nv[inverse]
        return ids, w, valid, shape, mass

    def save(self, path):
        np.savez_compressed(
            path, states=self.states, dense_to_compact=self.dense_to_compact,
            controls=self.controls, successors=self.successors, weights=self.weights,
            allowed=self.allowed, retained_mass=self.retained_mass,
            **{"axis" + str(i): a for i, a in enumerate(self.axes)},
        )

    @classmethod
    def load(cls, path):
        z = np.load(path)
        return cls(tuple(z["axis" + str(i)] for i in range(4)),
                   *[z[k] for k in ("states", "dense_to_compact", "controls", "successors",
                                    "weights", "allowed", "retained_mass")])


def grid_axes(half_width, *, position_step=0.3, positions=None, headings=8, speeds=()):
    """Axes (x, y, heading, speed).  With `positions`, x = y =
    linspace(-half, half, positions); otherwise the reference's cell-centred
    axis (n + 2 points spaced position_step, grid_game.py:163-164)."""
    if positions is None:
        n = int(round(2 * half_width / position_step))
        x = (np.arange(n + 2) - (n + 1) / 2) * position_step
    else:
        x = np.linspace(-half_width, half_width, positions)
    return (x, x.copy(), np.arange(headings) * 2 * np.pi / headings - np.pi,
            np.asarray(speeds, float))


def make_grid(arena, limits, role, *, position_step=0.3, headings=8, speeds=None, dt=0.05,
              positions=None, half_width=None):
    """Legal states, commands {-A,0,A} x {-K,0,K} (acceleration-major) and the
    successor stencils of one car (grid_game.py:158-196)."""
    return make_grids([arena], limits, role, position_step=position_step, headings=headings,
                      speeds=speeds, dt=dt, positions=positions, half_width=half_width)[0]


def make_grids(arenas, limits, role, *, position_step=0.3, headings=8, speeds=None, dt=0.05,
               positions=None, half_width=None):
    """make_grid for several arenas of one room size with one car model: the
    legality and the action tables of all maps each run as one GPU launch
    (pe_geom_backups_multi, pe_geom_actions_multi), then each map's
    stencils are formed on the host.  Equal to make_grid map by map."""
    from .native import actions_many, backup_safe_many

    if speeds is None:
        speeds = [-1, 0, 0.5, 1, 2] if role == "pursuer" else [-0.5, 0, 0.25, 0.5, 1]
    halves = {a.half_width for a in arenas}
    if half_width is None and len(halves) != 1:
        raise ValueError("One room half-width is required for a batch of arenas")
    half = arenas[0].half_width if half_width is None else half_width
    axes = grid_axes(half, position_step=position_step, positions=positions,
                     headings=headings, speeds=speeds)
    states = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 4)
    engines = [GeometryEngine(a, limits, role, control_dt=dt) for a in arenas]
    legal_all = backup_safe_many(engines, [states] * len(arenas))
    a, k = limits.acceleration, limits.max_curvature
    controls = np.array([[u, c] for u in (-a, 0, a) for c in (-k, 0, k)], float)
    kept_all = [states[legal] for legal in legal_all]
    acts = actions_many(engines, kept_all, controls, dt)
    return [_finish_grid(axes, states, legal, kept, controls, allowed, nxt, arena, limits,
                         role, dt)
            for legal, kept, (allowed, nxt), arena in zip(legal_all, kept_all, acts, arenas)]


def _finish_grid(axes, states, legal, kept, controls, allowed, nxt, arena, limits, role, dt):
    mapping = np.full(len(states), -1, np.int32)
    mapping[legal] = np.arange(legal.sum(), dtype=np.int32)
    grid = VehicleGrid(axes, kept, mapping, controls, None, None, None, None)
    grid.bind_geometry(arena, limits, role, dt)
    idx, w, valid, _, mass = grid.stencil(nxt)
    grid.successors = idx.reshape(len(kept), len(controls), 16)
    grid.weights = w.reshape(len(kept), len(controls), 16)
    grid.allowed = allowed & valid.reshape(allowed.shape)
    grid.retained_mass = mass.reshape(allowed.shape)
    if not grid.allowed.any(axis=1).all():
        raise ValueError("A retained state has no legal interpolable action")
    return grid


@dataclass
class PairDomain:
    """Packed layout of the pairs inside the exit radius (grid_game.py:336-386):
    per distinct P position the list of E states within the radius; pair
    (ip, ie) is stored at offsets[ip] + lookup[p_position[ip], ie]."""

    p_position: np.ndarray
    lookup: np.ndarray
    e_indices: np.ndarray
    counts: np.ndarray
    offsets: np.ndarray
    exit_radius: float

    @property
    def states(self):
        return int(self.offsets[-1])

    @classmethod
    def build(cls, p, e, exit_radius=3.0):
        positions, which = np.unique(p.states[:, :2], axis=0, return_inverse=True)
        which = which.ravel()
        lookup = np.full((len(positions), len(e.states)), -1, np.int32)
        members = []
        for i, xy in enumerate(positions):
            ids = np.flatnonzero(np.linalg.norm(e.states[:, :2] - xy, axis=1) <= exit_radius + 1e-8)
            members.append(ids)
            lookup[i, ids] = np.arange(len(ids), dtype=np.int32)
        counts = np.array([len(m) for m in members], np.int32)
        inverse = np.zeros((len(positions), int(counts.max())), np.int32)
        for i, m in enumerate(members):
            inverse[i, : len(m)] = m
        offsets = np.concatenate([[0], np.cumsum(counts[which], dtype=np.int64)])
        return cls(which.astype(np.int32), lookup, inverse, counts, offsets, float(exit_radius))

    def pack(self, dense):
        """Dense [p.n, e.n] field -> packed vector of the inside pairs."""
        out = np.empty(self.states, np.float32)
        for ip in range(len(self.p_position)):
            pos = self.p_position[ip]
            ids = self.e_indices[pos, : self.counts[pos]]
            out[self.offsets[ip]: self.offsets[ip + 1]] = dense[ip, ids]
        return out

    def unpack(self, packed, exit_value, ne):
        dense = np.full((len(self.p_position), ne), exit_value, np.float32)
        for ip in range(len(self.p_position)):
            pos = self.p_position[ip]
            ids = self.e_indices[pos, : self.counts[pos]]
            dense[ip, ids] = packed[self.offsets[ip]: self.offsets[ip + 1]]
        return dense

    def save(self, path):
        np.savez_compressed(path, p_position=self.p_position, lookup=self.lookup,
                            e_indices=self.e_indices, counts=self.counts,
                            offsets=self.offsets, exit_radius=self.exit_radius)

    @classmethod
    def load(cls, path):
        z = np.load(path)
        return cls(*[z[k] for k in ("p_position", "lookup", "e_indices", "counts", "offsets")],
                   float(z["exit_radius"]))


def successor_reach(grid):
    """Largest (x, y) distance from a state to a vertex of its successor
    stencils with positive weight (grid_game.py:464-472)."""
    reach = 0.0
    for start in range(0, len(grid.states), 1024):
        ids = grid.successors[start: start + 1024]
        d = np.linalg.norm(grid.states[ids, :2] - grid.states[start: start + 1024, None, None, :2],
                           axis=-1)
        reach = max(reach, float(np.where(grid.weights[start: start + 1024] > 0, d, 0).max()))
    return reach


def inside_table(p, e, exit_radius, slack=0.0):
    """[P cells, E cells]: centre distance <= exit radius (+ slack), in the
    PairDomain rule (norm <= radius + 1e-8)."""
    pp, ep = p.positions(), e.positions()
    d = np.linalg.norm(pp[:, None, :] - ep[None, :, :], axis=-1)
    return (d <= exit_radius + 1e-8 + slack).astype(np.uint8)


def game_tables(p, e, arena, exit_radius=3.0):
    """The C-ABI tables (pe_sl_game) of one map."""
    def car(g):
        return dict(states=g.states, successors=g.successors, weights=g.weights,
                    allowed=g.allowed, position=g.position_index(),
                    positions=len(g.axes[0]) * len(g.axes[1]))

    reach = successor_reach(e) + 1e-4
    return dict(
        p=car(p), e=car(e),
        inside=inside_table(p, e, exit_radius),
        near=inside_table(p, e, exit_radius, slack=reach + 1e-6),
        exit_value=float(exit_radius),
        p_body=(arena.pursuer.length, arena.pursuer.width),
        e_body=(arena.evader.length, arena.evader.width),
    )


def dense_inside(p, e, exit_radius):
    """[p.n, e.n] mask of the pairs the backup updates."""
    t = inside_table(p, e, exit_radius)
    return t[p.position_index()[:, None], e.position_index()[None, :]].astype(bool)


class GameSolver(_native.SL):
    """Device-resident value iteration of one or more contact-game maps."""

    def __init__(self, tables, precision=0):
        super().__init__(tables if isinstance(tables, (list, tuple)) else [tables], precision)
