"""Car grids, the pair domain and the GPU backup of the compact 8D contact
game.

The public names and stored fields are the reference's interface
(local_game/grid_game.py): `VehicleGrid` (fields axes, states,
dense_to_compact, controls, successors, weights, allowed, retained_mass;
`stencil`, `save`, `load`, `bind_geometry`), `make_grid`, and `PairDomain`
(`build`, `save`, `load`).  The computation is this package's own:

* legality, successors and the connection tests run as FP64 HIP kernels
  (pe_geom_*, several maps per launch);
* each successor's 16-vertex stencil over the legal states, with the
  nearest-supported fallback, is formed on the GPU (pe_car_stencil);
* `GameSolver` keeps the value field of one or more maps on the device
  (pe_sl_*).

Device value layout: dense [p.n, e.n] over the product of the two cars'
legal states; pairs beyond the exit radius hold the exit value.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .. import _native
from .native import GeometryEngine, actions_many, backup_safe_many

# Fields of a VehicleGrid written by save() besides its four axes.
_STORED = ("states", "dense_to_compact", "controls", "successors", "weights", "allowed",
           "retained_mass")


@dataclass
class VehicleGrid:
    """Legal single-car states of a regular (x, y, heading, speed) grid and,
    per command, the stencil of the held-command successor."""

    axes: tuple
    states: np.ndarray
    dense_to_compact: np.ndarray
    controls: np.ndarray
    successors: np.ndarray
    weights: np.ndarray
    allowed: np.ndarray
    retained_mass: np.ndarray
    geometry: GeometryEngine | None = field(default=None, repr=False, compare=False)

    def bind_geometry(self, arena, limits, role, dt=0.05):
        """Attach the map: enables the nearest-supported fallback."""
        self.geometry = GeometryEngine(arena, limits, role, control_dt=dt)
        return self

    @property
    def shape(self):
        return tuple(len(a) for a in self.axes)

    def position_index(self):
        """Index of each legal state's (x, y) cell, x-major."""
        dense = np.flatnonzero(np.asarray(self.dense_to_compact) >= 0)
        return (dense // (self.shape[2] * self.shape[3])).astype(np.int32)

    def positions(self):
        """(x, y) of every cell, index = ix * ny + iy."""
        return np.stack(np.meshgrid(self.axes[0], self.axes[1], indexing="ij"), -1).reshape(-1, 2)

    def stencil(self, states):
        """Stencil of states[..., 4] on this grid (pe_car_stencil): compact
        ids [m,16] int32, weights [m,16] float32, validity [m], the batch
        shape and the retained multilinear mass [m] (the reference's return
        order, grid_game.py:81-123)."""
        q = np.asarray(states, np.float64)
        batch = q.shape[:-1]
        g = self.geometry
        spec = None if g is None else (g.obstacles, g.arena.half_width, g.params)
        ids, w, valid, mass, _ = _native.car_stencil(self.axes, self.dense_to_compact,
                                                     q.reshape(-1, 4), spec)
        return ids, w, valid, batch, mass

    def save(self, path):
        arrays = {name: getattr(self, name) for name in _STORED}
        arrays.update({f"axis{k}": a for k, a in enumerate(self.axes)})
        np.savez_compressed(path, **arrays)

    @classmethod
    def load(cls, path):
        with np.load(path) as z:
            return cls(axes=tuple(z[f"axis{k}"] for k in range(4)),
                       **{name: z[name] for name in _STORED})


def grid_axes(half_width, *, position_step=0.3, positions=None, headings=8, speeds=()):
    """Axes (x, y, heading, speed).  With `positions`, x = y =
    linspace(-half, half, positions); otherwise the reference's cell-centred
    axis: n + 2 points spaced position_step about 0, n = round(2 half / step)
    (grid_game.py:163-164)."""
    if positions is None:
        cells = int(round(2 * half_width / position_step))
        x = position_step * (np.arange(cells + 2) - 0.5 * (cells + 1))
    else:
        x = np.linspace(-half_width, half_width, positions)
    # heading values as the reference forms them (k * 2 * pi / n - pi), so
    # that states, successors and tables agree to the bit
    heading = np.arange(headings) * 2 * np.pi / headings - np.pi
    return (x, x.copy(), heading, np.asarray(speeds, float))


def command_set(limits):
    """Held commands (acceleration, curvature), acceleration-major over
    {-A, 0, A} x {-K, 0, K} (grid_game.py:177-184)."""
    acc = np.array([-limits.acceleration, 0.0, limits.acceleration])
    cur = np.array([-limits.max_curvature, 0.0, limits.max_curvature])
    return np.stack(np.meshgrid(acc, cur, indexing="ij"), -1).reshape(-1, 2)


def make_grid(arena, limits, role, *, position_step=0.3, headings=8, speeds=None, dt=0.05,
              positions=None, half_width=None):
    """One car's grid on one map (the reference's make_grid,
    grid_game.py:158-196)."""
    return make_grids([arena], limits, role, position_step=position_step, headings=headings,
                      speeds=speeds, dt=dt, positions=positions, half_width=half_width)[0]


def make_grids(arenas, limits, role, *, position_step=0.3, headings=8, speeds=None, dt=0.05,
               positions=None, half_width=None):
    """make_grid for several maps of one room size and one car model:
    legality and successors of every map in one GPU launch each, then each
    map's stencils on the GPU.  Map by map equal to make_grid."""
    if speeds is None:
        speeds = [-1, 0, 0.5, 1, 2] if role == "pursuer" else [-0.5, 0, 0.25, 0.5, 1]
    if half_width is None:
        sizes = {a.half_width for a in arenas}
        if len(sizes) != 1:
            raise ValueError("One room half-width is required for a batch of arenas")
        half_width = sizes.pop()
    axes = grid_axes(half_width, position_step=position_step, positions=positions,
                     headings=headings, speeds=speeds)
    full = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 4)
    engines = [GeometryEngine(a, limits, role, control_dt=dt) for a in arenas]
    legal = backup_safe_many(engines, [full] * len(arenas))
    commands = command_set(limits)
    kept = [full[mask] for mask in legal]
    moves = actions_many(engines, kept, commands, dt)
    grids = []
    for mask, states, (ok, nxt), engine in zip(legal, kept, moves, engines):
        numbering = np.full(len(full), -1, np.int32)
        numbering[mask] = np.arange(len(states), dtype=np.int32)
        grid = VehicleGrid(axes, states, numbering, commands, None, None, None, None, engine)
        ids, w, valid, _, mass = grid.stencil(nxt.reshape(-1, 4))
        shape = (len(states), len(commands))
        grid.successors = ids.reshape(shape + (16,))
        grid.weights = w.reshape(shape + (16,))
        grid.allowed = ok & valid.reshape(shape)
        grid.retained_mass = mass.reshape(shape)
        if not grid.allowed.any(axis=1).all():
            raise ValueError("A retained state has no legal interpolable action")
        grids.append(grid)
    return grids


@dataclass
class PairDomain:
    """Packed layout of the pairs inside the exit radius (the reference's
    PairDomain, grid_game.py:336-386): per distinct pursuer position, the
    ascending list of evader states within the radius; pair (ip, ie) is
    stored at offsets[ip] + lookup[p_position[ip], ie]."""

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
        """Membership: Euclidean centre distance <= exit_radius + 1e-8."""
        cells, cell_of = np.unique(p.states[:, :2], axis=0, return_inverse=True)
        cell_of = cell_of.reshape(-1).astype(np.int32)
        delta = e.states[None, :, :2] - cells[:, None, :]
        member = np.sqrt((delta * delta).sum(-1)) <= exit_radius + 1e-8
        counts = member.sum(1).astype(np.int32)
        lookup = np.where(member, np.cumsum(member, 1) - 1, -1).astype(np.int32)
        width = int(counts.max()) if len(counts) else 0
        # members first, each group in ascending E order
        order = np.argsort(~member, axis=1, kind="stable")[:, :width]
        slot = np.arange(width)[None, :] < counts[:, None]
        e_indices = np.where(slot, order, 0).astype(np.int32)
        offsets = np.zeros(len(cell_of) + 1, np.int64)
        np.cumsum(counts[cell_of], out=offsets[1:])
        return cls(cell_of, lookup, e_indices, counts, offsets, float(exit_radius))

    def _index(self):
        """Row and column of every packed pair in the dense [p.n, e.n] field."""
        per_row = self.counts[self.p_position]
        rows = np.repeat(np.arange(len(self.p_position)), per_row)
        slot = np.arange(self.e_indices.shape[1])[None, :] < per_row[:, None]
        cols = self.e_indices[self.p_position][slot]
        return rows, cols

    def pack(self, dense):
        """Dense [p.n, e.n] field -> packed vector of the inside pairs."""
        rows, cols = self._index()
        return np.asarray(dense, np.float32)[rows, cols]

    def unpack(self, packed, exit_value, ne):
        dense = np.full((len(self.p_position), ne), exit_value, np.float32)
        rows, cols = self._index()
        dense[rows, cols] = packed
        return dense

    def save(self, path):
        np.savez_compressed(path, exit_radius=self.exit_radius,
                            **{k: getattr(self, k) for k in ("p_position", "lookup", "e_indices",
                                                             "counts", "offsets")})

    @classmethod
    def load(cls, path):
        with np.load(path) as z:
            return cls(z["p_position"], z["lookup"], z["e_indices"], z["counts"], z["offsets"],
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
