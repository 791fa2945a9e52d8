"""Finite-body geometry of the contact game: footprints, the rectangle-contact
payoff, and arenas of circular obstacles (host-side descriptions).

Same types and conventions as the reference:
  Footprint                 reversible/footprint_map.py:11-28
  rectangle_separation      local_game/geometry.py:15-36
  CircleArena               local_game/geometry.py:39-163
and a seeded random-rectangle map generator for the benchmark maps, drawn
on a 0.05 m occupancy raster and turned into equal-radius cylinders with the
reference's BARN convention (CircleArena.from_barn, geometry.py:105-115).
The per-state geometry (clearance, held-command successors, braking
safety) runs on the GPU in libpe_amd.so (local_game/native.py).
"""

from __future__ import annotations

from dataclasses import asdict, dataclass, field
import hashlib

import numpy as np


@dataclass(frozen=True)
class Footprint:
    """Oriented rectangle; pose refers to its center."""

    length: float
    width: float

    def __post_init__(self):
        if not np.isfinite([self.length, self.width]).all() or min(self.length, self.width) <= 0:
            raise ValueError("Footprint dimensions must be positive and finite")

    @property
    def radius(self) -> float:
        return 0.5 * float(np.hypot(self.length, self.width))

    def to_dict(self):
        return asdict(self)


def rectangle_separation(p, e, p_body=Footprint(0.8, 0.5), e_body=Footprint(0.4, 0.25)):
    """Separating-axis payoff: > 0 iff the two oriented rectangles are apart,
    <= 0 means contact.  Max of the four axis gaps, two in each body frame."""
    p = np.asarray(p, float)
    e = np.asarray(e, float)
    dx = e[..., 0] - p[..., 0]
    dy = e[..., 1] - p[..., 1]
    rel = e[..., 2] - p[..., 2]
    ac, as_ = np.abs(np.cos(rel)), np.abs(np.sin(rel))
    gaps = []
    for pose, own, other in ((p, p_body, e_body), (e, e_body, p_body)):
        c, s = np.cos(pose[..., 2]), np.sin(pose[..., 2])
        along = dx * c + dy * s
        across = -dx * s + dy * c
        gaps.append(np.abs(along) - 0.5 * own.length - 0.5 * (other.length * ac + other.width * as_))
        gaps.append(np.abs(across) - 0.5 * own.width - 0.5 * (other.length * as_ + other.width * ac))
    return np.maximum(np.maximum(gaps[0], gaps[1]), np.maximum(gaps[2], gaps[3]))


@dataclass
class CircleArena:
    """Square room of half-width `half_width` with circular obstacles."""

    circles: np.ndarray
    half_width: float
    pursuer: Footprint = field(default_factory=lambda: Footprint(0.8, 0.5))
    evader: Footprint = field(default_factory=lambda: Footprint(0.4, 0.25))
    source: dict | None = None

    def __post_init__(self):
        self.circles = np.asarray(self.circles, np.float64).reshape(-1, 3)
        if not np.isfinite(self.circles).all() or np.any(self.circles[:, 2] <= 0):
            raise ValueError("Finite circles[x,y,r] with positive radii required")
        if not np.isfinite(self.half_width) or self.half_width <= 0:
            raise ValueError("Positive bounded arena size required")

    def to_dict(self):
        return dict(
            circles=self.circles.tolist(),
            half_width=self.half_width,
            pursuer=self.pursuer.to_dict(),
            evader=self.evader.to_dict(),
            source=self.source,
            capture_rule="oriented_rectangle_contact",
            pose_reference="rectangle_center",
        )

    @classmethod
    def from_dict(cls, d):
        return cls(d["circles"], d["half_width"], Footprint(**d["pursuer"]),
                   Footprint(**d["evader"]), d.get("source"))

    @classmethod
    def from_raster(cls, grid, resolution, **kw):
        """Occupied cells of a square binary raster (row 0 at +y) become
        cylinders of radius resolution/2 at the cell centers; the room is the
        raster's square (the BARN convention of geometry.py:105-115)."""
        grid = np.asarray(grid)
        if grid.ndim != 2 or grid.shape[0] != grid.shape[1] or not np.isin(grid, [0, 1]).all():
            raise ValueError("Expected a square binary raster")
        size = grid.shape[0]
        row, col = np.nonzero(grid)
        centers = np.stack(
            [(col + 0.5 - size / 2) * resolution,
             (size / 2 - row - 0.5) * resolution,
             np.full(len(row), resolution / 2)], -1)
        return cls(centers, size * resolution / 2, **kw)

    def clearance(self, poses, role="pursuer"):
        """Signed rectangle/circle and rectangle/wall separation (numpy)."""
        s = np.asarray(poses, float)
        body = getattr(self, role)
        c, sn = np.cos(s[..., 2]), np.sin(s[..., 2])
        ex = (body.length * abs(c) + body.width * abs(sn)) / 2
        ey = (body.length * abs(sn) + body.width * abs(c)) / 2
        out = np.minimum(self.half_width - abs(s[..., 0]) - ex,
                         self.half_width - abs(s[..., 1]) - ey)
        for x, y, r in self.circles:
            dx, dy = x - s[..., 0], y - s[..., 1]
            qx = abs(dx * c + dy * sn) - body.length / 2
            qy = abs(-dx * sn + dy * c) - body.width / 2
            value = np.hypot(np.maximum(qx, 0), np.maximum(qy, 0)) + np.minimum(np.maximum(qx, qy), 0) - r
            out = np.minimum(out, value)
        return out

    def capture_value(self, state):
        s = np.asarray(state, float)
        return rectangle_separation(s[..., :4], s[..., 4:], self.pursuer, self.evader)


def random_rectangle_raster(seed, *, count=30, side=(0.2, 0.8), half_width=3.0,
                            resolution=0.05):
    """Binary raster of `count` axis-aligned rectangles with sides drawn
    U(side) and centers U(-half_width, half_width)^2 from
    numpy.random.default_rng(seed).  A cell is occupied when its center lies
    inside a rectangle.  Row 0 is the +y edge."""
    rng = np.random.default_rng(seed)
    size = int(round(2 * half_width / resolution))
    sides = rng.uniform(side[0], side[1], (count, 2))
    centers = rng.uniform(-half_width, half_width, (count, 2))
    coords = (np.arange(size) + 0.5 - size / 2) * resolution
    cx = coords[None, :]            # column -> x
    cy = -coords[:, None]           # row -> y (row 0 at +y)
    grid = np.zeros((size, size), np.uint8)
    for (w, h), (x, y) in zip(sides, centers):
        grid |= ((abs(cx - x) <= w / 2) & (abs(cy - y) <= h / 2)).astype(np.uint8)
    return grid, dict(seed=int(seed), count=count, sides=sides.tolist(),
                      centers=centers.tolist(), resolution=resolution)


def random_rectangle_arena(seed, *, pursuer=Footprint(0.5, 0.3), evader=Footprint(0.5, 0.3),
                           count=30, half_width=3.0, resolution=0.05):
    """The benchmark's seeded 30-rectangle map as a CircleArena."""
    grid, info = random_rectangle_raster(seed, count=count, half_width=half_width,
                                         resolution=resolution)
    info.update(kind="seeded random rectangles on an occupancy raster",
                sha256=hashlib.sha256(grid.tobytes()).hexdigest(),
                boundary="square walls at the raster edge")
    return CircleArena.from_raster(grid, resolution, pursuer=pursuer, evader=evader, source=info)
