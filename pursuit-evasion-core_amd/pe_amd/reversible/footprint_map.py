"""A square room with axis-aligned box obstacles, seen by oriented
rectangular cars: the map of the footprint-map 8D game (gpu_map_game).

Interface of the reference's FootprintMap (reversible/footprint_map.py):
constructor fields, `to_dict` / `from_dict`, `clearance` and `surfaces`.
A pose is the rectangle's centre and heading.  The clearance of a pose is a
signed separation (positive iff the rectangle touches neither a wall nor a
box), formed by the separating-axis test: for each box the largest gap
along the four candidate axes (the world axes and the body axes), then the
minimum over the walls and the boxes.  It is an implicit collision
function, not a Euclidean distance everywhere.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..local_game.geometry import Footprint


@dataclass(frozen=True)
class FootprintMap:
    half_width: float = 2.5
    obstacles: tuple = ()          # boxes (xmin, xmax, ymin, ymax)
    pursuer: Footprint = Footprint(0.8, 0.5)
    evader: Footprint = Footprint(0.4, 0.25)

    def __post_init__(self):
        if not (np.isfinite(self.half_width) and self.half_width > 0):
            raise ValueError("Positive square half-width required")
        for box in self.obstacles:
            b = np.asarray(box, float)
            if b.shape != (4,) or not np.isfinite(b).all() or b[0] >= b[1] or b[2] >= b[3]:
                raise ValueError("Obstacles are [xmin,xmax,ymin,ymax]")

    def to_dict(self):
        return dict(half_width=self.half_width, obstacles=[list(b) for b in self.obstacles],
                    pursuer=self.pursuer.to_dict(), evader=self.evader.to_dict(),
                    pose_reference="rectangle_center")

    @classmethod
    def from_dict(cls, data):
        return cls(data["half_width"], tuple(tuple(b) for b in data["obstacles"]),
                   Footprint(**data["pursuer"]), Footprint(**data["evader"]))

    def _boxes(self):
        """Centres [nb, 2] and half sides [nb, 2] of the boxes."""
        b = np.asarray(self.obstacles, float).reshape(-1, 4)
        centre = np.stack([b[:, 0] + b[:, 1], b[:, 2] + b[:, 3]], -1) / 2
        half = np.stack([b[:, 1] - b[:, 0], b[:, 3] - b[:, 2]], -1) / 2
        return centre, half

    def clearance(self, pose, role="pursuer"):
        """Signed separation of the `role` rectangle at pose [..., x, y,
        heading] from the walls and the boxes."""
        s = np.asarray(pose, dtype=float)
        if s.shape[-1] < 3 or not np.isfinite(s).all():
            raise ValueError("Pose requires finite x,y,heading")
        body = getattr(self, role)
        half_len, half_wid = body.length / 2, body.width / 2
        cos_h, sin_h = np.cos(s[..., 2]), np.sin(s[..., 2])
        ac, asn = np.abs(cos_h), np.abs(sin_h)
        # half extents of the rotated rectangle along the world x and y axes
        reach_x = half_len * ac + half_wid * asn
        reach_y = half_len * asn + half_wid * ac
        walls = np.minimum(self.half_width - np.abs(s[..., 0]) - reach_x,
                           self.half_width - np.abs(s[..., 1]) - reach_y)
        if not self.obstacles:
            return walls
        centre, half = self._boxes()
        dx = s[..., 0, None] - centre[:, 0]
        dy = s[..., 1, None] - centre[:, 1]
        hx, hy = half[:, 0], half[:, 1]
        c, sn, a, b = (v[..., None] for v in (cos_h, sin_h, ac, asn))
        gaps = np.stack([
            np.abs(dx) - hx - reach_x[..., None],                 # world x
            np.abs(dy) - hy - reach_y[..., None],                 # world y
            np.abs(dx * c + dy * sn) - half_len - hx * a - hy * b,   # body length axis
            np.abs(-dx * sn + dy * c) - half_wid - hx * b - hy * a,  # body width axis
        ], axis=-1)
        return np.minimum(walls, gaps.max(-1).min(-1))

    def surfaces(self, state, capture_radius):
        """The first-fault-loses game's surfaces at joint states [..., 8]:
        target = min(centre distance - capture radius, evader clearance)
        (capture or an evader fault ends the game for the pursuer), avoid =
        1e-6 - pursuer clearance (a pursuer fault loses); returns
        (max(target, avoid), avoid)."""
        s = np.asarray(state, float)
        dx, dy = s[..., 0] - s[..., 4], s[..., 1] - s[..., 5]
        gap = np.sqrt(dx * dx + dy * dy) - capture_radius
        target = np.minimum(gap, self.clearance(s[..., 4:], "evader"))
        avoid = 1e-6 - self.clearance(s[..., :4], "pursuer")
        return np.maximum(target, avoid), avoid
