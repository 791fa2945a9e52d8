"""Oriented rectangular footprints in an axis-aligned obstacle map
(reversible/footprint_map.py:11-145): the geometry behind the footprint-map
8D game.  Positions refer to footprint centres; the signed separation is an
implicit collision function, not a Euclidean distance everywhere.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..local_game.geometry import Footprint


@dataclass(frozen=True)
class FootprintMap:
    """Square room of half-width `half_width` with boxes
    [xmin, xmax, ymin, ymax] (footprint_map.py:31-66)."""

    half_width: float = 2.5
    obstacles: tuple = ()
    pursuer: Footprint = Footprint(0.8, 0.5)
    evader: Footprint = Footprint(0.4, 0.25)

    def __post_init__(self):
        if not np.isfinite(self.half_width) or self.half_width <= 0:
            raise ValueError("Positive square half-width required")
        for box in self.obstacles:
            if len(box) != 4 or not np.isfinite(box).all() or box[0] >= box[1] or box[2] >= box[3]:
                raise ValueError("Obstacles are [xmin,xmax,ymin,ymax]")

    def to_dict(self):
        return dict(half_width=self.half_width, obstacles=[list(b) for b in self.obstacles],
                    pursuer=self.pursuer.to_dict(), evader=self.evader.to_dict(),
                    pose_reference="rectangle_center")

    @classmethod
    def from_dict(cls, data):
        return cls(data["half_width"], tuple(tuple(b) for b in data["obstacles"]),
                   Footprint(**data["pursuer"]), Footprint(**data["evader"]))

    def clearance(self, pose, role="pursuer"):
        """Positive iff the rectangle at pose [..., x, y, heading] touches no
        obstacle or wall: the minimum over the walls and the boxes of the
        largest separating-axis gap (world axes and the two body axes)
        (footprint_map.py:68-97)."""
        s = np.asarray(pose, dtype=float)
        if s.shape[-1] < 3 or not np.isfinite(s).all():
            raise ValueError("Pose requires finite x,y,heading")
        body = getattr(self, role)
        x, y, h = s[..., 0], s[..., 1], s[..., 2]
        c, sn = np.cos(h), np.sin(h)
        ac, asn = abs(c), abs(sn)
        hl, hw = body.length / 2, body.width / 2
        ex, ey = hl * ac + hw * asn, hl * asn + hw * ac  # half extents on x, y
        out = np.minimum(self.half_width - abs(x) - ex, self.half_width - abs(y) - ey)
        for xmin, xmax, ymin, ymax in self.obstacles:
            ox, oy = (xmin + xmax) / 2, (ymin + ymax) / 2
            hx, hy = (xmax - xmin) / 2, (ymax - ymin) / 2
            dx, dy = x - ox, y - oy
            gaps = np.stack([
                abs(dx) - hx - ex,
                abs(dy) - hy - ey,
                abs(dx * c + dy * sn) - hl - hx * ac - hy * asn,
                abs(-dx * sn + dy * c) - hw - hx * asn - hy * ac,
            ], axis=-1)
            out = np.minimum(out, gaps.max(axis=-1))
        return out

    def point_sdf(self, xy):
        """Signed point clearance of the room and boxes (footprint_map.py:99-111)."""
        xy = np.asarray(xy, float)
        x, y = xy[..., 0], xy[..., 1]
        out = self.half_width - np.maximum(abs(x), abs(y))
        for xmin, xmax, ymin, ymax in self.obstacles:
            qx = abs(x - (xmin + xmax) / 2) - (xmax - xmin) / 2
            qy = abs(y - (ymin + ymax) / 2) - (ymax - ymin) / 2
            box = np.hypot(np.maximum(qx, 0), np.maximum(qy, 0)) + np.minimum(np.maximum(qx, qy), 0)
            out = np.minimum(out, box)
        return out

    def surfaces(self, state, capture_radius):
        """First-fault-loses game: target capture or E fault, avoid P fault
        (footprint_map.py:113-121); returns (max(target, avoid), avoid)."""
        s = np.asarray(state, float)
        cap = np.linalg.norm(s[..., :2] - s[..., 4:6], axis=-1) - capture_radius
        target = np.minimum(cap, self.clearance(s[..., 4:], "evader"))
        avoid = -self.clearance(s[..., :4], "pursuer") + 1e-6
        return np.maximum(target, avoid), avoid

    def event(self, state, capture_radius):
        """Terminal event of a joint state; a P fault wins ties
        (footprint_map.py:123-131)."""
        s = np.asarray(state, float)
        if self.clearance(s[:4], "pursuer") <= 0:
            return "pursuer_collision"
        if np.linalg.norm(s[:2] - s[4:6]) <= capture_radius:
            return "capture"
        if self.clearance(s[4:], "evader") <= 0:
            return "evader_collision"
        return None
