"""Reversible-car game contract: vehicle limits and the game configuration.

Same fields, defaults, validation and JSON form as the reference contract
(reversible/model.py:13-79), so a configuration file such as
configs/reversible_free_a3.json loads unchanged.  World state per car is
[x, y, heading, signed speed]; a command is [acceleration, curvature].
"""

from __future__ import annotations

from dataclasses import asdict, dataclass, field
import math

import numpy as np


@dataclass(frozen=True)
class VehicleLimits:
    """Signed speed interval, acceleration bound and curvature bound of a car
    (reversible/model.py:13-33)."""

    speed_min: float
    speed_max: float
    acceleration: float
    max_curvature: float

    def __post_init__(self):
        vals = (self.speed_min, self.speed_max, self.acceleration, self.max_curvature)
        if not all(math.isfinite(v) for v in vals) or not (
            self.speed_min < 0 < self.speed_max
        ):
            raise ValueError("Signed finite speed bounds must straddle zero.")
        if self.acceleration <= 0 or self.max_curvature <= 0:
            raise ValueError("Acceleration and curvature bounds must be positive.")

    @property
    def max_abs_speed(self) -> float:
        return max(-self.speed_min, self.speed_max)

    @property
    def minimum_turn_radius(self) -> float:
        return 1.0 / self.max_curvature


_RULES = ("first_fault_loses", "collision_constrained")
_CAPTURE = ("center_distance", "rectangle_contact")


@dataclass(frozen=True)
class ReversibleConfig:
    """Game configuration (reversible/model.py:36-79)."""

    name: str = "reversible_cars"
    pursuer: VehicleLimits = field(default=VehicleLimits(-1.0, 2.0, 2.0, 1.0))
    evader: VehicleLimits = field(default=VehicleLimits(-0.5, 1.0, 1.0, 2.0))
    capture_radius: float = 0.3
    control_dt: float = 0.05
    arena_half_width: float = 3.0
    boundary_rule: str = "first_fault_loses"
    point_vehicles: bool = True
    capture_rule: str = "center_distance"

    def __post_init__(self):
        positive = (self.capture_radius, self.control_dt, self.arena_half_width)
        if not all(math.isfinite(v) for v in positive) or min(positive) <= 0:
            raise ValueError(
                "Radius, control interval and square extent must be positive."
            )
        if self.boundary_rule not in _RULES or not isinstance(self.point_vehicles, bool):
            raise ValueError(
                "Use a declared boundary rule and an explicit point/footprint flag."
            )
        if self.capture_rule not in _CAPTURE:
            raise ValueError("Unknown capture rule")
        if self.point_vehicles and (
            self.capture_rule != "center_distance"
            or self.boundary_rule != "first_fault_loses"
        ):
            raise ValueError("Finite-body game rules require point_vehicles=False")

    def to_dict(self) -> dict:
        return asdict(self)

    @classmethod
    def from_dict(cls, value: dict) -> "ReversibleConfig":
        data = dict(value)
        for role in ("pursuer", "evader"):
            if role in data:
                data[role] = VehicleLimits(**data[role])
        return cls(**data)


def wrap_angle(value):
    """Angle wrapped into [-pi, pi) (reversible/model.py:82-83)."""
    return (np.asarray(value) + np.pi) % (2 * np.pi) - np.pi


def _acceleration(v, a, limits):
    # Acceleration vanishes when it would push past a speed bound
    # (reversible/model.py:186-191).
    out = np.where(
        ((v <= limits.speed_min) & (a < 0)) | ((v >= limits.speed_max) & (a > 0)),
        0.0,
        a,
    )
    return out


def world_rhs(states, p_controls, e_controls, config):
    """Dynamics of the joint state [xP, yP, thP, vP, xE, yE, thE, vE] under
    commands [acceleration, curvature] (reversible/model.py:194-206)."""
    s = np.asarray(states, float)
    p = np.asarray(p_controls, float)
    e = np.asarray(e_controls, float)
    parts = []
    for base, u, limits in ((0, p, config.pursuer), (4, e, config.evader)):
        heading, v = s[..., base + 2], s[..., base + 3]
        parts += [v * np.cos(heading), v * np.sin(heading), v * u[..., 1],
                  _acceleration(v, u[..., 0], limits)]
    return np.stack(np.broadcast_arrays(*parts), axis=-1)


def relative_rhs(states, p_controls, e_controls, config):
    """Relative dynamics of [x, y, psi, vP, vE] (reversible/model.py:209-220)."""
    r = np.asarray(states, float)
    p = np.asarray(p_controls, float)
    e = np.asarray(e_controls, float)
    x, y, psi, vp, ve = (r[..., k] for k in range(5))
    wp = vp * p[..., 1]
    we = ve * e[..., 1]
    parts = (
        ve * np.cos(psi) - vp + wp * y,
        ve * np.sin(psi) - wp * x,
        we - wp,
        _acceleration(vp, p[..., 0], config.pursuer),
        _acceleration(ve, e[..., 0], config.evader),
    )
    return np.stack(np.broadcast_arrays(*parts), axis=-1)
