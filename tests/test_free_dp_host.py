"""Host logic of the free_dp mirror: grid layout, validation, guards and the
exceptions of reference reversible/free_dp.py:23-176 (no GPU needed)."""

import json
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from pe_amd.reversible.free_dp import (
    RelativeGrid,
    output_times,
    physics_signature,
    solve,
)
from pe_amd.reversible.model import ReversibleConfig, VehicleLimits

ROOT = Path(__file__).resolve().parents[1]


def a3():
    data = {
        "name": "reversible_free_a3",
        "pursuer": {"speed_min": -1.0, "speed_max": 2.0, "acceleration": 3.0, "max_curvature": 1.0},
        "evader": {"speed_min": -0.5, "speed_max": 1.0, "acceleration": 3.0, "max_curvature": 2.0},
        "capture_radius": 0.3,
        "control_dt": 0.05,
        "arena_half_width": 3.0,
        "boundary_rule": "first_fault_loses",
        "point_vehicles": True,
    }
    return ReversibleConfig.from_dict(json.loads(json.dumps(data)))


def test_grid_layout_and_axes():
    g = RelativeGrid(17, 12, 7, 7, 4.0)
    assert g.shape == (7, 7, 12, 17, 17)
    assert g.states == 169_932
    vp, ve, psi, y, x = g.axes(a3())
    assert vp[0] == -1 and vp[-1] == 2 and ve[0] == -0.5 and ve[-1] == 1
    assert psi[0] == -np.pi and len(psi) == 12 and psi[-1] < np.pi
    np.testing.assert_allclose(np.diff(psi), 2 * np.pi / 12)
    assert x[0] == -4 and x[-1] == 4
    c1 = RelativeGrid(65, 36, 15, 15, 8.0)
    assert c1.states == 34_222_500


@pytest.mark.parametrize(
    "kwargs", [dict(positions=2), dict(headings=3), dict(p_speeds=2.5), dict(e_speeds=True)]
)
def test_grid_rejects_bad_sizes(kwargs):
    with pytest.raises(ValueError):
        RelativeGrid(**kwargs)
    with pytest.raises(ValueError):
        RelativeGrid(extent=float("inf"))


def test_output_times_and_signature():
    np.testing.assert_allclose(output_times(0.2, 0.1), [0, 0.1, 0.2])
    np.testing.assert_allclose(output_times(0.25, 0.1), [0, 0.1, 0.2, 0.25])
    assert len(output_times(3.0, 0.1)) == 31
    sig = physics_signature(a3())
    assert sig["capture_radius"] == 0.3 and "capture_rule" not in sig
    assert sig["pursuer"]["acceleration"] == 3.0


def test_config_contract():
    with pytest.raises(ValueError):
        VehicleLimits(0.0, 1.0, 1.0, 1.0)  # must straddle zero
    with pytest.raises(ValueError):
        ReversibleConfig(capture_rule="rectangle_contact")  # point vehicles
    c = replace(a3(), point_vehicles=False, capture_rule="rectangle_contact",
                boundary_rule="collision_constrained")
    assert ReversibleConfig.from_dict(c.to_dict()) == c


def test_budget_preflight_does_not_allocate(tmp_path):
    with pytest.raises(MemoryError):
        solve(ReversibleConfig(), RelativeGrid(101, 48, 13, 13, 4), tmp_path / "large",
              max_working_gb=0.001)
    assert not (tmp_path / "large").exists()


def test_refuses_overwrite(tmp_path):
    out = tmp_path / "field"
    out.mkdir()
    (out / "x").write_text("keep")
    with pytest.raises(FileExistsError):
        solve(ReversibleConfig(), RelativeGrid(9, 8, 7, 7, 1.0), out, horizon=0.1)


@pytest.mark.parametrize(
    "kwargs",
    [dict(order=3), dict(cfl=1.5), dict(horizon=-1.0), dict(threads=0),
     dict(output_dt=0.0), dict(local_precision="float16")],
)
def test_invalid_options(tmp_path, kwargs):
    with pytest.raises(ValueError):
        solve(ReversibleConfig(), RelativeGrid(9, 8, 7, 7, 1.0), tmp_path / "o", **kwargs)


def test_capture_rule_options(tmp_path):
    from pe_amd.local_game.geometry import Footprint

    rect = replace(ReversibleConfig(), point_vehicles=False, capture_rule="rectangle_contact",
                   boundary_rule="collision_constrained")
    with pytest.raises(ValueError):
        solve(rect, RelativeGrid(9, 8, 7, 7, 1.0), tmp_path / "a")
    with pytest.raises(ValueError):
        solve(ReversibleConfig(), RelativeGrid(9, 8, 7, 7, 1.0), tmp_path / "b",
              capture_footprints=(Footprint(0.5, 0.3), Footprint(0.5, 0.3)))
    with pytest.raises(ValueError):
        solve(ReversibleConfig(), RelativeGrid(9, 8, 7, 7, 1.0), tmp_path / "c", exit_radius=2.0)
