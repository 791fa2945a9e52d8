"""The dense 8D world-game oracle (oracle/world8d.c) against golden vectors
written by the reference executable itself (reversible/native_world.cpp,
compiled unmodified; tests/golden/make_golden.py): final field, arrival
times, every saved frame and the CFL limit, bit for bit, for RK2 with the
linear speed boundary, RK1 with the legacy signed one, and an even position
count.  Plus invariants the scheme fixes: the completed update never rises
(tube projection) and never falls below the pursuer wall-fault avoid
surface.
"""

import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["world8d_5x4x3x3_order2_linear", "world8d_5x4x3x3_order1_legacy",
         "world8d_6x6x4x3_order2_linear"]


def cfg_from(oracle_lib, case, threads=4):
    return oracle_lib.World8DCfg(
        case["positions"], case["headings"], case["p_speeds"], case["e_speeds"], case["order"],
        case["linear_speed_boundary"], threads, case["extent"], case["arena"], case["radius"],
        case["p_min"], case["p_max"], case["e_min"], case["e_max"], case["p_accel"],
        case["e_accel"], case["p_curvature"], case["e_curvature"], case["cfl"])


def bits(a):
    return np.nan_to_num(np.asarray(a, np.float32)).view(np.uint32)


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_world_binary(oracle_lib, name):
    z = np.load(GOLDEN / f"{name}.npz")
    case = json.loads(str(z["case"]))
    cfg = cfg_from(oracle_lib, case)
    frames, final, arrival, steps, maxdt = oracle_lib.world8d_solve(
        cfg, case["horizon"], case["output_dt"], history=bool(case["history"]))
    assert steps == int(z["internal_steps"])
    assert maxdt == pytest.approx(float(z["max_internal_dt"]), rel=1e-11)
    np.testing.assert_array_equal(bits(final), bits(z["final"]))
    np.testing.assert_array_equal(np.isnan(arrival), np.isnan(z["arrival"]))
    np.testing.assert_array_equal(bits(arrival), bits(z["arrival"]))
    if case["history"]:
        assert frames.shape == z["fields"].shape
        np.testing.assert_array_equal(bits(frames), bits(z["fields"]))


def test_world_invariants(oracle_lib):
    z = np.load(GOLDEN / f"{CASES[0]}.npz")
    case = json.loads(str(z["case"]))
    cfg = cfg_from(oracle_lib, case)
    frames, final, arrival, _, _ = oracle_lib.world8d_solve(cfg, 0.4, 0.1)
    n, h = case["positions"], case["headings"]
    shape = (n, n, h, case["p_speeds"], n, n, h, case["e_speeds"])
    assert (np.diff(frames, axis=0) <= 0).all()
    x = np.linspace(-case["extent"], case["extent"], n).astype(np.float32)
    avoid = (np.maximum(abs(x)[:, None], abs(x)[None, :]) - np.float32(case["arena"])
             + np.float32(1e-6))
    f = final.reshape(shape)
    assert (f >= avoid[:, :, None, None, None, None, None, None] - 1e-7).all()
    # captured nodes have an arrival time no later than the horizon
    captured = final <= 0
    assert np.isfinite(arrival[captured]).all() and (arrival[captured] <= 0.4 + 1e-6).all()


def grid_of(case):
    from pe_amd.reversible.world_dp import WorldGrid

    return WorldGrid(case["positions"], case["headings"], case["p_speeds"], case["e_speeds"],
                     case["extent"])


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("speed_boundary", ["linear", "legacy_signed"])
def test_oracle_step_matches_independent_numpy(oracle_lib, order, speed_boundary):
    """One RK step of the oracle against the full-array float64 statement
    (tests/test_reversible_world_dp.py:113-151 applies it to the reference
    executable with the same 2e-6 tolerance)."""
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import WorldGrid, global_rate_bound
    from world8d_numpy import one_step

    config, grid, dt = ReversibleConfig(), WorldGrid(5, 4, 3, 3, 3.5), 0.01
    assert dt < 0.8 / global_rate_bound(config, grid)
    case = dict(positions=5, headings=4, p_speeds=3, e_speeds=3, order=order,
                linear_speed_boundary=int(speed_boundary == "linear"), extent=3.5, arena=3.0,
                radius=0.3, p_min=-1.0, p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0,
                e_accel=1.0, p_curvature=1.0, e_curvature=2.0, cfl=0.8)
    _, final, _, steps, _ = oracle_lib.world8d_solve(cfg_from(oracle_lib, case), dt, dt,
                                                    history=False)
    initial, avoid, expected = one_step(grid, config, dt, order, speed_boundary)
    assert steps == 1
    np.testing.assert_allclose(final.reshape(grid.shape), expected, atol=2e-6, rtol=2e-6)
    assert np.max(final.reshape(grid.shape) - initial) < 2e-6
    assert np.min(final.reshape(grid.shape) - avoid) > -2e-6


def test_oracle_speed_limit_generator(oracle_lib):
    """phi = 10 - xP - vP at a node at the pursuer's top speed, heading 0,
    RK1: the pursuer's best hold (no acceleration) gives dphi/dt = -vP = -2
    exactly (tests/test_reversible_world_dp.py:204-226 computes the same -2
    from exact held-control steps); the inward one-sided speed slope
    reproduces it, the legacy signed extrapolation gives 0 there."""
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import WorldGrid

    config, grid = ReversibleConfig(), WorldGrid(5, 4, 4, 4)
    axes = grid.axes(config)
    mesh = np.broadcast_arrays(*np.meshgrid(*axes, indexing="ij", sparse=True))
    phi = (10 - mesh[0] - mesh[3]).astype(np.float32)
    node = (2, 2, 2, 3, 2, 2, 2, 3)
    dt = 0.005
    for linear, expected in ((1, -2.0), (0, 0.0)):
        case = dict(positions=5, headings=4, p_speeds=4, e_speeds=4, order=1,
                    linear_speed_boundary=linear, extent=3.5, arena=3.0, radius=0.3,
                    p_min=-1.0, p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0, e_accel=1.0,
                    p_curvature=1.0, e_curvature=2.0, cfl=0.8)
        arrival = np.full(phi.size, np.nan, np.float32)
        _, final, _, steps, _ = oracle_lib.world8d_solve(
            cfg_from(oracle_lib, case), dt, dt, field=phi.reshape(-1), arrival=arrival,
            history=False)
        assert steps == 1
        rate = (float(final.reshape(grid.shape)[node]) - float(phi[node])) / dt
        assert rate == pytest.approx(expected, abs=2e-3)


def test_probe_goldens_are_sequential_float64_sums():
    """probe_fields.f64 of the executable = sum over the 256 corners, in
    order, of double(field[index]) * weight, at every saved frame; pins the
    summation the GPU probe kernel restates (native_world.cpp:241-248)."""
    z = np.load(GOLDEN / f"{CASES[0]}.npz")
    frames, ii, ww = z["fields"], z["probe_index"].astype(np.int64), z["probe_weight"]
    assert z["probe_fields"].shape == (len(frames), len(ii))
    for t, frame in enumerate(frames):
        acc = np.zeros(len(ii))
        for k in range(256):
            acc = acc + frame[ii[:, k]].astype(np.float64) * ww[:, k]
        np.testing.assert_array_equal(acc, z["probe_fields"][t])


def test_surfaces_match_oracle_initial_field(oracle_lib):
    from pe_amd.reversible.model import ReversibleConfig

    z = np.load(GOLDEN / f"{CASES[2]}.npz")
    case = json.loads(str(z["case"]))
    grid = grid_of(case)
    from pe_amd.reversible.world_dp import surfaces

    # surfaces() takes its axes from float64 linspace, the executable from
    # float steps -extent + i*dx: they differ by an ulp on some nodes.
    target, avoid = surfaces(ReversibleConfig(), grid)
    field, arrival = oracle_lib.world8d_init(cfg_from(oracle_lib, case))
    np.testing.assert_allclose(np.maximum(target, avoid).reshape(-1), field, atol=1e-6, rtol=0)
    np.testing.assert_array_equal(np.isnan(arrival), field > 0)


def test_cfl_limit_is_cfl_over_global_rate_bound(oracle_lib):
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import global_rate_bound

    for name in CASES:
        z = np.load(GOLDEN / f"{name}.npz")
        case = json.loads(str(z["case"]))
        bound = global_rate_bound(ReversibleConfig(), grid_of(case))
        assert float(z["max_internal_dt"]) == pytest.approx(case["cfl"] / bound, rel=1e-6)
