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
