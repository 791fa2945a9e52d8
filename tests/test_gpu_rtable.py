"""The 5D table reader's GPU batch queries (pe_rtable through the C ABI):
the reference's reader tests (tests/test_reversible_free_dp.py:159-215) and
the oracle on a solver-written history, to 1e-12."""

import math

import numpy as np
import pytest

import rtable_oracle as ro
from rtable_fixtures import synthetic_table

pytestmark = pytest.mark.gpu


def test_signed_history_query_order_periodicity_and_unknown(gpu, tmp_path):
    print("[rtable] synthetic table", flush=True)
    table, c = synthetic_table(tmp_path)
    states = np.array([[1, 0.7, 0.4, -0.5, 0.4], [0.8, -0.4, 3.2, 0.5, 0.1]])
    expected = states[:, 0] + 0.2 * states[:, 1] + 0.3 * states[:, 3] - 0.4 * states[:, 4]
    np.testing.assert_allclose(table.value(states), expected, atol=2e-7)
    shifted = states.copy()
    shifted[:, 2] += 2 * np.pi
    np.testing.assert_allclose(table.value(shifted), expected, atol=2e-7)
    assert np.isnan(table.value([2.1, 0, 0, 0, 0]))
    assert np.isnan(table.value([1, 0, 0, 2.1, 0]))
    assert np.isnan(table.value([np.nan, 0, 0, 0, 0]))
    assert np.isnan(table.field(states, -0.01)).all()
    assert np.isnan(table.field(states, 2.01)).all()
    assert table.value([0.1, 0.1, 0, -1, 1]) == 0
    from dataclasses import replace

    table.assert_compatible(replace(c, arena_half_width=10))
    with pytest.raises(ValueError):
        table.assert_compatible(replace(c, capture_radius=0.4))


def test_hamiltonian_controls_flip_curvature_with_signed_speed(gpu, tmp_path):
    table, _ = synthetic_table(tmp_path)
    s = np.array([[1, 0.7, 0.4, -0.5, 0.4], [1, 0.7, 0.4, 0.5, 0.4], [1, 0.7, 0.4, -1, -0.5]])
    p, e = table.hamiltonian_controls(s, 1)
    np.testing.assert_allclose(p[:2, 1], [1, -1])
    np.testing.assert_allclose(p[:2, 0], [-2, -2])
    assert p[2, 0] == 0 and e[2, 0] == 0
    np.testing.assert_allclose(table.gradient(s, 1),
                               np.broadcast_to([1, 0.2, 0, 0.3, -0.4], (3, 5)), atol=1e-6)


def test_budget_view_limits_queries_without_copying_history(gpu, tmp_path):
    from pe_amd.reversible.relative_table import RelativeTable

    full, c = synthetic_table(tmp_path)
    short = RelativeTable(full.folder, c, max_budget=1.0)
    state = [1.2, 0.3, 0, 1, 0.2]
    assert np.isfinite(full.value(state)) and np.isnan(short.value(state))
    np.testing.assert_array_equal(short.fields, full.fields[:3])
    assert np.isnan(short.field(state, 1.1))
    assert full.horizon == 2 and short.horizon == 1


def test_gpu_queries_match_oracle_on_a_solved_history(gpu, tmp_path):
    """A free_dp solve with history (RelativeGrid(9, 8, 7, 7, 1.0), horizon
    0.3, 7 frames); 3000 seeded states over and past the grid (NaN and
    periodic headings included) at seeded budgets."""
    from pe_amd.reversible.free_dp import RelativeGrid, solve
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.relative_table import RelativeTable

    print("[rtable] solved history vs oracle", flush=True)
    c = ReversibleConfig()
    solve(c, RelativeGrid(9, 8, 7, 7, 1.0), tmp_path / "field", horizon=0.3, output_dt=0.05,
          threads=2)
    table = RelativeTable(tmp_path / "field", c)
    rng = np.random.default_rng(21)
    n = 3000
    states = np.column_stack([rng.uniform(-1.1, 1.1, n), rng.uniform(-1.1, 1.1, n),
                              rng.uniform(-10, 10, n), rng.uniform(-1.05, 2.05, n),
                              rng.uniform(-0.55, 1.05, n)])
    states[:5, 0] = [1.0, -1.0, 0.0, 0.25, 0.5]
    states[5, 1] = np.nan
    budgets = rng.uniform(-0.02, 0.32, n)
    budgets[:3] = [0.0, 0.3, 0.05]
    fields = np.asarray(table.fields).reshape(len(table.times), -1)
    times = list(table.times)
    got = table.field(states, budgets)
    want = ro.field(fields, times, table.axes, states.tolist(), budgets.tolist())
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0, equal_nan=True)
    got_v = table.value(states)
    want_v = ro.value(fields, times, table.axes, table.physics, states.tolist())
    np.testing.assert_array_equal(np.isnan(got_v), np.isnan(want_v))
    np.testing.assert_allclose(got_v, want_v, atol=1e-12, rtol=0, equal_nan=True)
    print(f"[rtable] resolved {np.isfinite(got_v).sum()} of {n}", flush=True)
    assert np.isfinite(got_v).sum() > 100 and np.isnan(got_v).sum() > 100


def test_gpu_stencil_equals_oracle(gpu, tmp_path):
    """RelativeTable.stencil (pe_grid_stencil) against the oracle's
    per-query stencil, bit for bit."""
    table, _ = synthetic_table(tmp_path)
    rng = np.random.default_rng(9)
    states = np.column_stack([rng.uniform(-2.2, 2.2, 300), rng.uniform(-2.2, 2.2, 300),
                              rng.uniform(-9, 9, 300), rng.uniform(-1.1, 2.1, 300),
                              rng.uniform(-0.6, 1.1, 300)])
    states[0, 2] = np.pi
    states[1, 0] = 2.0
    states[2, 3] = np.nan
    ii, ww, valid, shape = table.stencil(states)
    assert shape == (300,)
    for k, s in enumerate(states):
        oi, ow, ov = ro.stencil_one(list(s), table.axes)
        assert valid[k] == ov
        np.testing.assert_array_equal(ii[k], oi)
        np.testing.assert_array_equal(ww[k], ow)
