"""The 5D table reader without a GPU: the oracle (oracle/rtable_oracle.py)
against the reader's known answers, and
the constructor checks (reversible/relative_table.py:14-74)."""

import json
import math

import numpy as np
import pytest

import rtable_oracle as ro
from rtable_fixtures import synthetic_table


def oracle_parts(table):
    fields = np.asarray(table.fields).reshape(len(table.times), -1)
    return fields, list(table.times), table.axes, table.physics


def test_oracle_known_answers(tmp_path):
    """(tests/test_reversible_free_dp.py:159-177) on the oracle: on a field
    linear in x, y, vP, vE and decreasing at unit rate, the first crossing is
    x + 0.2 y + 0.3 vP - 0.4 vE exactly; psi is periodic; outside the grid,
    NaN; inside the capture radius, 0."""
    table, _ = synthetic_table(tmp_path)
    fields, times, axes, physics = oracle_parts(table)
    states = [[1, 0.7, 0.4, -0.5, 0.4], [0.8, -0.4, 3.2, 0.5, 0.1]]
    expected = [s[0] + 0.2 * s[1] + 0.3 * s[3] - 0.4 * s[4] for s in states]
    np.testing.assert_allclose(ro.value(fields, times, axes, physics, states), expected,
                               atol=2e-7)
    shifted = [[s[0], s[1], s[2] + 2 * math.pi, s[3], s[4]] for s in states]
    np.testing.assert_allclose(ro.value(fields, times, axes, physics, shifted), expected,
                               atol=2e-7)
    for bad in ([2.1, 0, 0, 0, 0], [1, 0, 0, 2.1, 0], [math.nan, 0, 0, 0, 0]):
        assert math.isnan(ro.value(fields, times, axes, physics, [bad])[0])
    assert all(math.isnan(v) for v in ro.field(fields, times, axes, states, [-0.01] * 2))
    assert all(math.isnan(v) for v in ro.field(fields, times, axes, states, [2.01] * 2))
    assert ro.value(fields, times, axes, physics, [[0.1, 0.1, 0, -1, 1]]) == [0.0]
    # field at a budget between frames: linear in time, exact on this field
    s = [0.3, -0.2, 1.0, 0.25, -0.1]
    f = ro.field(fields, times, axes, [s], [0.75])[0]
    assert f == pytest.approx(0.3 + 0.2 * -0.2 + 0.3 * 0.25 - 0.4 * -0.1 - 0.75, abs=2e-7)


def test_oracle_stencil_is_multilinear_exact():
    """Brute force: the stencil reproduces any function multilinear in the
    non-periodic coordinates, at random states."""
    from pe_amd.reversible.free_dp import RelativeGrid
    from pe_amd.reversible.model import ReversibleConfig

    axes = RelativeGrid(9, 8, 7, 7, 2).axes(ReversibleConfig())
    vp, ve, psi, y, x = np.meshgrid(*axes, indexing="ij")
    fn = lambda x, y, vp, ve: (x + 1) * (2 - y) * (vp + 3) - 0.5 * ve * x
    grid_values = fn(x, y, vp, ve).reshape(-1)
    rng = np.random.default_rng(3)
    for _ in range(50):
        s = [rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(-7, 7), rng.uniform(-1, 2),
             rng.uniform(-0.5, 1)]
        ii, ww, valid = ro.stencil_one(s, axes)
        assert valid and sum(ww) == pytest.approx(1, abs=1e-14)
        got = sum(grid_values[i] * w for i, w in zip(ii, ww))
        assert got == pytest.approx(fn(s[0], s[1], s[3], s[4]), abs=1e-12)


def test_constructor_checks(tmp_path):
    from pe_amd.reversible.relative_table import RelativeTable

    table, c = synthetic_table(tmp_path)
    folder = table.folder
    meta = json.loads((folder / "meta.json").read_text())
    short = RelativeTable(folder, c, max_budget=1.0)
    assert short.horizon == 1 and len(short.times) == 3
    np.testing.assert_array_equal(short.fields, table.fields[:3])
    with pytest.raises(ValueError):
        RelativeTable(folder, c, max_budget=0.7)
    with pytest.raises(ValueError):
        RelativeTable(folder, c, max_budget=0.0)
    for key, bad in (("schema", "x"), ("status", "running"), ("history_saved", False),
                     ("tube_projection", "other")):
        (folder / "meta.json").write_text(json.dumps(dict(meta, **{key: bad})))
        with pytest.raises(ValueError):
            RelativeTable(folder, c)
    (folder / "meta.json").write_text(json.dumps(dict(meta, times=meta["times"][:-1])))
    with pytest.raises(ValueError, match="byte length"):
        RelativeTable(folder, c)
    (folder / "meta.json").write_text(json.dumps(meta))
    from dataclasses import replace

    with pytest.raises(ValueError):
        RelativeTable(folder, replace(c, capture_radius=0.4))
    with pytest.raises(ValueError):
        table.stencil(np.zeros((2, 4)))


def test_vectorised_oracle_equals_loop_oracle(tmp_path):
    table, _ = synthetic_table(tmp_path)
    fields, times, axes, _ = oracle_parts(table)
    rng = np.random.default_rng(4)
    states = np.column_stack([rng.uniform(-2.2, 2.2, 200), rng.uniform(-2.2, 2.2, 200),
                              rng.uniform(-9, 9, 200), rng.uniform(-1.1, 2.1, 200),
                              rng.uniform(-0.6, 1.1, 200)])
    budgets = rng.uniform(-0.1, 2.1, 200)
    got = ro.field_np(fields, times, axes, states, budgets)
    want = ro.field(fields, times, axes, states.tolist(), budgets.tolist())
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0, equal_nan=True)
