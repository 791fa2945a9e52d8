"""Host logic of the field readers (no GPU): first_crossing against hand-
computed crossings (reversible/map_game_field.py:83-98)."""

import numpy as np
import pytest

from pe_amd.reversible.map_game_field import first_crossing


def test_first_crossing_interpolates_between_budgets():
    times = [0.0, 1.0, 2.0]
    values = np.array([[1.0, 1.0, -1.0, 2.0, 1.0],
                       [0.5, -1.0, -1.0, 1.0, np.nan],
                       [-0.5, -2.0, -1.0, 0.5, -1.0]])
    got = first_crossing(times, values)
    # col 0: 0.5 -> -0.5 between t=1 and 2 -> 1.5; col 1: 1 -> -1 -> 0.5;
    # col 2: already captured at t=0; col 3: never; col 4: the frame before
    # the crossing is NaN, so no interpolation: the earlier budget, t=1
    assert got[0] == pytest.approx(1.5)
    assert got[1] == pytest.approx(0.5)
    assert got[2] == 0.0
    assert np.isnan(got[3])
    assert got[4] == 1.0


def test_first_crossing_valid_mask_and_shape_checks():
    times = [0.0, 0.5]
    values = np.array([[1.0, 1.0], [-1.0, -1.0]])
    got = first_crossing(times, values, valid=[True, False])
    assert got[0] == pytest.approx(0.25) and np.isnan(got[1])
    with pytest.raises(ValueError):
        first_crossing([0.0, 0.0], values)
    with pytest.raises(ValueError):
        first_crossing([0.0, 1.0, 2.0], values)
