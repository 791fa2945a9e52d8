"""Budget-indexed field helpers shared by the field readers."""

from __future__ import annotations

import numpy as np


def first_crossing(times, values, valid=None):
    """First budget at which each query's value reaches zero, linearly
    interpolated between the bracketing saved budgets; NaN where it never
    does (or the query is invalid).  values: [budgets, queries]
    (reversible/map_game_field.py:83-98)."""
    times = np.asarray(times, float)
    values = np.asarray(values, float)
    if values.ndim != 2 or len(times) != len(values) or np.any(np.diff(times) <= 0):
        raise ValueError("Expected increasing budgets and fields[budget,query]")
    hit = np.isfinite(values) & (values <= 0)
    found = hit.any(axis=0)
    j = hit.argmax(axis=0)
    prev = np.maximum(j - 1, 0)
    cols = np.arange(values.shape[1])
    a, b = values[prev, cols], values[j, cols]
    gap = a - b
    frac = np.zeros_like(a)
    np.divide(a, gap, out=frac, where=np.abs(gap) > 1e-12)
    out = times[prev] + np.clip(frac, 0.0, 1.0) * (times[j] - times[prev])
    if valid is not None:
        found &= np.asarray(valid, bool)
    return np.where(found, out, np.nan)
