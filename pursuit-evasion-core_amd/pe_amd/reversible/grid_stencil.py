"""Multilinear stencils on the uniform grids of the reversible-car games,
formed on the GPU (pe_grid_stencil).

`interpolation_stencil(states, axes, periodic_axes)` keeps the signature and
the answers of the reference's helper (reversible/world_table.py:13-45): node
indices [m, 2^d], weights [m, 2^d], validity [m] and the batch shape, with
the corner order of itertools.product([0, 1], repeat=d), periodic axes
wrapping with period (axis[1] - axis[0]) * n and the other axes clipped,
states more than 1e-10 outside them invalid.
"""

from __future__ import annotations

import numpy as np

from .. import _native


def axis_spec(axes, periodic_axes, period=None):
    """(n, lo, hi, step, period, mask) of uniformly spaced axes for
    pe_grid_stencil; the period of a periodic axis defaults to step * n."""
    n = [len(a) for a in axes]
    lo = [float(a[0]) for a in axes]
    hi = [float(a[-1]) for a in axes]
    step = [float(a[1] - a[0]) for a in axes]
    per = [0.0] * len(axes)
    mask = 0
    for k in periodic_axes:
        mask |= 1 << k
        per[k] = step[k] * n[k] if period is None else float(period[k])
    return n, lo, hi, step, per, mask


def interpolation_stencil(states, axes, periodic_axes):
    s = np.asarray(states, dtype=float)
    d = len(axes)
    if s.shape[-1] != d:
        raise ValueError(f"Expected final state dimension {d}.")
    batch = s.shape[:-1]
    n, lo, hi, step, per, mask = axis_spec(axes, periodic_axes)
    index, weight, valid = _native.grid_stencil(s.reshape(-1, d), n, lo, hi, step, per, mask,
                                                1e-10)
    return index, weight, valid, batch
