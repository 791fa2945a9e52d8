"""Independent full-array NumPy (float64) statement of one update of the
dense 8D square game, for tiny grids: the check the reference applies to its
own executable (tests/test_reversible_world_dp.py:42-110).

It is formulated differently from the oracle and the kernel: shifted copies
of the whole array per axis and offset (np.roll on the periodic headings,
padding by signed linear extrapolation elsewhere) instead of per-node
stencil walks, float64 instead of float32.
"""

import numpy as np

from pe_amd.reversible.world_dp import hamiltonian, surfaces


def _shifted(value, axis, offset, periodic):
    """value[..., i + offset, ...] along `axis`, continued past the ends."""
    if periodic:
        return np.roll(value, -offset, axis=axis)
    n = value.shape[axis]
    idx = np.arange(n) + offset
    out = np.take(value, np.clip(idx, 0, n - 1), axis=axis)
    first, second = np.take(value, 0, axis=axis), np.take(value, 1, axis=axis)
    last, before = np.take(value, n - 1, axis=axis), np.take(value, n - 2, axis=axis)
    for i in np.flatnonzero((idx < 0) | (idx >= n)):
        if idx[i] < 0:
            edge, inner, dist = first, second, -idx[i]
        else:
            edge, inner, dist = last, before, idx[i] - (n - 1)
        sl = [slice(None)] * value.ndim
        sl[axis] = i
        out[tuple(sl)] = edge + dist * np.abs(inner - edge) * np.sign(edge)
    return out


def _at(array, axis, index):
    sl = [slice(None)] * array.ndim
    sl[axis] = index
    return tuple(sl)


def rhs(value, grid, config, order, speed_boundary="linear"):
    """H(x, grad) + Lax-Friedrichs dissipation at every node."""
    axes = grid.axes(config)
    mesh = np.meshgrid(*axes, indexing="ij", sparse=True)
    states = np.stack(np.broadcast_arrays(*mesh), axis=-1)
    p, e = config.pursuer, config.evader
    alpha = [np.abs(states[..., 3] * np.cos(states[..., 2])),
             np.abs(states[..., 3] * np.sin(states[..., 2])),
             np.abs(states[..., 3]) * p.max_curvature, p.acceleration,
             np.abs(states[..., 7] * np.cos(states[..., 6])),
             np.abs(states[..., 7] * np.sin(states[..., 6])),
             np.abs(states[..., 7]) * e.max_curvature, e.acceleration]
    grads, diss = [], 0.0
    for k, h in enumerate(grid.spacings(config)):
        periodic = k in (2, 6)
        linear = speed_boundary == "linear" and k in (3, 7)
        m1, p1 = _shifted(value, k, -1, periodic), _shifted(value, k, 1, periodic)
        left, right = (value - m1) / h, (p1 - value) / h
        if order == 2:
            dc = (p1 - 2 * value + m1) / h ** 2
            dl = (value - 2 * m1 + _shifted(value, k, -2, periodic)) / h ** 2
            dr = (_shifted(value, k, 2, periodic) - 2 * p1 + value) / h ** 2
            if linear:
                dl[_at(dl, k, 1)] = dc[_at(dc, k, 1)]
                dr[_at(dr, k, -2)] = dc[_at(dc, k, -2)]
            left = left + 0.5 * h * np.where(np.abs(dl) <= np.abs(dc), dl, dc)
            right = right - 0.5 * h * np.where(np.abs(dc) <= np.abs(dr), dc, dr)
        if linear:
            lo = (np.take(value, 1, axis=k) - np.take(value, 0, axis=k)) / h
            hi = (np.take(value, -1, axis=k) - np.take(value, -2, axis=k)) / h
            for arr in (left, right):
                arr[_at(arr, k, 0)] = lo
                arr[_at(arr, k, -1)] = hi
        grads.append(0.5 * (left + right))
        diss = diss + 0.5 * alpha[k] * (right - left)
    return hamiltonian(states, np.stack(grads, axis=-1), config) + diss


def one_step(grid, config, dt, order, speed_boundary="linear"):
    """(initial, avoid, field after one RK step with the tube projection)."""
    target, avoid = surfaces(config, grid, dtype=np.float64)
    initial = np.maximum(target, avoid)
    stage = initial + dt * rhs(initial, grid, config, order, speed_boundary)
    if order == 2:
        stage = 0.5 * (initial + stage + dt * rhs(stage, grid, config, order, speed_boundary))
    return initial, avoid, np.maximum(np.minimum(stage, initial), avoid)
