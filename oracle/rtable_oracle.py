"""ORACLE — test infrastructure only.

Pure-Python restatement, one query at a time, of the 5D table reader
reversible/relative_table.py:76-161 (stencil, field, value) for small
batches.  Only tests/ and bench.py's cpu_baseline leg use it, as the checker.

Pinned by the reader's own known answers (tests/test_reversible_free_dp.py:
132-215: a field linear in x, y, vP, vE and time, whose interpolation and
first crossing are exact; heading periodicity; NaN outside the grid or the
budget range; 0 inside the capture radius; the max_budget view) and by
brute-force multilinear exactness.
"""

from __future__ import annotations

import math


def stencil_one(state, axes):
    """(indices[32], weights[32], valid) of one state [x, y, psi, vP, vE];
    axes in storage order [vP, vE, psi, y, x] (relative_table.py:82-109)."""
    q = [state[3], state[4], state[2], state[1], state[0]]
    valid = all(math.isfinite(v) for v in q)
    q = [v if math.isfinite(v) else 0.0 for v in q]
    shape = [len(a) for a in axes]
    strides = [1] * 5
    for k in range(3, -1, -1):
        strides[k] = strides[k + 1] * shape[k + 1]
    low, frac = [], []
    for k, axis in enumerate(axes):
        step = float(axis[1] - axis[0])
        n = shape[k]
        if k == 2:
            f = ((q[k] + math.pi) % (2 * math.pi)) / step
            j = int(math.floor(f)) % n
            w = f - math.floor(f)
        else:
            lo, hi = float(axis[0]), float(axis[-1])
            valid = valid and lo - 1e-10 <= q[k] <= hi + 1e-10
            f = (min(max(q[k], lo), hi) - lo) / step
            j = min(int(math.floor(f)), n - 2)
            w = f - j
        low.append(j)
        frac.append(w)
    indices, weights = [], []
    for c in range(32):
        idx, w = 0, 1.0
        for k in range(5):
            up = (c >> (4 - k)) & 1
            node = low[k] + up
            if k == 2:
                node %= shape[k]
            idx += node * strides[k]
            w *= frac[k] if up else 1 - frac[k]
        indices.append(idx)
        weights.append(w)
    return indices, weights, valid


def _sample(frame, indices, weights):
    return sum(float(frame[i]) * w for i, w in zip(indices, weights))


def field(fields, times, axes, states, budgets):
    """fields [frames, N] (flat); relative_table.py:111-131."""
    out = []
    horizon = times[-1]
    for s, tau in zip(states, budgets):
        ii, ww, valid = stencil_one(s, axes)
        if not (valid and math.isfinite(tau) and 0 <= tau <= horizon + 1e-10):
            out.append(math.nan)
            continue
        k = sum(1 for t in times if t <= tau) - 1
        k = min(max(k, 0), len(times) - 2)
        fraction = (tau - times[k]) / (times[k + 1] - times[k])
        a = _sample(fields[k], ii, ww)
        b = _sample(fields[k + 1], ii, ww)
        out.append(a + fraction * (b - a))
    return out


def value(fields, times, axes, physics, states):
    """First crossing (relative_table.py:133-161) with a plain scan for the
    first frame at or below zero (equal to the reader's bisection on
    histories that are non-increasing in time)."""
    out = []
    p, e = physics["pursuer"], physics["evader"]
    for s in states:
        physical = all(math.isfinite(v) for v in s) and (
            p["speed_min"] - 1e-10 <= s[3] <= p["speed_max"] + 1e-10) and (
            e["speed_min"] - 1e-10 <= s[4] <= e["speed_max"] + 1e-10)
        if physical and math.sqrt(s[0] * s[0] + s[1] * s[1]) <= physics["capture_radius"]:
            out.append(0.0)
            continue
        ii, ww, valid = stencil_one(s, axes)
        samples = [_sample(f, ii, ww) for f in fields]
        if not valid or not samples[-1] <= 0:
            out.append(math.nan)
            continue
        hi = next(t for t, v in enumerate(samples) if v <= 0)
        lo = max(hi - 1, 0)
        a, b = samples[lo], samples[hi]
        fraction = a / (a - b) if abs(a - b) > 1e-14 else 0.0
        out.append(times[lo] + min(max(fraction, 0.0), 1.0) * (times[hi] - times[lo]))
    return out


def stencil_np(states, axes):
    """Vectorised stencil of many states (numpy, as the reader computes it,
    relative_table.py:76-109): indices [m, 32], weights [m, 32], valid [m]."""
    import numpy as np

    s = np.asarray(states, dtype=float).reshape(-1, 5)[:, [3, 4, 2, 1, 0]]
    valid = np.isfinite(s).all(1)
    s = np.where(np.isfinite(s), s, 0)
    shape = [len(a) for a in axes]
    strides = np.cumprod([1] + shape[:0:-1])[::-1]
    bits = (np.arange(32)[:, None] >> (4 - np.arange(5))[None, :]) & 1
    ii = np.zeros((len(s), 32), dtype=np.int64)
    ww = np.ones((len(s), 32))
    for k, axis in enumerate(axes):
        step, n = axis[1] - axis[0], shape[k]
        if k == 2:
            f = ((s[:, k] + np.pi) % (2 * np.pi)) / step
            j = np.floor(f).astype(np.int64) % n
            w = f - np.floor(f)
        else:
            valid &= (s[:, k] >= axis[0] - 1e-10) & (s[:, k] <= axis[-1] + 1e-10)
            f = (np.clip(s[:, k], axis[0], axis[-1]) - axis[0]) / step
            j = np.minimum(np.floor(f).astype(np.int64), n - 2)
            w = f - j
        node = j[:, None] + bits[None, :, k]
        if k == 2:
            node %= n
        ii += node * strides[k]
        ww *= np.where(bits[None, :, k], w[:, None], 1 - w[:, None])
    return ii, ww, valid


def field_np(fields, times, axes, states, budgets):
    """Vectorised field (relative_table.py:111-131); fields [frames, N]."""
    import numpy as np

    ii, ww, valid = stencil_np(states, axes)
    times = np.asarray(times, dtype=float)
    tau = np.asarray(budgets, dtype=float).reshape(-1)
    valid &= np.isfinite(tau) & (tau >= 0) & (tau <= times[-1] + 1e-10)
    tau = np.where(np.isfinite(tau), tau, 0)
    k = np.clip(np.searchsorted(times, tau, side="right") - 1, 0, len(times) - 2)
    fraction = (tau - times[k]) / (times[k + 1] - times[k])
    a = (fields[k[:, None], ii] * ww).sum(1)
    b = (fields[(k + 1)[:, None], ii] * ww).sum(1)
    return np.where(valid, a + fraction * (b - a), np.nan)
