"""pe_rtable_field_noalloc (the 5D reader's field query without device
allocation) against pe_rtable_field through the C ABI: the same kernel on the
same history, so the answers must be bit-identical, on two calls in a row
(the second smaller, so stale staging from the first would show) and after
the limit is exceeded (rejected before any device work)."""

import numpy as np
import pytest

from pe_amd import _native

pytestmark = pytest.mark.gpu


def _history(rng):
    shape = (5, 4, 6, 9, 9)  # vP, vE, psi, y, x
    axes = [np.linspace(-1, 2, 5), np.linspace(-0.5, 1, 4),
            -np.pi + np.arange(6) * 2 * np.pi / 6, np.linspace(-2, 2, 9), np.linspace(-2, 2, 9)]
    times = np.array([0.0, 0.1, 0.2, 0.35])
    fields = rng.normal(size=(len(times),) + shape).astype(np.float32)
    return fields, axes, times


def _queries(rng, n):
    s = np.stack([rng.uniform(-2.2, 2.2, n), rng.uniform(-2.2, 2.2, n),
                  rng.uniform(-7, 7, n), rng.uniform(-1.1, 2.1, n),
                  rng.uniform(-0.6, 1.1, n)], 1)
    b = rng.uniform(-0.02, 0.37, n)
    return s, b


def test_noalloc_field_matches_allocating_field(gpu):
    rng = np.random.default_rng(7)
    fields, axes, times = _history(rng)
    t = _native.RelTable(fields, axes, times)
    try:
        for n in (50_000, 1_000):
            print(f"[no_alloc] rtable field, {n} queries", flush=True)
            s, b = _queries(rng, n)
            want = t.field(s, b)
            got = t.field_noalloc(s, b)
            assert np.isfinite(want).sum() > n // 2
            np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
        s, b = _queries(rng, _native.PE_RTABLE_NOALLOC_MAX + 1)
        with pytest.raises(ValueError, match="exceed the limit"):
            t.field_noalloc(s, b)
        s, b = _queries(rng, 3)
        np.testing.assert_array_equal(t.field_noalloc(s, b).view(np.uint64),
                                      t.field(s, b).view(np.uint64))
    finally:
        t.close()
