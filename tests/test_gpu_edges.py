"""Edge cases of the query and table entry points, through the C ABI on the
GPU: empty batches return empty results without a launch, mismatched or
out-of-range inputs are refused with the library's error (never read past a
buffer), and the allocation-free reader refuses batches above its limit."""

import numpy as np
import pytest

from pe_amd import _native

pytestmark = pytest.mark.gpu


def small_history():
    axes = [np.linspace(-1, 2, 3), np.linspace(-0.5, 1, 3),
            np.arange(8) * 2 * np.pi / 8 - np.pi, np.linspace(-1, 1, 5), np.linspace(-1, 1, 5)]
    shape = [len(a) for a in axes]
    frames = np.stack([np.full(shape, 1.0 - 0.1 * k, np.float32).ravel() for k in range(3)])
    return _native.RelTable(frames, axes, np.array([0.0, 0.1, 0.2])), axes


def test_empty_batches(gpu):
    print("[gpu] empty batches", flush=True)
    table, axes = small_history()
    try:
        assert table.field(np.zeros((0, 5)), np.zeros(0)).shape == (0,)
        assert table.field_noalloc(np.zeros((0, 5)), np.zeros(0)).shape == (0,)
        assert table.value(np.zeros((0, 5)), 0.3, [-1, 2, -0.5, 1]).shape == (0,)
    finally:
        table.close()
    n = [len(a) for a in axes]
    idx, w, valid = _native.grid_stencil(np.zeros((0, 5)), n, [a[0] for a in axes],
                                         [a[-1] for a in axes], [a[1] - a[0] for a in axes])
    assert idx.shape == (0, 32) and w.shape == (0, 32) and valid.shape == (0,)
    car_axes = (np.linspace(-1, 1, 3), np.linspace(-1, 1, 3),
                np.arange(4) * np.pi / 2 - np.pi, np.linspace(0, 1, 2))
    d2c = np.arange(3 * 3 * 4 * 2, dtype=np.int32)
    ids, w, valid, mass, fb = _native.car_stencil(car_axes, d2c, np.zeros((0, 4)))
    assert ids.shape == (0, 16) and fb == 0
    hist = _native.History(np.zeros((2, 10), np.float16))
    try:
        out = hist.query(np.zeros((0, 4), np.int64), np.zeros((0, 4)), np.zeros((0, 4)), 3.0)
        assert out.shape == (2, 0)
    finally:
        hist.close()


def test_refused_inputs(gpu):
    print("[gpu] refused inputs", flush=True)
    table, _ = small_history()
    try:
        with pytest.raises(ValueError):
            table.field(np.zeros((4, 5)), np.zeros(3))
        m = _native.PE_RTABLE_NOALLOC_MAX
        with pytest.raises(ValueError):
            table.field_noalloc(np.zeros((m + 1, 5)), np.zeros(m + 1))
        # the largest allowed batch is served
        got = table.field_noalloc(np.zeros((m, 5)), np.full(m, 0.05))
        np.testing.assert_allclose(got, 0.95, atol=1e-6)
    finally:
        table.close()
    hist = _native.History(np.zeros((2, 10), np.float16))
    try:
        with pytest.raises(ValueError):  # a kept pair index past the history
            hist.query(np.full((1, 4), 10, np.int64), np.full((1, 4), 0.25),
                       np.ones((1, 4)), 3.0)
    finally:
        hist.close()
    car_axes = (np.linspace(-1, 1, 3), np.linspace(-1, 1, 3),
                np.arange(4) * np.pi / 2 - np.pi, np.linspace(0, 1, 2))
    with pytest.raises(ValueError):  # no legal state
        _native.car_stencil(car_axes, np.full(72, -1, np.int32), np.zeros((1, 4)))
    with pytest.raises(ValueError):  # repeated compact index
        _native.car_stencil(car_axes, np.zeros(72, np.int32), np.zeros((1, 4)))
