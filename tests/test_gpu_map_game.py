"""Footprint-map 8D game on the GPU (pe_world8d in its footprint-map mode):
the reference's independent-numpy check of one update
(tests/test_footprint_map.py:45-103)."""

from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [1, 2])
def test_gpu_map_update_matches_independent_numpy(gpu, order, tmp_path):
    from pe_amd.reversible.footprint_map import FootprintMap
    from pe_amd.reversible.gpu_map_game import solve
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import WorldGrid
    from world8d_numpy import rhs

    print(f"[map] order {order}", flush=True)
    c = ReversibleConfig()
    c = replace(c, pursuer=replace(c.pursuer, acceleration=3),
                evader=replace(c.evader, acceleration=3), capture_radius=0.5,
                point_vehicles=False, arena_half_width=2.5)
    g = WorldGrid(5, 4, 3, 3, 3)
    m = FootprintMap(2.5, ((-0.25, 0.25, -1.4, 0.1),))
    states = np.stack(np.broadcast_arrays(*np.meshgrid(*g.axes(c), indexing="ij",
                                                       sparse=True)), axis=-1)
    physical_initial, avoid = m.surfaces(states, 0.5)
    dt = 0.01
    folder = tmp_path / "gpu"
    meta = solve(c, g, m, folder, horizon=dt, output_dt=dt, order=order, save_final=True,
                 save_history=True, query_indices=np.arange(g.states))
    labels = np.load(folder / "labels.npz")
    initial = labels["values"][0].reshape(g.shape)
    np.testing.assert_allclose(initial, physical_initial, atol=3e-7, rtol=3e-7)
    stage = (initial + dt * rhs(initial, g, c, order)).astype(np.float32)
    expected = stage if order == 1 else 0.5 * (initial + stage + dt * rhs(stage, g, c, order))
    expected = np.maximum(np.minimum(expected, initial), avoid)
    actual = np.fromfile(folder / "final.f32", np.float32).reshape(g.shape)
    np.testing.assert_allclose(actual, expected, atol=3e-6, rtol=3e-6)
    history = np.fromfile(folder / "fields.f32", np.float32).reshape(2, *g.shape)
    np.testing.assert_array_equal(history[0], initial)
    np.testing.assert_array_equal(history[1], actual)
    assert meta["steps"] == 1
    assert meta["target"] == "capture_OR_evader_collision"
    np.testing.assert_array_equal(labels["values"][-1], actual.ravel()[labels["indices"]])


def test_gpu_map_continuous_queries_and_monotone_history(gpu, tmp_path):
    """Continuous query states sample the multilinear interpolation of each
    saved frame; the history never rises and never falls below the avoid
    floor."""
    from pe_amd.reversible.footprint_map import FootprintMap
    from pe_amd.reversible.gpu_map_game import solve
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import WorldGrid
    from pe_amd.reversible.grid_stencil import interpolation_stencil

    print("[map] continuous queries", flush=True)
    c = replace(ReversibleConfig(), point_vehicles=False, arena_half_width=2.5)
    g = WorldGrid(7, 6, 4, 4, 3)
    m = FootprintMap(2.5, ((-0.25, 0.25, -1.4, 0.1), (0.8, 1.4, 0.5, 1.2)))
    rng = np.random.default_rng(3)
    q = np.column_stack([rng.uniform(-2.5, 2.5, (50, 2)), rng.uniform(-3, 3, 50),
                         rng.uniform(-1, 2, 50), rng.uniform(-2.5, 2.5, (50, 2)),
                         rng.uniform(-3, 3, 50), rng.uniform(-0.5, 1, 50)])
    meta = solve(c, g, m, tmp_path / "q", horizon=0.3, output_dt=0.1, save_history=True,
                 query_states=q)
    history = np.fromfile(tmp_path / "q" / "fields.f32", np.float32).reshape(4, -1)
    assert (np.diff(history, axis=0) <= 0).all()
    ii, ww, valid, _ = interpolation_stencil(q, g.axes(c), (2, 6))
    labels = np.load(tmp_path / "q" / "labels.npz")
    for t in range(4):
        np.testing.assert_allclose(labels["values"][t], (history[t][ii] * ww).sum(1),
                                   atol=1e-6)
    assert meta["status"] == "complete" and meta["query_count"] == 50
