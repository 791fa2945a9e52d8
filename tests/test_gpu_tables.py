"""GPU table build and solves at the benchmark grids, against the oracle.

* pe_car_stencil against the oracle's stencil (oracle/stencil.c, pinned in
  test_oracle_tables.py) on the same successor states: ids, weights, mass
  and validity bit for bit, fallback rows included, on seeded random grids
  and on real maps.
* The whole table chain of the C4 grid (9 x 9 positions, 8 headings, 4
  speeds per car, 30 rectangles; seeds 1000-1003) and of the C2 grid
  (11 x 11 x 12 x 5, seed 0): legality, command admissibility and
  stencils built from the GPU geometry against the oracle's chain (oracle
  geometry, then oracle stencil).  Legality and admissibility must be
  equal; the successors agree to 1e-12, so the stencils are compared as
  operators: the interpolated value of seeded fields agrees to 1e-6.
* C4 FP32 sweeps against the FP64 oracle sweep over the first sweeps
  (2e-6), and each C4 map converged in FP32 to residual 1e-4 against the
  FP64 GPU solve after the same number of sweeps (max|dV| <= 1e-3,
  capture-set flips only where |V64| <= 1e-3; the flips are listed).
"""

import numpy as np
import pytest

from pe_amd import _native
from pe_amd.local_game.grid_game import GameSolver, command_set, dense_inside
from pe_amd.local_game.reference import benchmark_games, contact_config

pytestmark = pytest.mark.gpu
EARLY, ORDERED, GRAPH = _native.PE_SL_EARLY_EXIT, _native.PE_SL_ORDERED, _native.PE_SL_GRAPH
C4 = dict(positions=9, headings=8, speeds=4)
C4_SEEDS = [1000, 1001, 1002, 1003]


def dense_states(axes):
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 4)


def oracle_geometry(g):
    return g.geometry.obstacles, g.geometry.arena.half_width, g.geometry.params


@pytest.mark.parametrize("seed", [0, 1])
def test_car_stencil_matches_oracle_on_random_grids(gpu, oracle_lib, seed):
    from pe_amd.local_game.geometry import Footprint, random_rectangle_arena

    rng = np.random.default_rng(seed)
    x = np.linspace(-3, 3, 9)
    axes = (x, x.copy(), np.arange(12) * 2 * np.pi / 12 - np.pi, np.linspace(-1, 2, 5))
    full = dense_states(axes)
    legal = rng.random(len(full)) > 0.3
    legal &= ~((np.abs(full[:, 0]) < 1.2) & (np.abs(full[:, 1]) < 1.2))  # forces fallbacks
    d2c = np.where(legal, np.cumsum(legal) - 1, -1).astype(np.int32)
    q = np.stack([rng.uniform(-3.3, 3.3, 20000), rng.uniform(-3.3, 3.3, 20000),
                  rng.uniform(-7, 7, 20000), rng.uniform(-1.2, 2.2, 20000)], 1)
    arena = random_rectangle_arena(5 + seed, pursuer=Footprint(0.5, 0.3),
                                   evader=Footprint(0.5, 0.3), count=12)
    par = np.array([-1.0, 2.0, 3.0, 1.0, 0.5, 0.3, 0.05])
    print(f"[gpu] car stencil vs oracle, random grid seed {seed}", flush=True)
    ids, w, valid, mass, fb = _native.car_stencil(axes, d2c, q, (arena.circles, 3.0, par))
    oids, ow, ovalid, omass, _, ofb = oracle_lib.car_stencil(axes, d2c, q, arena.circles, 3.0,
                                                             par)
    assert fb == ofb and fb > 50
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    np.testing.assert_array_equal(valid, ovalid)
    np.testing.assert_array_equal(mass.view(np.uint64), omass.view(np.uint64))
    # without geometry: the fallback rows keep zero weights and are invalid
    ids2, w2, valid2, _, fb2 = _native.car_stencil(axes, d2c, q)
    miss = (mass <= 1e-9) & np.isfinite(mass)
    assert fb2 == fb and not valid2[miss].any() and (w2[miss] == 0).all()
    np.testing.assert_array_equal(w2[~miss], w[~miss])


@pytest.fixture(scope="module")
def c4_games():
    print("[gpu] C4 tables for seeds 1000-1003", flush=True)
    return benchmark_games(C4_SEEDS, **C4)


def oracle_chain(oracle_lib, config, g, limits):
    """Legality, admissibility and stencils from the oracle alone."""
    axes = g.axes
    full = dense_states(axes)
    obs, half, par = oracle_geometry(g)
    legal = oracle_lib.geom_backups(full, obs, half, par)
    d2c = np.where(legal, np.cumsum(legal) - 1, -1).astype(np.int32)
    kept = full[legal]
    commands = command_set(limits)
    ok, nxt = oracle_lib.geom_actions(kept, commands, config.control_dt, obs, half, par)
    ids, w, valid, mass, _, _ = oracle_lib.car_stencil(axes, d2c, nxt.reshape(-1, 4),
                                                              obs, half, par)
    shape = ok.shape
    return dict(legal=legal, kept=kept, nxt=nxt, allowed=ok & valid.reshape(shape),
                ids=ids.reshape(shape + (16,)), w=w.reshape(shape + (16,)),
                mass=mass.reshape(shape))


def compare_chain(oracle_lib, config, g, limits, label):
    ref = oracle_chain(oracle_lib, config, g, limits)
    np.testing.assert_array_equal(g.dense_to_compact >= 0, ref["legal"], err_msg=label)
    np.testing.assert_array_equal(g.states, ref["kept"], err_msg=label)
    np.testing.assert_array_equal(g.allowed, ref["allowed"], err_msg=label)
    # the same stencil arithmetic on the GPU's own successors: bit for bit
    obs, half, par = oracle_geometry(g)
    ok_gpu, nxt_gpu = g.geometry.actions(g.states, g.controls, config.control_dt)
    err = nxt_gpu - ref["nxt"]
    err[..., 2] = (err[..., 2] + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(err).max() < 1e-12, label
    oids, ow, ovalid, omass, _, _ = oracle_lib.car_stencil(
        g.axes, g.dense_to_compact, nxt_gpu.reshape(-1, 4), obs, half, par)
    np.testing.assert_array_equal(g.successors.reshape(-1, 16), oids, err_msg=label)
    np.testing.assert_array_equal(g.weights.reshape(-1, 16).view(np.uint32), ow.view(np.uint32),
                                  err_msg=label)
    np.testing.assert_array_equal(g.allowed, ok_gpu & ovalid.reshape(ok_gpu.shape),
                                  err_msg=label)
    # the two chains as interpolation operators on seeded fields
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(3):
        f = rng.normal(size=len(g.states))
        a = (g.weights.astype(np.float64) * f[g.successors]).sum(-1)
        b = (ref["w"].astype(np.float64) * f[ref["ids"]]).sum(-1)
        worst = max(worst, float(np.abs(a - b)[g.allowed].max()))
    same = np.mean(g.successors == ref["ids"])
    print(f"{label}: {len(g.states)} legal states, allowed {int(g.allowed.sum())}, "
          f"ids equal {same:.6f}, operator max diff {worst:.2e}", flush=True)
    assert worst <= 1e-6, label


def test_c4_table_chain_matches_oracle(gpu, oracle_lib, c4_games):
    config = contact_config()
    for seed, (_, arena, p, e, _) in zip(C4_SEEDS, c4_games):
        compare_chain(oracle_lib, config, p, config.pursuer, f"C4 seed {seed} pursuer")
        compare_chain(oracle_lib, config, e, config.evader, f"C4 seed {seed} evader")


def test_c2_table_chain_matches_oracle(gpu, oracle_lib):
    config = contact_config()
    print("[gpu] C2 tables seed 0", flush=True)
    (_, arena, p, e, _), = benchmark_games([0])
    compare_chain(oracle_lib, config, p, config.pursuer, "C2 seed 0 pursuer")
    compare_chain(oracle_lib, config, e, config.evader, "C2 seed 0 evader")


def test_c4_sweeps_match_fp64_oracle(gpu, oracle_lib, c4_games):
    _, arena, p, e, tables = c4_games[0]
    inside = dense_inside(p, e, 3.0)
    V = oracle_lib.sl_init(p.states, e.states, inside, 3.0, arena.pursuer, arena.evader)
    print(f"[gpu] C4 seed 1000: FP32 sweeps vs FP64 oracle ({len(p.states)}x{len(e.states)})",
          flush=True)
    with GameSolver(tables) as s:
        np.testing.assert_allclose(s.read(), V, atol=5e-7)
        for k in range(4):
            V, change = oracle_lib.sl_sweep(V, p, e, inside, threads=16)
            s.sweep(1, EARLY | ORDERED)
            got = s.read()
            d = np.abs(got - V)
            assert d.max() <= 2e-6, (k, d.max())
            assert s.trace()[-1, 0] == pytest.approx(change, abs=2e-6)


def test_c4_converged_fp32_vs_fp64(gpu, c4_games):
    """Each C4 map: FP32 run to residual max|dV| <= 1e-4 against the FP64
    GPU solve (bit-exact with the FP64 oracle) after the same sweeps."""
    flags = EARLY | ORDERED | GRAPH
    for seed, (_, arena, p, e, tables) in zip(C4_SEEDS, c4_games):
        print(f"[gpu] C4 seed {seed} converged FP32 vs FP64", flush=True)
        with GameSolver(tables) as s32:
            done, ok = s32.run(1e-4, 30000, 16, flags)
            v32 = s32.read()
        assert ok
        with GameSolver(tables, precision=1) as s64:
            s64.sweep(done, flags)
            v64 = s64.read()
            res64 = s64.trace()[-1, 0]
        d = np.abs(v32 - v64)
        flip = np.flatnonzero((v32.ravel() <= 0) != (v64.ravel() <= 0))
        print(f"C4 seed {seed}: {done} sweeps (FP64 residual {res64:.2e}), "
              f"max|dV|={d.max():.3e} mean={d.mean():.3e}, flips {len(flip)} "
              f"{flip[:20].tolist()} at |V64| {np.abs(v64.ravel()[flip])[:20].tolist()}",
              flush=True)
        assert d.max() <= 1e-3
        assert (np.abs(v64.ravel()[flip]) <= 1e-3).all()


def test_c4_full_batch_equals_maps_swept_alone(gpu):
    """The C4 configuration as batched: all 128 maps (seeds 1000-1127) in one
    handle, 24 graph-replayed sweeps, then 24 more with converged maps
    frozen at a loose tolerance.  Three maps (first, middle, last) swept
    alone the same way give the same fields bit for bit, and the same
    residual trace while every map is still swept."""
    seeds = list(range(1000, 1128))
    print("[gpu] C4 128-map batch: tables", flush=True)
    games = benchmark_games(seeds, **C4)
    flags = EARLY | ORDERED | GRAPH
    with GameSolver([g[4] for g in games]) as batch:
        assert batch.stats()["maps"] == 128
        batch.sweep(24, flags)
        trace = batch.trace()
        done, _ = batch.run(2e-2, 24, 8, flags | _native.PE_SL_FREEZE)
        fields = {k: batch.read(k) for k in (0, 63, 127)}
    print(f"[gpu] C4 128-map batch: {done} frozen-run sweeps", flush=True)
    for k, f in fields.items():
        with GameSolver(games[k][4]) as alone:
            alone.sweep(24, flags)
            np.testing.assert_array_equal(alone.trace()[:, 0], trace[:, k])
            alone.run(2e-2, 24, 8, flags | _native.PE_SL_FREEZE)
            np.testing.assert_array_equal(alone.read().view(np.uint32), f.view(np.uint32))
