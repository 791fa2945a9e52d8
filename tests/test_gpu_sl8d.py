"""GPU parity of the 8D contact-game path (through the C ABI).

* Geometry kernels (FP64) against the oracle, which is bit-identical to the
  reference's own OpenMP library: successors to 1e-12, clearance to 1e-12,
  backup-safety, action-legality and connection bytes exactly equal.
* Backup sweeps (FP32) against the FP64 oracle sweep: seeded random tables
  (the reference's fake_grid test, tests/test_compact_dp.py:53-118) to 2e-6,
  a small real map over several sweeps, and C2 at full size (one sweep
  compared on every pair, 40 sweeps of properties).  Tolerance stated in
  each test; capture-set membership must agree except within tolerance of 0.
* The early exit, the control ordering, the hipGraph replay and the choice
  of E-side kernel (G rows staged in LDS, or gathered from global memory)
  change nothing: results are bit-identical to the plain sweep.
* FP64 solver handles do the oracle's arithmetic and match it bit for bit
  (seeded tables over several sweeps, one full C2 sweep).  With them, the
  FP32 solve of C2 run to residual 1e-4 is compared with the FP64 solve
  after the same sweeps: max|dV| <= 1e-3, capture set equal except within
  1e-3 of zero.
"""

from types import SimpleNamespace

import numpy as np
import pytest

from pe_amd import _native
from pe_amd.local_game.geometry import Footprint, random_rectangle_arena
from pe_amd.local_game.grid_game import (
    GameSolver,
    PairDomain,
    dense_inside,
    game_tables,
)
from pe_amd.local_game.native import GeometryEngine
from pe_amd.local_game.reference import benchmark_game, build_grids, contact_config
from pe_amd.reversible.model import VehicleLimits

pytestmark = pytest.mark.gpu
EARLY, ORDERED, GRAPH = _native.PE_SL_EARLY_EXIT, _native.PE_SL_ORDERED, _native.PE_SL_GRAPH
GLOBAL = _native.PE_SL_GLOBAL_GATHER
WIDE = _native.PE_SL_WIDE_IDS
ONE_PASS, U_MAJOR = _native.PE_SL_ONE_PASS, _native.PE_SL_U_MAJOR
FREEZE = _native.PE_SL_FREEZE
CONTROLS = np.array([[a, k] for a in (-3, 0, 3) for k in (-1, 0, 1)], float)


def seeded_states(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(-3, 3, n), rng.uniform(-3, 3, n),
                     rng.uniform(-np.pi, np.pi, n), rng.uniform(-1, 2, n)], 1)


def test_geometry_kernels_match_oracle(gpu, oracle_lib):
    arena = random_rectangle_arena(3, pursuer=Footprint(0.5, 0.3), evader=Footprint(0.5, 0.3))
    lim = VehicleLimits(-1.0, 2.0, 3.0, 1.0)
    eng = GeometryEngine(arena, lim, "pursuer", control_dt=0.05)
    obs, par = eng.obstacles, eng.params
    s = seeded_states(4000, 0)
    print("[gpu] geometry kernels", flush=True)
    c_gpu = eng.clearance(s)
    c_ref = oracle_lib.geom_clearance(s, obs, 3.0, par)
    assert np.abs(c_gpu - c_ref).max() < 1e-12
    b_gpu = eng.backup_safe(s)
    b_ref = oracle_lib.geom_backups(s, obs, 3.0, par)
    np.testing.assert_array_equal(b_gpu, b_ref)
    ok_gpu, nxt_gpu = eng.actions(s[:1500], CONTROLS, 0.05)
    ok_ref, nxt_ref = oracle_lib.geom_actions(s[:1500], CONTROLS, 0.05, obs, 3.0, par)
    err = nxt_gpu - nxt_ref
    err[..., 2] = (err[..., 2] + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(err).max() < 1e-12
    np.testing.assert_array_equal(ok_gpu, ok_ref)
    t = s[::-1] + np.random.default_rng(1).normal(0, 0.2, s.shape)
    k_gpu = eng.connections(s, t)
    k_ref = oracle_lib.geom_connections(s, t, obs, 3.0, par)
    np.testing.assert_array_equal(k_gpu, k_ref)


def fake_grid(n, na, rng):
    states = rng.uniform(-1, 1, (n, 4))
    ids = rng.integers(n, size=(n, na, 16), dtype=np.int32)
    w = rng.random((n, na, 16)).astype(np.float32)
    w /= w.sum(-1, keepdims=True)
    allowed = rng.random((n, na)) > 0.3
    allowed[:, 0] = True
    return SimpleNamespace(states=states, successors=ids, weights=w, allowed=allowed)


def fake_tables(p, e, radius):
    """Each state is its own cell, so the inside table is the pair mask."""
    d = np.linalg.norm(p.states[:, None, :2] - e.states[None, :, :2], axis=-1)
    inside = (d <= radius).astype(np.uint8)

    def car(g):
        return dict(states=g.states, successors=g.successors, weights=g.weights,
                    allowed=g.allowed, position=np.arange(len(g.states), dtype=np.int32),
                    positions=len(g.states))

    return dict(p=car(p), e=car(e), inside=inside, near=np.ones_like(inside),
                exit_value=float(radius), p_body=(0.8, 0.5), e_body=(0.4, 0.25)), inside


# The last case has E ids beyond 16 bits (32-bit tables, global-gather E side).
@pytest.mark.parametrize("seed,np_,ne,na", [(114, 5, 7, 3), (57, 5, 7, 4), (8, 300, 257, 9),
                                           (5, 3, 70000, 3)])
def test_sweeps_match_fp64_oracle_on_random_tables(gpu, oracle_lib, seed, np_, ne, na):
    rng = np.random.default_rng(seed)
    p, e = fake_grid(np_, na, rng), fake_grid(ne, na, rng)
    tables, inside = fake_tables(p, e, 1.2)
    print(f"[gpu] random tables seed={seed} {np_}x{ne}", flush=True)
    with GameSolver(tables) as s:
        V0 = s.read().astype(np.float64)
        expect = oracle_lib.sl_init(p.states, e.states, inside, 1.2,
                                    Footprint(0.8, 0.5), Footprint(0.4, 0.25))
        np.testing.assert_allclose(V0, expect, atol=5e-7)
        V = V0
        for k in range(3):
            V, change = oracle_lib.sl_sweep(V, p, e, inside)
            s.sweep(1, 0)
            got = s.read()
            np.testing.assert_allclose(got, V, atol=2e-6)
            assert s.trace()[-1, 0] == pytest.approx(change, abs=2e-6)


def run_variants(tables, sweeps):
    out = {}
    for name, flags in [("plain", 0), ("early", EARLY), ("ordered", EARLY | ORDERED),
                        ("graph", EARLY | ORDERED | GRAPH), ("global", EARLY | ORDERED | GLOBAL),
                        ("global_plain", GLOBAL), ("wide_ids", EARLY | ORDERED | WIDE),
                        ("wide_ids_global", EARLY | ORDERED | GLOBAL | WIDE),
                        ("one_pass", EARLY | ORDERED | ONE_PASS),
                        ("u_major", EARLY | ORDERED | U_MAJOR),
                        ("one_pass_full", ORDERED | ONE_PASS),
                        ("u_major_global", EARLY | ORDERED | U_MAJOR | GLOBAL),
                        ("one_pass_global_wide", EARLY | ORDERED | ONE_PASS | GLOBAL | WIDE),
                        ("separate_p", EARLY | ORDERED | _native.PE_SL_SEPARATE_P_SIDE)]:
        with GameSolver(tables) as s:
            s.sweep(sweeps, flags)
            out[name] = (s.read(), s.trace()[:, 0])
    return out


# 9 commands per state take the one-pass kernel with 9-slot running maxima,
# 12 the one with room for 15.
@pytest.mark.parametrize("na", [9, 12])
def test_early_exit_ordering_and_graph_are_exact(gpu, na):
    rng = np.random.default_rng(21)
    p, e = fake_grid(400, na, rng), fake_grid(300, na, rng)
    tables, _ = fake_tables(p, e, 1.0)
    print(f"[gpu] flag variants, {na} commands per state", flush=True)
    res = run_variants(tables, 6)
    base, trace = res["plain"]
    for name in ("early", "ordered", "graph", "global", "global_plain", "wide_ids",
                 "wide_ids_global", "one_pass", "u_major", "one_pass_full", "u_major_global",
                 "one_pass_global_wide", "separate_p"):
        np.testing.assert_array_equal(res[name][0], base, err_msg=name)
        np.testing.assert_array_equal(res[name][1], trace, err_msg=name)


@pytest.fixture(scope="module")
def small_game():
    config = contact_config()
    body = Footprint(0.5, 0.3)
    arena = random_rectangle_arena(11, pursuer=body, evader=body, count=12)
    print("[gpu] small game grids", flush=True)
    p, e = build_grids(config, arena, headings=8, positions=7, half_width=3.0,
                       p_speeds=3, e_speeds=3)
    return config, arena, p, e, game_tables(p, e, arena, 3.0)


def test_small_map_sweeps_match_oracle(gpu, oracle_lib, small_game):
    config, arena, p, e, tables = small_game
    inside = dense_inside(p, e, 3.0)
    with GameSolver(tables) as s:
        V = oracle_lib.sl_init(p.states, e.states, inside, 3.0, arena.pursuer, arena.evader)
        np.testing.assert_allclose(s.read(), V, atol=5e-7)
        for k in range(20):
            V, _ = oracle_lib.sl_sweep(V, p, e, inside)
        s.sweep(20, EARLY | ORDERED)
        got = s.read()
    d = np.abs(got - V)
    print(f"small map ({len(p.states)}x{len(e.states)}) 20 sweeps: max|dV|={d.max():.3e} "
          f"mean={d.mean():.3e}")
    assert d.max() <= 1e-5
    flip = np.flatnonzero((got.ravel() <= 0) != (V.ravel() <= 0))
    assert (np.abs(V.ravel()[flip]) <= 1e-5).all()


def test_small_map_visiting_orders_are_exact(gpu, small_game):
    """On a real map (structured zero weights, walls, a stored command order
    that changes as values settle), the one-pass and u-major E sides and
    the plain index order give the same values over 300 sweeps."""
    tables = small_game[4]
    out = {}
    for name, flags in [("plain", 0), ("one_pass", EARLY | ORDERED | ONE_PASS | GRAPH),
                        ("u_major", EARLY | ORDERED | U_MAJOR | GRAPH)]:
        with GameSolver(tables) as s:
            s.sweep(300, flags)
            out[name] = (s.read(), s.trace()[:, 0])
    for name in ("one_pass", "u_major"):
        np.testing.assert_array_equal(out[name][0], out["plain"][0], err_msg=name)
        np.testing.assert_array_equal(out[name][1], out["plain"][1], err_msg=name)


@pytest.mark.parametrize("game", ["small_game", "c2_game"])
def test_union_row_p_side_agrees_with_stencil_p_side(gpu, oracle_lib, game, request):
    """The FP32 P side that adds each of a P state's union rows to every
    command's sum (the default) against the per-command stencil sums in
    vertex order (PE_SL_STENCIL_P_SIDE): the same values to rounding over
    20 sweeps, and both within the FP32 bar of the FP64 oracle sweep."""
    config, arena, p, e, tables = request.getfixturevalue(game)
    inside = dense_inside(p, e, 3.0)
    out = {}
    for name, flags in [("union", EARLY | ORDERED | GRAPH),
                        ("stencil", EARLY | ORDERED | GRAPH | _native.PE_SL_STENCIL_P_SIDE)]:
        with GameSolver(tables) as s:
            v0 = s.read()
            s.sweep(1, flags & ~GRAPH)
            out[name + "1"] = s.read()
            s.sweep(19, flags)
            out[name] = s.read()
    d = np.abs(out["union"] - out["stencil"])
    print(f"{game}: union vs stencil P side after 20 sweeps: max|dV| {d.max():.3e}")
    assert d.max() <= 5e-6  # a few FP32 ulps of values up to 3 (measured 1.7e-6)
    ref1, _ = oracle_lib.sl_sweep(v0.astype(np.float64), p, e, inside, threads=16)
    for name in ("union1", "stencil1"):
        assert np.abs(out[name] - ref1).max() <= 1e-5, name


@pytest.mark.parametrize("game", ["small_game", "c2_game"])
def test_fused_p_side_equals_separate_p_side(gpu, game, request):
    """The default FP32 sweep computes each P state's values inside its
    E-side block (no G buffer); with PE_SL_SEPARATE_P_SIDE the union-row P
    side writes G and the E side stages it.  Same sums, so the same bits:
    fields, residual traces and stored command orders over 30 sweeps."""
    tables = request.getfixturevalue(game)[4]
    out = {}
    for name, flags in [("fused", EARLY | ORDERED | GRAPH),
                        ("separate", EARLY | ORDERED | GRAPH | _native.PE_SL_SEPARATE_P_SIDE),
                        ("fused_u_major", EARLY | ORDERED | U_MAJOR),
                        ("separate_u_major", EARLY | ORDERED | U_MAJOR |
                         _native.PE_SL_SEPARATE_P_SIDE)]:
        with GameSolver(tables) as s:
            s.sweep(30, flags)
            out[name] = (s.read(), s.trace()[:, 0])
    for a, b in (("fused", "separate"), ("fused_u_major", "separate_u_major"),
                 ("fused", "fused_u_major")):
        np.testing.assert_array_equal(out[a][0], out[b][0], err_msg=f"{a} {b}")
        np.testing.assert_array_equal(out[a][1], out[b][1], err_msg=f"{a} {b}")


def test_lds_block_sizes_are_exact(gpu, small_game):
    """The LDS E side gives the same values at 256, 512 and 1024 threads per
    block, in both visiting orders, with graph replay; other sizes are
    rejected."""
    tables = small_game[4]
    out = {}
    for threads in (256, 512, 1024):
        for name, flags in [("one_pass", EARLY | ORDERED | GRAPH),
                            ("u_major", EARLY | ORDERED | U_MAJOR)]:
            with GameSolver(tables) as s:
                s.sweep(4, flags)  # the handle's own size first
                s.set_lds_threads(threads)
                assert s.stats()["lds_e_side"] == threads
                s.sweep(60, flags)
                out[threads, name] = (s.read(), s.trace()[:, 0])
    base = out[1024, "one_pass"]
    for key, (v, tr) in out.items():
        np.testing.assert_array_equal(v, base[0], err_msg=str(key))
        np.testing.assert_array_equal(tr, base[1], err_msg=str(key))
    with GameSolver(tables) as s:
        with pytest.raises(ValueError):
            s.set_lds_threads(128)


def test_given_field_outside_the_radius_is_kept(gpu, small_game):
    """An initial field whose pairs beyond the exit radius hold another
    value than the exit value: those pairs keep it through an odd number of
    sweeps (both value buffers), and the LDS E side (which writes inside
    pairs only) gives the same field as the global-gather E side (which
    rewrites every pair)."""
    config, arena, p, e, tables = small_game
    inside = dense_inside(p, e, 3.0)
    with GameSolver(tables) as s:
        v0 = s.read().copy()
    v0[~inside] = 7.0
    out = {}
    for name, flags in [("lds", EARLY | ORDERED), ("global", EARLY | ORDERED | GLOBAL)]:
        with GameSolver(tables) as s:
            s.initialize(v0)
            s.sweep(3, flags)
            out[name] = s.read()
    assert (out["lds"][~inside] == 7.0).all()
    np.testing.assert_array_equal(out["lds"], out["global"])


def test_conflicting_sweep_flags_are_rejected(gpu, small_game):
    with GameSolver(small_game[4]) as s:
        with pytest.raises(ValueError):
            s.sweep(1, EARLY | ORDERED | ONE_PASS | U_MAJOR)
        with pytest.raises(ValueError):
            s.sweep(1, 1 << 20)


def test_small_map_converged_matches_fp64_oracle(gpu, oracle_lib, small_game):
    """Converged solve: the FP32 GPU iteration run to residual max|dV| <= 1e-4
    against the FP64 oracle after the same number of sweeps.  Bar:
    max|dV| <= 1e-3, capture-set (V <= 0) membership equal except where the
    FP64 value is within 1e-3 of zero (those pairs are printed)."""
    import time

    config, arena, p, e, tables = small_game
    inside = dense_inside(p, e, 3.0)
    with GameSolver(tables) as s:
        done, ok = s.run(1e-4, 20000, 1, EARLY | ORDERED)
        got = s.read()
        gpu_res = s.trace()[-1, 0]
    print(f"[gpu] small map converged in {done} sweeps (residual {gpu_res:.2e})", flush=True)
    assert ok
    V = oracle_lib.sl_init(p.states, e.states, inside, 3.0, arena.pursuer, arena.evader)
    t0 = time.perf_counter()
    for k in range(done):
        V, change = oracle_lib.sl_sweep(V, p, e, inside, threads=16)
        if k % 500 == 499:
            print(f"oracle sweep {k + 1}/{done} ({time.perf_counter() - t0:.0f} s)", flush=True)
    d = np.abs(got - V)
    flip = np.flatnonzero((got.ravel() <= 0) != (V.ravel() <= 0))
    print(f"converged small map: max|dV|={d.max():.3e} mean={d.mean():.3e}, FP64 residual "
          f"{change:.2e}, capture-set flips {len(flip)} at |V64| "
          f"{np.abs(V.ravel()[flip]).tolist()[:10]}")
    assert d.max() <= 1e-3
    assert (np.abs(V.ravel()[flip]) <= 1e-3).all()


def test_query_and_packed_layout(gpu, small_game):
    config, arena, p, e, tables = small_game
    domain = PairDomain.build(p, e, 3.0)
    with GameSolver(tables) as s:
        s.sweep(5, EARLY)
        field = s.read()
        packed = domain.pack(field)
        np.testing.assert_array_equal(domain.unpack(packed, 3.0, len(e.states)), field)
        rng = np.random.default_rng(0)
        ip = rng.integers(len(p.states), size=64)
        ie = rng.integers(len(e.states), size=64)
        pi, pw, _, _, _ = p.stencil(p.states[ip])
        ei, ew, _, _, _ = e.stencil(e.states[ie])
        got = s.query(pi, pw, ei, ew)
    np.testing.assert_allclose(got, field[ip, ie], atol=1e-6)


def _continuous_states(g, m, rng):
    """Seeded states spread over a car grid's box: any heading, positions
    and speeds between the axis ends."""
    lo, hi = [a[0] for a in g.axes], [a[-1] for a in g.axes]
    return np.stack([rng.uniform(lo[0], hi[0], m), rng.uniform(lo[1], hi[1], m),
                     rng.uniform(-np.pi, np.pi, m), rng.uniform(lo[3], hi[3], m)], 1)


def oracle_stencil(oracle_lib, g, q):
    geo = g.geometry
    ids, w, valid, _, _, _ = oracle_lib.car_stencil(g.axes, g.dense_to_compact, q,
                                                    geo.obstacles, geo.arena.half_width,
                                                    geo.params)
    return ids, w, valid


def test_continuous_queries_match_product_formula(gpu, oracle_lib, small_game):
    """pe_sl_query (query_field, reference.py:127-135) at 4,000 continuous
    pair states: the product of the two cars' 16-vertex stencils, taken
    from the oracle's stencil, summed in float64 over the solver's field.
    The kernel sums in float32 (up to 256 terms of magnitude <= 4, a few
    float32 ulps): agreement to 4e-6."""
    config, arena, p, e, tables = small_game
    rng = np.random.default_rng(1)
    pi, pw, pv = oracle_stencil(oracle_lib, p, _continuous_states(p, 4000, rng))
    ei, ew, ev = oracle_stencil(oracle_lib, e, _continuous_states(e, 4000, rng))
    keep = pv & ev
    pi, pw, ei, ew = pi[keep], pw[keep], ei[keep], ew[keep]
    with GameSolver(tables) as s:
        s.sweep(12, EARLY | ORDERED)
        field = s.read()
        got = s.query(pi, pw, ei, ew)
    pair = field[np.maximum(pi, 0)[:, :, None], np.maximum(ei, 0)[:, None, :]]
    want = np.einsum("na,nb,nab->n", pw.astype(np.float64), ew.astype(np.float64),
                     pair.astype(np.float64))
    d = np.abs(got - want)
    print(f"continuous queries: {keep.sum()} valid of 4000, max|dV|={d.max():.2e}")
    assert keep.sum() > 1000 and d.max() <= 4e-6


C2_SEED = 0


@pytest.fixture(scope="module")
def c2_game():
    print("[gpu] C2 game build (11x11x12x5 per car, 30 rectangles)", flush=True)
    return benchmark_game(C2_SEED)


def test_c2_full_size(gpu, oracle_lib, c2_game):
    config, arena, p, e, tables = c2_game
    inside = dense_inside(p, e, 3.0)
    print(f"C2: {len(p.states)} x {len(e.states)} legal states, {inside.sum()} pairs inside "
          f"of {len(p.states) * len(e.states)}", flush=True)
    with GameSolver(tables) as s:
        V0 = s.read()
        s.sweep(1, EARLY | ORDERED)
        V1 = s.read()
        ref1, _ = oracle_lib.sl_sweep(V0.astype(np.float64), p, e, inside, threads=16)
        d = np.abs(V1 - ref1)
        print(f"C2 one sweep vs FP64 oracle: max|dV|={d.max():.3e} mean={d.mean():.3e}")
        assert d.max() <= 1e-5
        s.sweep(39, EARLY | ORDERED | GRAPH)
        V40 = s.read()
        trace = s.trace()[:, 0]
    assert (V40 <= V1 + 1e-7).all() and (V1 <= V0 + 1e-7).all()
    assert (V40[~inside] == 3.0).all()
    assert (V40[V0 <= 0] <= 0).all()
    assert len(trace) == 40 and (trace >= 0).all()
    print(f"C2 40 sweeps: residual {trace[-1]:.3e}, captured fraction "
          f"{np.mean(V40[inside] <= 0):.4f}")


def test_batched_maps_equal_individual_solves(gpu):
    """C4 layout: several maps in one handle (one launch per kernel, graph
    replay) give bit-identical fields to solving each map alone."""
    from pe_amd.local_game.reference import benchmark_game

    print("[gpu] batched maps", flush=True)
    tables = [benchmark_game(1000 + k, positions=7, headings=8, speeds=3, rectangles=12)[4]
              for k in range(4)]
    with GameSolver(tables) as s:
        s.sweep(10, EARLY | ORDERED | GRAPH)
        batched = [s.read(k) for k in range(4)]
        traces = s.trace()
    for k, t in enumerate(tables):
        with GameSolver(t) as s1:
            s1.sweep(10, EARLY | ORDERED)
            np.testing.assert_array_equal(s1.read(), batched[k])
            np.testing.assert_array_equal(s1.trace()[:, 0], traces[:, k])


def test_multi_map_table_build_equals_single_maps(gpu):
    """benchmark_games (legality and action tables of all maps in one launch
    per phase, pe_geom_*_multi) gives the tables of benchmark_game map by
    map, bit for bit."""
    from pe_amd.local_game.reference import benchmark_game, benchmark_games

    print("[gpu] multi-map table build", flush=True)
    seeds = [1000, 1001, 1002]
    many = benchmark_games(seeds, positions=7, headings=8, speeds=3, rectangles=12)
    for seed, (_, _, p, e, t) in zip(seeds, many):
        _, _, p1, e1, t1 = benchmark_game(seed, positions=7, headings=8, speeds=3, rectangles=12)
        for a, b in ((p, p1), (e, e1)):
            for name in ("states", "dense_to_compact", "successors", "weights", "allowed"):
                np.testing.assert_array_equal(getattr(a, name), getattr(b, name),
                                              err_msg=f"seed {seed} {name}")
        np.testing.assert_array_equal(t["inside"], t1["inside"])
        np.testing.assert_array_equal(t["near"], t1["near"])
    # The multi-map kernels against the single-map entry points.
    from pe_amd.local_game.native import actions_many, backup_safe_many

    engines = [g[2].geometry for g in many]
    dense = np.stack(np.meshgrid(*many[0][2].axes, indexing="ij"), -1).reshape(-1, 4)
    for eng, legal in zip(engines, backup_safe_many(engines, [dense] * len(engines))):
        np.testing.assert_array_equal(legal, eng.backup_safe(dense))
    kept = [g[2].states for g in many]
    controls = many[0][2].controls
    for eng, x, (ok, nxt) in zip(engines, kept, actions_many(engines, kept, controls, 0.05)):
        ok1, nxt1 = eng.actions(x, controls, 0.05)
        np.testing.assert_array_equal(ok, ok1)
        np.testing.assert_array_equal(nxt, nxt1)


def test_freeze_stops_each_map_as_if_solved_alone(gpu):
    """pe_sl_run with PE_SL_FREEZE on a batch: each map stops being swept at
    the first check where its own residual is at most tol, so its field and
    its stopping sweep equal a run of that map alone; afterwards every map
    is swept again, from the field it stopped at."""
    from pe_amd.local_game.reference import benchmark_game

    print("[gpu] freeze converged maps", flush=True)
    tables = [benchmark_game(1000 + k, positions=7, headings=8, speeds=3, rectangles=12)[4]
              for k in range(4)]
    flags = EARLY | ORDERED | GRAPH
    tol, every = 1e-4, 16
    alone, stops, after = [], [], []
    for t in tables:
        with GameSolver(t) as s1:
            done, ok = s1.run(tol, 20_000, every, flags)
            assert ok
            alone.append(s1.read())
            stops.append(done)
            s1.sweep(6, flags)
            after.append(s1.read())
    assert len(set(stops)) > 1, "maps should converge at different sweeps"
    with GameSolver(tables) as s:
        done, ok = s.run(tol, 20_000, every, flags | FREEZE)
        assert ok and done == max(stops)
        trace = s.trace()
        for k in range(4):
            np.testing.assert_array_equal(s.read(k), alone[k], err_msg=f"map {k}")
            # swept up to its own stop, residual zero (not swept) after it
            assert (trace[:stops[k], k] > 0).any() and (trace[stops[k]:, k] == 0).all()
        s.sweep(6, flags)
        for k in range(4):
            np.testing.assert_array_equal(s.read(k), after[k], err_msg=f"map {k} after")


@pytest.mark.parametrize("flags", [0, EARLY | ORDERED, EARLY | ORDERED | GRAPH,
                                   EARLY | ORDERED | GLOBAL])
def test_fp64_sweeps_are_bit_exact_with_oracle(gpu, oracle_lib, flags):
    """FP64 handles do the oracle's arithmetic (each stencil sum a separate
    multiply and add, left to right over the positive weights), so from the
    oracle's own initial field every sweep agrees bit for bit, whatever the
    early exit, control order, graph replay or E-side kernel."""
    rng = np.random.default_rng(8)
    p, e = fake_grid(300, 9, rng), fake_grid(257, 9, rng)
    tables, inside = fake_tables(p, e, 1.2)
    V = oracle_lib.sl_init(p.states, e.states, inside, 1.2, Footprint(0.8, 0.5),
                           Footprint(0.4, 0.25))
    print(f"[gpu] fp64 sweeps flags={flags}", flush=True)
    with GameSolver(tables, precision=1) as s:
        s.initialize(V)
        for k in range(4):
            V, change = oracle_lib.sl_sweep(V, p, e, inside)
            s.sweep(1, flags & ~GRAPH)
        s.sweep(2, flags)
        V, _ = oracle_lib.sl_sweep(V, p, e, inside)
        V, change = oracle_lib.sl_sweep(V, p, e, inside)
        got = s.read()
        trace = s.trace()[:, 0]
    assert got.dtype == np.float64
    np.testing.assert_array_equal(got, V)
    assert trace[-1] == change


def test_c2_fp64_sweep_bit_exact_and_converged_fp32_vs_fp64(gpu, oracle_lib, c2_game):
    """Full C2: one FP64 GPU sweep equals the FP64 oracle bit for bit; then
    the FP32 solve run to residual max|dV| <= 1e-4 against the FP64 solve
    after the same number of sweeps.  Bar: max|dV| <= 1e-3, capture-set
    (V <= 0) membership equal except where |V64| <= 1e-3 (listed)."""
    config, arena, p, e, tables = c2_game
    inside = dense_inside(p, e, 3.0)
    V0 = oracle_lib.sl_init(p.states, e.states, inside, 3.0, arena.pursuer, arena.evader)
    print("[gpu] C2 FP64 sweep vs oracle", flush=True)
    with GameSolver(tables, precision=1) as s64:
        s64.initialize(V0)
        s64.sweep(1, EARLY | ORDERED)
        ref1, _ = oracle_lib.sl_sweep(V0, p, e, inside, threads=16)
        np.testing.assert_array_equal(s64.read(), ref1)
        print("[gpu] C2 FP32 to residual 1e-4", flush=True)
        with GameSolver(tables) as s32:
            s32.initialize(V0.astype(np.float32))
            done, ok = s32.run(1e-4, 20000, 1, EARLY | ORDERED)
            v32 = s32.read()
        assert ok
        s64.initialize(V0)
        s64.sweep(done, EARLY | ORDERED | GRAPH)
        v64 = s64.read()
        res64 = s64.trace()[-1, 0]
    d = np.abs(v32 - v64)
    flip = np.flatnonzero((v32.ravel() <= 0) != (v64.ravel() <= 0))
    print(f"C2 converged ({done} sweeps, FP64 residual {res64:.2e}): FP32 vs FP64 "
          f"max|dV|={d.max():.3e} mean={d.mean():.3e}, capture-set flips {len(flip)} "
          f"at |V64| {np.abs(v64.ravel()[flip])[:10].tolist()}")
    assert d.max() <= 1e-3
    assert (np.abs(v64.ravel()[flip]) <= 1e-3).all()


def identity_car(count, x):
    """`count` states at (x, 0, 0, 0), one command whose 16 stencil vertices
    are the state itself with the whole weight on the first."""
    states = np.zeros((count, 4))
    states[:, 0] = x
    ids = np.repeat(np.arange(count, dtype=np.int32)[:, None, None], 16, axis=2)
    w = np.zeros((count, 1, 16), np.float32)
    w[:, :, 0] = 1
    return SimpleNamespace(states=states, successors=ids, weights=w,
                           allowed=np.ones((count, 1), bool))


@pytest.mark.parametrize("n_p,n_e", [(65536, 1), (1, 65536)])
def test_more_than_65535_vehicle_states(gpu, oracle_lib, n_p, n_e):
    """(tests/test_compact_dp.py:121-133) 65,536 states of one car at x = 0
    against the other car at x = 1: bodies 0.8 and 0.4 long give the
    initial value 1 - 0.4 - 0.2 = 0.4 on every pair.  With 65,536 E states
    the 16-bit packed E records cannot index the E side and the 32-bit
    tables are used.  Identity stencils keep every value after a sweep; the
    sweep is compared with the FP64 oracle's on every pair, for both E-side
    kernels."""
    p, e = identity_car(n_p, 0.0), identity_car(n_e, 1.0)
    tables, inside = fake_tables(p, e, 3.0)
    print(f"[gpu] identity tables {n_p}x{n_e}", flush=True)
    expect0 = oracle_lib.sl_init(p.states, e.states, inside, 3.0,
                                 Footprint(0.8, 0.5), Footprint(0.4, 0.25))
    np.testing.assert_allclose(expect0, 0.4, atol=1e-12)
    expect1, _ = oracle_lib.sl_sweep(expect0, p, e, inside)
    for flags in (EARLY, EARLY | GLOBAL):
        with GameSolver(tables) as s:
            V0 = s.read()
            assert V0.shape == (n_p, n_e)
            np.testing.assert_allclose(V0, 0.4, atol=1e-6)
            s.sweep(1, flags)
            np.testing.assert_allclose(s.read(), expect1, atol=1e-6)
