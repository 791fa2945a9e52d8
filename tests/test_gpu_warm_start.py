"""Learned warm start of the 8D contact-game iteration, on small maps
(7 x 7 positions, 8 headings, 3 speeds per car) and on maps of the C2 grid
(11 x 11 positions over +-3 m, 12 headings, 5 speeds per car).

The network is trained on the GPU by the library's own BF16-MFMA kernels on
plain-DP fields of two maps, its lift delta is calibrated on a third map
(the largest underestimate over its inside pairs), and the warm field
min(phi, net + delta) is built on the device (pe_sl_warm_start).  On a
fourth, held-out map the cold and the warm run are swept side by side.

What holds whatever the network learned, and is asserted:

  * the warm start never exceeds the payoff phi, and every warm iterate
    stays at or below the cold iterate of the same sweep count.  The sweep
    is monotone in floating point too (non-negative weights, one summation
    order, min and max), so this holds bit for bit, not only up to rounding;
  * no warm value ever rises above its start (the sweep never raises a
    value), so a pair that starts below V* ends below it: the plain DP is
    the judge, and a warm start whose delta left pairs more than 1e-3 below
    V* is a failed warm start, not a slower one;
  * where the start is nowhere below V*, the warm run is within 1e-3 of V*
    no later than the cold run (V* <= warm <= cold at every sweep).

What the network gains is measured and printed, not asserted: on these
maps it is no sweep (the calibrated delta is of the order of the values'
range, so min(phi, net + delta) is phi on the slow pairs).

The network also orders each pair's commands (pe_sl_order_from_mlp) for
the first sweep of a cold run: values and residuals are asserted equal bit
for bit with and without it, and the first sweep's E side is timed.
"""

import numpy as np
import pytest

from pe_amd import _native
from pe_amd.learned import warm_start as ws
from pe_amd.local_game.grid_game import GameSolver

pytestmark = pytest.mark.gpu

CASES = {
    "small": dict(grid=dict(positions=7, headings=8, speeds=3), steps=600, vstar_cap=40_000,
                  chunk=128),
    # the C2 grid is the grid of the warm-start configuration's maps
    "c2_grid": dict(grid={}, steps=400, vstar_cap=30_000, chunk=256),
}
TRAIN, VALIDATE, HELD_OUT = (0, 1), 2, 3


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, gpu):
    name = request.param
    spec = CASES[name]
    grid = spec["grid"]
    print(f"[gpu] warm start {name}: plain DP on maps {TRAIN}", flush=True)
    maps = [ws.solve_plain(seed, **grid) for seed in TRAIN]
    rng = np.random.default_rng(0)
    xs, ys = zip(*(ws.sample_pairs(m, 16384, rng) for m in maps))
    print(f"[gpu] warm start {name}: training", flush=True)
    net, info = ws.train(np.concatenate(xs), np.concatenate(ys), maps[0].p, maps[0].e,
                         steps=spec["steps"], batch=4096)
    validation = ws.solve_plain(VALIDATE, **grid)
    delta = float(max(ws.underestimates(net, validation).max(), 0.0))
    print(f"[gpu] warm start {name}: loss {info['loss_first']:.3f} -> "
          f"{info['loss_last']:.3f}, delta {delta:.3f}", flush=True)
    p, e, tables = ws.game(HELD_OUT, **grid)
    with GameSolver(tables) as s:
        done, _ = s.run(1e-6, spec["vstar_cap"], 32, ws.FLAGS)
        vstar, residual = s.read(), float(s.trace()[-1, 0])
    print(f"[gpu] warm start {name}: held-out map {HELD_OUT}: V* after {done} sweeps, "
          f"residual {residual:.1e}", flush=True)
    inside = ws.dense_inside(p, e, ws.EXIT_RADIUS)
    yield dict(name=name, chunk=spec["chunk"], net=net, delta=delta, info=info, tables=tables,
               vstar=vstar, inside=inside, vstar_sweeps=done, vstar_residual=residual)
    net.close()


def test_training_lowers_the_loss(case):
    assert case["info"]["loss_last"] < 0.5 * case["info"]["loss_first"]
    assert np.isfinite(case["delta"]) and case["delta"] >= 0
    assert case["vstar_residual"] <= 1e-5


def test_warm_start_is_bounded_by_cold_and_judged_by_the_plain_dp(case):
    vstar, inside = case["vstar"], case["inside"]
    with GameSolver(case["tables"]) as cold, GameSolver(case["tables"]) as warm:
        phi = cold.read()
        warm.warm_start(case["net"], case["delta"])
        v0 = warm.read()
        assert (v0 <= phi).all()
        below = float(max((vstar - v0)[inside].max(), 0.0))
        first = {"cold": None, "warm": None}
        sweeps = 0
        # the cold iterate after vstar_sweeps sweeps is V* itself
        while sweeps < case["vstar_sweeps"] and first["cold"] is None:
            cold.sweep(case["chunk"], ws.FLAGS)
            warm.sweep(case["chunk"], ws.FLAGS)
            sweeps += case["chunk"]
            vc, vw = cold.read(), warm.read()
            assert (vw <= vc).all(), f"warm above cold after {sweeps} sweeps"
            assert (vw <= v0).all(), f"a warm value rose above its start after {sweeps} sweeps"
            for name, v in (("cold", vc), ("warm", vw)):
                if first[name] is None and np.abs(v - vstar)[inside].max() <= 1e-3:
                    first[name] = sweeps
    judge_ok = below <= 1e-3 and first["warm"] is not None
    print(f"[gpu] warm start {case['name']}, map {HELD_OUT}, delta {case['delta']:.3f}: start "
          f"below V* by up to {below:.3e} at {int(((vstar - v0)[inside] > 0).sum())} pairs; "
          f"sweeps to 1e-3 of V*: cold {first['cold']}, warm {first['warm']}; "
          f"judge {'ok' if judge_ok else 'FAIL (the plain answer is used)'}", flush=True)
    assert first["cold"] is not None
    if below == 0.0:
        # V* <= warm <= cold at every sweep
        assert first["warm"] is not None and first["warm"] <= first["cold"]
    if below > 1e-3:
        # a pair that started more than 1e-3 below V* never comes within
        # 1e-3 of it: the plain DP's judgement stands
        assert first["warm"] is None


def test_network_command_order_leaves_the_values(case):
    """pe_sl_order_from_mlp stores, for every inside pair, the commands of
    one backup evaluated on min(phi, network); the sweeps then try them
    first.  Early exit and the visiting order never change a value, so a
    cold run from phi gives the same fields and residuals bit for bit with
    the network's order as with none; the first sweep's E side is timed
    both ways."""
    flags = ws.FLAGS & ~_native.PE_SL_GRAPH
    with GameSolver(case["tables"]) as plain, GameSolver(case["tables"]) as ordered:
        ordered.order_from(case["net"])
        first = {}
        for name, s in (("index order", plain), ("network order", ordered)):
            s.set_profiling(True)
            s.sweep(1, flags)
            st = s.stats()
            first[name] = st["e_ms"] / max(1, st["e_launches"])
            s.set_profiling(False)
        plain.sweep(63, ws.FLAGS)
        ordered.sweep(63, ws.FLAGS)
        np.testing.assert_array_equal(plain.read().view(np.uint32),
                                      ordered.read().view(np.uint32))
        np.testing.assert_array_equal(plain.trace(), ordered.trace())
    print(f"[gpu] {case['name']}: first sweep's E side {first['index order']:.3f} ms in index "
          f"order, {first['network order']:.3f} ms in the network's order", flush=True)


def test_constructed_starts_around_vstar(case):
    """The judge's two facts on constructed starts (no network): a start
    between V* and phi moves towards V* and never rises; a start lowered
    below V* at one inside pair stays there, so the answer ends more than
    1e-3 from the plain DP's and the judge rejects it.  The small-map case
    only (the C2-grid case checks the same sweep)."""
    if case["name"] != "small":
        pytest.skip("the constructed starts run on the small maps")
    vstar, inside = case["vstar"], case["inside"]
    rng = np.random.default_rng(7)
    with GameSolver(case["tables"]) as s:
        phi = s.read()
        lift = np.float32(0.05) * rng.random(vstar.shape, np.float32)
        above = np.where(inside, np.minimum(phi, vstar + lift), phi).astype(np.float32)
        lowered = above.copy()
        cand = np.flatnonzero((inside & (vstar > -0.4)).ravel())
        at = int(cand[len(cand) // 2])
        lowered.ravel()[at] = vstar.ravel()[at] - np.float32(0.1)
        for name, v0 in (("between V* and phi", above), ("one pair 0.1 below V*", lowered)):
            s.initialize(v0)
            gap0 = float(np.abs(v0 - vstar)[inside].max())
            prev = v0
            for _ in range(8):
                s.sweep(case["chunk"], ws.FLAGS)
                v = s.read()
                assert (v <= prev).all(), f"{name}: a value rose"
                prev = v
            gap = float(np.abs(prev - vstar)[inside].max())
            print(f"[gpu] constructed start {name}: max |V - V*| {gap0:.3e} -> {gap:.3e} "
                  f"after {8 * case['chunk']} sweeps", flush=True)
            if name == "between V* and phi":
                # V* is the plain field at residual 1e-6: later sweeps may
                # still lower it by the remaining tail, far below 2e-3 here
                assert (prev >= vstar - 2e-3)[inside].all()
                assert gap <= gap0
            else:
                assert prev.ravel()[at] <= lowered.ravel()[at]
                assert gap >= 0.1 - 1e-6
