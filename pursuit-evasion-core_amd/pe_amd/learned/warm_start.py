"""Learned warm start of the contact-game value iteration.

Pipeline (all compute on the GPU):
  1. solve seeded random-rectangle maps with the plain DP (cold start from
     phi) to a small residual (1e-6 in the batched comparison, --label-tol
     on the command line): these FP32 value tables are the training labels;
  2. sample pairs inside the exit radius; the network inputs (8D state, phi,
     both clearances) are built on the device from the solver's tables;
  3. train the BF16-MFMA MLP (weighted MSE, optionally weighting
     underestimates more, AdamW) on the training maps;
  4. calibrate the lift delta on validation maps so that min(phi, net + delta)
     stays above their plain-DP values (largest underestimate over every
     inside pair);
  5. on held-out maps, compare warm and cold starts: sweeps and GPU time until
     the residual max|dV| <= 1e-4, and until max|V - V*| <= 1e-3, where V* is
     the plain DP's own field at a small residual (1e-6 for one map, 1e-7 in
     the batched comparison).  The plain DP is the judge: a
     warm-started answer that ends farther than 1e-3 from V* is reported as a
     failure (and the plain answer is the one to use).

Why a warm start cannot change the answer when it sits between the value
and the payoff: the sweep F(V) = min(V, T V) is monotone, V* = lim F^k(phi)
is its greatest fixed point below phi, so V* <= V0 <= phi gives
V* <= F^k(V0) <= F^k(phi) -> V*.  A start below V* anywhere stays below it
there (F never raises a value), which is why the judge is needed.  Control
ordering: the first warm sweep records each pair's argmin/argmax commands on
the network's field; later sweeps try them first and exit early.  Early exit
never changes a result.

Command line (one GPU):
    python -m pe_amd.learned.warm_start train --train 0-55 --validate 56-63 --out net.npz
    python -m pe_amd.learned.warm_start evaluate --net net.npz --test 64-79 --out c3.json
"""

from __future__ import annotations

import argparse
from dataclasses import dataclass
import json
import sys
import time

import numpy as np

from .. import _native
from ..local_game.grid_game import GameSolver, dense_inside
from ..local_game.reference import benchmark_game
from .mlp import ValueMLP

FLAGS = _native.PE_SL_EARLY_EXIT | _native.PE_SL_ORDERED | _native.PE_SL_GRAPH
EXIT_RADIUS = 3.0


@dataclass
class SolvedMap:
    seed: int
    p: object
    e: object
    tables: dict
    value: np.ndarray      # plain-DP field [p.n, e.n] at the stopping residual
    inside: np.ndarray     # [p.n, e.n] pairs updated by the backup
    sweeps: int
    seconds: float         # GPU time of the sweeps


def game(seed, **grid):
    """Grids and device tables of one seeded map, with each car's clearance
    (the network's map features)."""
    config, arena, p, e, tables = benchmark_game(seed, **grid)
    for side, g in (("p", p), ("e", e)):
        tables[side]["clearance"] = g.geometry.clearance(g.states)
    return p, e, tables


def solve_plain(seed, *, tol=1e-4, max_sweeps=60_000, **grid):
    """Cold-start plain DP to residual <= tol."""
    p, e, tables = game(seed, **grid)
    with GameSolver(tables) as s:
        done, ok = s.run(tol, max_sweeps, 32, FLAGS)
        seconds = s.stats()["update_seconds"]
        value = s.read()
    if not ok:
        raise RuntimeError(f"map {seed}: plain DP did not reach residual {tol}")
    return SolvedMap(seed, p, e, tables, value, dense_inside(p, e, EXIT_RADIUS), done, seconds)


def map_inputs(m: SolvedMap):
    """Network inputs of every pair of a map [p.n * e.n, 11]."""
    with GameSolver(m.tables) as s:
        return s.pair_inputs(0, m.inside.size)


def sample_pairs(m: SolvedMap, count, rng):
    """Network inputs and labels of `count` random inside pairs of a map."""
    flat = np.flatnonzero(m.inside.ravel())
    pick = np.sort(rng.choice(flat, size=min(count, len(flat)), replace=False))
    return map_inputs(m)[pick], m.value.ravel()[pick]


def train(x, y, p, e, *, steps=8000, batch=8192, seed=0, **net_kw):
    net = ValueMLP.for_game(p, e, seed=seed, **net_kw)
    losses = net.fit(x, y, steps=steps, batch=batch, seed=seed + 1)
    return net, dict(samples=len(x), steps=steps, batch=batch,
                     loss_first=float(losses[:20].mean()), loss_last=float(losses[-20:].mean()))


def underestimates(net, m: SolvedMap):
    """V - net over every inside pair of a solved map."""
    x = map_inputs(m)
    keep = m.inside.ravel()
    return m.value.ravel()[keep] - net(x[keep])


def run_to_targets(solver, target, *, tol_err=1e-3, tol_res=1e-4, chunk=64, max_sweeps=60_000,
                   stop_res=None):
    """Sweeps in chunks from the solver's current field, recording for the
    first chunk boundary at which max|V - target| <= tol_err and at which the
    residual <= tol_res the sweep count and the GPU seconds.  Stops when both
    are met (and the residual is <= stop_res, if given) or at max_sweeps."""
    sweeps, t_start = 0, solver.stats()["update_seconds"]
    first_err = first_res = None
    err = float("inf")
    while sweeps < max_sweeps:
        solver.sweep(chunk, FLAGS)
        sweeps += chunk
        gpu = solver.stats()["update_seconds"] - t_start
        res = float(solver.trace()[-1, 0])
        if first_err is None and target is not None:
            err = float(np.abs(solver.read() - target).max())
            if err <= tol_err:
                first_err = (sweeps, gpu)
        if first_res is None and res <= tol_res:
            first_res = (sweeps, gpu)
        done_err = target is None or first_err is not None
        if done_err and first_res is not None and (stop_res is None or res <= stop_res):
            break
    return first_err, first_res, sweeps, solver.read()


def evaluate(net, delta, seed, *, vstar_tol=1e-6, max_sweeps=60_000, log=print, **grid):
    """Cold vs warm on one held-out map.  V* = the cold run continued to
    residual vstar_tol; since the cold iterates decrease monotonically to V*,
    max|V_k - V*| is non-increasing and is read off after the run from
    snapshots every 64 sweeps."""
    p, e, tables = game(seed, **grid)
    inside = dense_inside(p, e, EXIT_RADIUS)
    out = dict(seed=seed, inside_pairs=int(inside.sum()))
    with GameSolver(tables) as s:
        snaps, cold = [], dict(res=None)
        t0 = s.stats()["update_seconds"]
        done = 0
        while done < max_sweeps:
            s.sweep(64, FLAGS)
            done += 64
            snaps.append((done, s.stats()["update_seconds"] - t0, s.read()))
            res = float(s.trace()[-1, 0])
            if cold["res"] is None and res <= 1e-4:
                cold["res"] = (done, snaps[-1][1])
            if res <= vstar_tol:
                break
        vstar = snaps[-1][2]
        out["vstar_sweeps"], out["vstar_residual"] = done, res
        cold_err = next(((k, t) for k, t, v in snaps if np.abs(v - vstar).max() <= 1e-3), None)
        del snaps
        log(f"[c3] map {seed}: cold to res 1e-4 {cold['res']}, to 1e-3 of V* {cold_err}, "
            f"V* at {done} sweeps (res {res:.1e})")
        s.initialize()
        g0 = s.stats()["update_seconds"]
        t0 = time.perf_counter()
        s.warm_start(net, delta)
        init_wall = time.perf_counter() - t0
        v0 = s.read()
        limit = cold_err[0] if cold_err else done
        warm_err, warm_res, warm_done, warm_final = run_to_targets(
            s, vstar, max_sweeps=max(limit, cold["res"][0] if cold["res"] else 0) + 64)
    below = (vstar - v0)[inside]
    out.update(
        cold_sweeps_to_res=cold["res"][0] if cold["res"] else None,
        cold_seconds_to_res=cold["res"][1] if cold["res"] else None,
        cold_sweeps_to_1e3=cold_err[0] if cold_err else None,
        cold_seconds_to_1e3=cold_err[1] if cold_err else None,
        warm_init_seconds=init_wall,
        warm_sweeps_to_res=warm_res[0] if warm_res else None,
        warm_seconds_to_res=(warm_res[1] + init_wall) if warm_res else None,
        warm_sweeps_to_1e3=warm_err[0] if warm_err else None,
        warm_seconds_to_1e3=(warm_err[1] + init_wall) if warm_err else None,
        warm_start_below_vstar_pairs=int((below > 0).sum()),
        warm_start_below_vstar_max=float(max(below.max(), 0.0)),
        warm_start_mean_gap=float((v0 - vstar)[inside].mean()),
        warm_final_vs_vstar=float(np.abs(warm_final - vstar).max()),
        warm_sweeps_run=warm_done,
    )
    out["judge_ok"] = out["warm_final_vs_vstar"] <= 1e-3
    log(f"[c3] map {seed}: warm to res 1e-4 {warm_res}, to 1e-3 of V* {warm_err}, "
        f"start below V* at {out['warm_start_below_vstar_pairs']} pairs "
        f"(max {out['warm_start_below_vstar_max']:.3f}), judge {'ok' if out['judge_ok'] else 'FAIL'}")
    return out


def _batch(seeds, grid):
    """Tables of several seeded maps (geometry of all maps in one launch per
    phase), with each car's clearance as the network's map features."""
    from ..local_game.reference import benchmark_games

    out = []
    for _, arena, p, e, tables in benchmark_games(seeds, **grid):
        for side, g in (("p", p), ("e", e)):
            tables[side]["clearance"] = g.geometry.clearance(g.states)
        out.append((p, e, tables))
    return out


def _run_batch(solver, tol, max_sweeps, every=16):
    """Plain sweeps of every map of a batched solver until each map's
    residual is <= tol (converged maps frozen); GPU seconds, sweeps, and
    each map's stopping sweep."""
    t0 = solver.stats()["update_seconds"]
    done, ok = solver.run(tol, max_sweeps, every, FLAGS | _native.PE_SL_FREEZE)
    checks = solver.trace()[every - 1::every]
    stops = [int(every * (np.argmax(checks[:, k] <= tol) + 1)) if (checks[:, k] <= tol).any()
             else None for k in range(checks.shape[1])]
    return solver.stats()["update_seconds"] - t0, done, ok, stops


def _stop_residuals(trace, stops):
    """Each map's residual at the sweep where a frozen run stopped sweeping
    it (the last sweep for a map that never converged)."""
    return np.array([trace[(stop or len(trace)) - 1, k] for k, stop in enumerate(stops)])


def converge_batch(solver, vstar, masks, *, tol_err=1e-3, tol_res=1e-4, cap=200_000,
                   wait=None):
    """Sweeps every map of a batched solver from its current field (no
    freezing) until each map flagged in `wait` (default: all) is within
    tol_err of its V* over its inside pairs and the last sweep's residual is
    <= tol_res on every map, or until `cap` sweeps.  The distance is checked every 64 sweeps, or every 1/32 of
    the sweeps done so far when that is more (a 3 % resolution), by reading
    the maps back.  Per map: (sweeps, GPU seconds) at the first check within
    tol_err, and at the first sweep whose residual is <= tol_res (seconds
    of the check that follows it); None where the target is not reached."""
    base = len(solver.trace())
    t0 = solver.stats()["update_seconds"]
    n = len(vstar)
    wait = [True] * n if wait is None else list(wait)
    first_err = [None] * n
    checks, sweeps = [], 0
    while sweeps < cap:
        step = min(max(64, sweeps // 32 // 64 * 64), cap - sweeps)
        solver.sweep(step, FLAGS)
        sweeps += step
        st = solver.stats()
        checks.append((sweeps, st["update_seconds"] - t0))
        for k in range(n):
            if first_err[k] is None:
                gap = np.abs(solver.read(k) - vstar[k])[masks[k]]
                if gap.size == 0 or gap.max() <= tol_err:
                    first_err[k] = checks[-1]
        if all(f is not None for f, w in zip(first_err, wait) if w) and \
                st["last_residual"] <= tol_res:
            break
    trace = solver.trace()[base:]
    first_res = []
    for k in range(n):
        hit = np.flatnonzero(trace[:, k] <= tol_res)
        if len(hit) == 0:
            first_res.append(None)
            continue
        at = int(hit[0]) + 1
        first_res.append((at, next(t for s, t in checks if s >= at)))
    return first_err, first_res, sweeps, checks[-1][1] if checks else 0.0


def _targets(first, over=None):
    """Per-map sweeps and seconds of converge_batch's records, and the
    batch's (the slowest map's) figures over the maps flagged in `over`."""
    over = [True] * len(first) if over is None else over
    sweeps = [f[0] if f else None for f in first]
    seconds = [f[1] if f else None for f in first]
    chosen = [f for f, o in zip(first, over) if o]
    done = bool(chosen) and all(f is not None for f in chosen)
    return dict(map_sweeps=sweeps, map_gpu_seconds=seconds,
                batch_sweeps=max(f[0] for f in chosen) if done else None,
                batch_gpu_seconds=max(f[1] for f in chosen) if done else None)


def check_batch(net, delta, seeds, grid, *, tol_res=1e-4, tol_err=1e-3, max_sweeps=20_000):
    """The plain DP as the judge of a whole batch: the maps of `seeds` in one
    handle, run to residual tol_res with converged maps frozen, once from phi
    and once from the warm start min(phi, net + delta); per map the largest
    difference of the two answers over its inside pairs (the judge wants at
    most tol_err), and each run's sweeps and GPU seconds."""
    held = _batch(list(seeds), grid)
    tables = [g[2] for g in held]
    masks = [dense_inside(p, e, EXIT_RADIUS) for p, e, _ in held]
    runs = {}
    for name in ("cold", "warm"):
        with GameSolver(tables) as s:
            if name == "warm":
                s.warm_start(net, delta)
            seconds, done, ok, stops = _run_batch(s, tol_res, max_sweeps)
            runs[name] = dict(fields=[s.read(k) for k in range(len(tables))], seconds=seconds,
                              sweeps=done, converged=ok, map_sweeps=stops)
    diff = [float(np.abs(w - c)[m].max()) if m.any() else 0.0
            for w, c, m in zip(runs["warm"]["fields"], runs["cold"]["fields"], masks)]
    seeds = list(seeds)
    out = dict(maps=len(tables), tolerance_residual=tol_res, tolerance_vs_plain=tol_err,
               max_abs_diff_vs_plain=max(diff), maps_within_tolerance=int(sum(d <= tol_err
                                                                            for d in diff)),
               maps_over_tolerance={str(sd): d for sd, d in zip(seeds, diff) if d > tol_err})
    for name, r in runs.items():
        out[name] = dict(sweeps=r["sweeps"], gpu_seconds=r["seconds"], converged=r["converged"],
                         map_sweeps_total=int(sum(k or r["sweeps"] for k in r["map_sweeps"])))
    out["judge_ok"] = out["maps_within_tolerance"] == len(tables)
    return out


def accuracy_bar(tables, vstar, masks, lifts=(1.0, 0.1, 0.01), *, tol_err=1e-3, tol_res=1e-4,
                 cap=100_000, seed=0):
    """What a warm start would have to achieve before it pays: on maps whose
    V* is known, the plain sweeps from starts min(phi, V* + lift * u), u
    uniform in [0, 1) per inside pair (so never below V*, and on average
    lift / 2 above it), until every map is within tol_err of V* and at
    residual tol_res.  Sweeps per map to both targets for each lift."""
    rng = np.random.default_rng(seed)
    with GameSolver(tables) as s:
        phi = [s.read(k) for k in range(len(tables))]
    rows = []
    for lift in lifts:
        v0 = [np.where(m, np.minimum(f, z + np.float32(lift) * rng.random(z.shape, np.float32)),
                       f).astype(np.float32) for f, z, m in zip(phi, vstar, masks)]
        with GameSolver(tables) as s:
            s.initialize(v0)
            e_first, r_first, _, _ = converge_batch(s, vstar, masks, tol_err=tol_err,
                                                    tol_res=tol_res, cap=cap)
        rows.append(dict(lift=lift, to_vstar=_targets(e_first), to_residual=_targets(r_first)))
    return rows


def compare_batched(train_seeds, validate_seeds, test_seeds, *, grid, label_tol=1e-6,
                    label_cap=20_000, tol_err=1e-3, tol_res=1e-4, vstar_tol=1e-7,
                    vstar_cap=100_000, steps=3000, batch=8192, samples=32768, seed=0,
                    under_weight=4.0, check_seeds=None, log=print):
    """The learned warm start against the plain DP on batches of maps.

    Labels: the plain DP to residual `label_tol` (or `label_cap` sweeps) on
    the training and validation maps, all in one handle with converged maps
    frozen.  The network (8D state, phi, both clearances) is trained on the
    GPU by the library's MFMA kernels; delta is the largest underestimate
    over the inside pairs of the validation maps.  On the held-out test
    maps, V* is the plain DP continued to residual `vstar_tol` (or
    `vstar_cap` sweeps; a map that does not get there has no V* and is
    left out of the batch figures to V*).  The plain DP from phi and the
    warm start min(phi, net + delta) are then each swept until every map
    with a V* is within tol_err of it and every map has residual <=
    tol_res: sweeps and GPU seconds of each map and of the batch to both
    targets.  The judge: the warm start is nowhere more than tol_err below
    V* (a value below V* stays below it), and the warm run reaches tol_err
    of V* on every map that has one.  With `check_seeds`, the maps of those
    seeds are then warm-started as one batch and judged against the plain
    DP run beside them (check_batch); last, the starts between V* and V* +
    lift on the held-out maps with a V* (accuracy_bar)."""
    out = dict(train_maps=len(train_seeds), validate_maps=len(validate_seeds),
               test_maps=len(test_seeds), tolerance_vs_vstar=tol_err,
               tolerance_residual=tol_res)
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    games = _batch(list(train_seeds) + list(validate_seeds), grid)
    n_train = len(train_seeds)
    with GameSolver([g[2] for g in games]) as s:
        label_s, label_done, ok, stops = _run_batch(s, label_tol, label_cap)
        label_res = _stop_residuals(s.trace(), stops)
        fields = [s.read(k) for k in range(len(games))]
        inputs = [s.pair_inputs(0, f.size, k) for k, f in enumerate(fields)]
    out.update(labels=dict(tolerance=label_tol, gpu_seconds=label_s, sweeps=label_done,
                           converged=bool(ok), max_residual=float(label_res.max())))
    xs, ys = [], []
    for (p, e, _), f, x in zip(games[:n_train], fields, inputs):
        inside = np.flatnonzero(dense_inside(p, e, EXIT_RADIUS).ravel())
        pick = np.sort(rng.choice(inside, size=min(samples, len(inside)), replace=False))
        xs.append(x[pick])
        ys.append(f.ravel()[pick])
    p0, e0, _ = games[0]
    t1 = time.perf_counter()
    net, info = train(np.concatenate(xs), np.concatenate(ys), p0, e0, steps=steps, batch=batch,
                      seed=seed, under_weight=under_weight)
    info["under_weight"] = under_weight
    out.update(training=info, train_wall_seconds=time.perf_counter() - t1)
    worst, rmse = 0.0, []
    for (p, e, _), f, x in zip(games[n_train:], fields[n_train:], inputs[n_train:]):
        keep = dense_inside(p, e, EXIT_RADIUS).ravel()
        miss = f.ravel()[keep] - net(x[keep])
        worst = max(worst, float(miss.max()))
        rmse.append(float(np.sqrt(np.mean(miss ** 2))))
    delta = max(0.0, worst)
    out.update(validation_rmse=rmse, delta=delta)
    log(f"[learned] labels {label_s:.1f} s GPU ({label_done} sweeps), training "
        f"{out['train_wall_seconds']:.1f} s, validation RMSE {np.round(rmse, 3).tolist()}, "
        f"delta {delta:.3f}")
    held = _batch(list(test_seeds), grid)
    tables = [g[2] for g in held]
    masks = [dense_inside(p, e, EXIT_RADIUS) for p, e, _ in held]
    with GameSolver(tables) as ref:
        vstar_s, vstar_done, vstar_ok, vstar_stops = _run_batch(ref, vstar_tol, vstar_cap,
                                                                 every=64)
        vstar_res = _stop_residuals(ref.trace(), vstar_stops)
        vstar = [ref.read(k) for k in range(len(held))]
    has_vstar = [stop is not None for stop in vstar_stops]
    log(f"[learned] V*: {vstar_done} sweeps ({vstar_s:.1f} s GPU), residual "
        f"max {vstar_res.max():.1e}; {sum(has_vstar)} of {len(held)} maps at {vstar_tol:g}")
    with GameSolver(tables) as cold:
        c_err, c_res, c_done, c_s = converge_batch(cold, vstar, masks, tol_err=tol_err,
                                                   tol_res=tol_res, cap=vstar_cap,
                                                   wait=has_vstar)
    cold_t = dict(to_vstar=_targets(c_err, has_vstar), to_residual=_targets(c_res),
                  sweeps_run=c_done)
    log(f"[learned] cold: batch to {tol_err:g} of V* {cold_t['to_vstar']['batch_sweeps']} "
        f"sweeps, to residual {tol_res:g} {cold_t['to_residual']['batch_sweeps']} sweeps")
    limit = cold_t["to_vstar"]["batch_sweeps"] or vstar_cap
    with GameSolver(tables) as warm:
        w0 = time.perf_counter()
        warm.warm_start(net, delta)
        start_s = time.perf_counter() - w0
        v0 = [warm.read(k) for k in range(len(held))]
        w_err, w_res, w_done, w_s = converge_batch(warm, vstar, masks, tol_err=tol_err,
                                                   tol_res=tol_res,
                                                   cap=min(vstar_cap, limit + limit // 4 + 1024),
                                                   wait=has_vstar)
    warm_t = dict(to_vstar=_targets(w_err, has_vstar), to_residual=_targets(w_res),
                  sweeps_run=w_done, start_seconds=start_s)
    # measured only where V* is known: the capped iterate of a map that did
    # not converge lies above V*, so a start below it proves nothing
    below = [float(max((v - z)[keep].max(), 0.0)) if ok else None
             for z, v, keep, ok in zip(v0, vstar, masks, has_vstar)]
    worst_below = max([b for b in below if b is not None], default=0.0)
    warm_t.update(start_below_vstar_max=below,
                  start_mean_gap_to_vstar=[float((z - v)[keep].mean())
                                           for z, v, keep in zip(v0, vstar, masks)])

    # The network's command order for a cold run: the first sweep's E side
    # in the index order (nothing stored yet) and in the network's order,
    # and the cold run to residual `tol_res` from the network's order.
    order = {}
    with GameSolver(tables) as s:
        s.set_profiling(True)
        s.sweep(1, FLAGS & ~_native.PE_SL_GRAPH)
        st = s.stats()
        order["first_sweep_e_ms_index_order"] = st["e_ms"] / max(1, st["e_launches"])
        s.set_profiling(False)
        p_s, p_done, _, p_stops = _run_batch(s, tol_res, vstar_cap)
        order.update(plain_sweeps_to_residual=p_done + 1, plain_gpu_seconds_to_residual=p_s,
                     plain_map_sweeps=[None if k is None else k + 1 for k in p_stops])
    with GameSolver(tables) as s:
        o0 = time.perf_counter()
        s.order_from(net)
        order["order_seconds"] = time.perf_counter() - o0
        s.set_profiling(True)
        s.sweep(1, FLAGS & ~_native.PE_SL_GRAPH)
        st = s.stats()
        order["first_sweep_e_ms_network_order"] = st["e_ms"] / max(1, st["e_launches"])
        s.set_profiling(False)
        o_s, o_done, o_ok, o_stops = _run_batch(s, tol_res, vstar_cap)
        order.update(sweeps_to_residual=o_done + 1, gpu_seconds_to_residual=o_s,
                     map_sweeps=[None if k is None else k + 1 for k in o_stops])
    log(f"[learned] command order: first E side {order['first_sweep_e_ms_index_order']:.3f} ms "
        f"(index order) / {order['first_sweep_e_ms_network_order']:.3f} ms (network order)")

    def ratio(key):
        c, w = cold_t[key]["batch_gpu_seconds"], warm_t[key]["batch_gpu_seconds"]
        return c / (w + start_s) if c is not None and w is not None else None

    out.update(
        vstar=dict(tolerance=vstar_tol, cap=vstar_cap, sweeps=vstar_done, converged=vstar_ok,
                   gpu_seconds=vstar_s, map_sweeps=vstar_stops, maps_with_vstar=has_vstar,
                   residual=[float(r) for r in vstar_res]),
        cold=cold_t, warm=warm_t, network_command_order=order,
        # The plain DP judges the warm start: it must not start more than
        # tol_err below V* anywhere (such a value is never raised again),
        # and the warm run must come within tol_err of V* on every map.
        judge_ok=bool(worst_below <= tol_err and warm_t["to_vstar"]["batch_sweeps"] is not None),
        speedup_to_vstar=ratio("to_vstar"), speedup_to_residual=ratio("to_residual"),
        total_wall_seconds=time.perf_counter() - t0)
    log(f"[learned] warm: batch to {tol_err:g} of V* {warm_t['to_vstar']['batch_sweeps']} "
        f"sweeps, to residual {tol_res:g} {warm_t['to_residual']['batch_sweeps']} sweeps; "
        f"start below V* by up to {worst_below:.3f}; judge {out['judge_ok']}")
    if check_seeds is not None:
        chk = out["batch_checked_by_plain_dp"] = check_batch(net, delta, check_seeds, grid,
                                                              tol_res=tol_res, tol_err=tol_err)
        chk["seeds"] = f"{min(check_seeds)}-{max(check_seeds)}"
        log(f"[learned] batch of {chk['maps']} maps, warm and cold to residual {tol_res:g}: "
            f"max |warm - plain| {chk['max_abs_diff_vs_plain']:.2e}, "
            f"{chk['maps_within_tolerance']} maps within {tol_err:g}; map-sweeps "
            f"{chk['warm']['map_sweeps_total']} warm, {chk['cold']['map_sweeps_total']} cold; "
            f"GPU {chk['warm']['gpu_seconds']:.2f} s warm, {chk['cold']['gpu_seconds']:.2f} s cold")
    net.close()
    # The bar a network would have to clear: starts between V* and V* +
    # lift on the maps that have a V*, against the cold runs above.
    known = [k for k, ok in enumerate(has_vstar) if ok]
    if known:
        bar = accuracy_bar([tables[k] for k in known], [vstar[k] for k in known],
                           [masks[k] for k in known], tol_err=tol_err, tol_res=tol_res,
                           cap=vstar_cap, seed=seed)
        out["start_accuracy_bar"] = dict(
            maps=[int(list(test_seeds)[k]) for k in known],
            cold_to_vstar=[cold_t["to_vstar"]["map_sweeps"][k] for k in known],
            cold_to_residual=[cold_t["to_residual"]["map_sweeps"][k] for k in known],
            starts=bar)
        for row in bar:
            log(f"[learned] start within {row['lift']:g} above V*: to {tol_err:g} of V* "
                f"{row['to_vstar']['map_sweeps']} sweeps, to residual {tol_res:g} "
                f"{row['to_residual']['map_sweeps']} (cold {out['start_accuracy_bar']['cold_to_vstar']}, "
                f"{out['start_accuracy_bar']['cold_to_residual']})")
    return out


def seeds(text):
    lo, _, hi = text.partition("-")
    return list(range(int(lo), int(hi or lo) + 1))


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def cmd_train(args):
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    xs, ys, label = [], [], dict(sweeps=[], seconds=[])
    first = None
    for k in seeds(args.train):
        m = solve_plain(k, tol=args.label_tol)
        first = first or m
        x, y = sample_pairs(m, args.samples, rng)
        xs.append(x)
        ys.append(y)
        label["sweeps"].append(m.sweeps)
        label["seconds"].append(m.seconds)
        _log(f"[c3] label map {k}: {m.sweeps} sweeps, {m.seconds:.2f} s GPU "
             f"({time.perf_counter() - t0:.0f} s)")
    x, y = np.concatenate(xs), np.concatenate(ys)
    t1 = time.perf_counter()
    net, info = train(x, y, first.p, first.e, steps=args.steps, batch=args.batch,
                      seed=args.seed, under_weight=args.under_weight)
    info["train_seconds"] = time.perf_counter() - t1
    _log(f"[c3] trained: {json.dumps(info)}")
    under, val = [], []
    for k in seeds(args.validate):
        m = solve_plain(k, tol=args.label_tol)
        u = underestimates(net, m)
        under.append(u)
        val.append(dict(seed=k, rmse=float(np.sqrt(np.mean(u ** 2))), max_under=float(u.max()),
                        q999=float(np.quantile(u, 0.999))))
        _log(f"[c3] validation map {k}: {json.dumps(val[-1])}")
    under = np.concatenate(under)
    delta = float(max(0.0, np.quantile(under, args.quantile)))
    np.savez(args.out, params=net.get_params(), delta=delta,
             shift=np.array(net.config.state_shift[:]), scale=np.array(net.config.state_scale[:]),
             meta=json.dumps(dict(train=args.train, validate=args.validate, labels=label,
                                  training=info, validation=val, quantile=args.quantile,
                                  delta=delta, under_weight=args.under_weight, seed=args.seed)))
    _log(f"[c3] delta {delta:.4f} (quantile {args.quantile}); saved {args.out}")


def load_net(path):
    z = np.load(path)
    meta = json.loads(str(z["meta"]))
    net = ValueMLP(shift=z["shift"], scale=z["scale"], seed=meta["seed"],
                   under_weight=meta["under_weight"])
    net.set_params(z["params"])
    return net, float(z["delta"]), meta


def cmd_evaluate(args):
    net, delta, meta = load_net(args.net)
    if args.delta is not None:
        delta = args.delta
    rows = [evaluate(net, delta, k, log=_log) for k in seeds(args.test)]
    summary = dict(delta=delta, net=meta, maps=rows)
    with open(args.out, "w") as f:
        json.dump(summary, f, indent=1)
    _log(f"[c3] wrote {args.out}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--train", default="0-55")
    t.add_argument("--validate", default="56-63")
    t.add_argument("--label-tol", type=float, default=1e-4)
    t.add_argument("--samples", type=int, default=65536)
    t.add_argument("--steps", type=int, default=8000)
    t.add_argument("--batch", type=int, default=8192)
    t.add_argument("--under-weight", type=float, default=4.0)
    t.add_argument("--quantile", type=float, default=1.0)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", required=True)
    ev = sub.add_parser("evaluate")
    ev.add_argument("--net", required=True)
    ev.add_argument("--test", default="64-79")
    ev.add_argument("--delta", type=float, default=None)
    ev.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    (cmd_train if args.cmd == "train" else cmd_evaluate)(args)


if __name__ == "__main__":
    main()
