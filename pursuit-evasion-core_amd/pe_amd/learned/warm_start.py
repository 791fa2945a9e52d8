"""Learned warm start of the contact-game value iteration.

Pipeline (all compute on the GPU):
  1. solve seeded random-rectangle maps with the plain DP (cold start from
     phi) to residual 1e-4: these FP32 value tables are the training labels;
  2. sample pairs inside the exit radius; the network inputs (8D state, phi,
     both clearances) are built on the device from the solver's tables;
  3. train the BF16-MFMA MLP (weighted MSE, optionally weighting
     underestimates more, AdamW) on the training maps;
  4. calibrate the lift delta on validation maps so that min(phi, net + delta)
     stays above their plain-DP values (largest underestimate over every
     inside pair);
  5. on held-out maps, compare warm and cold starts: sweeps and GPU time until
     the residual max|dV| <= 1e-4, and until max|V - V*| <= 1e-3, where V* is
     the plain DP's own field at residual 1e-6.  The plain DP is the judge: a
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


def compare_batched(train_seeds, validate_seeds, test_seeds, *, grid, tol=1e-4,
                    max_sweeps=20_000, vstar_tol=1e-6, vstar_cap=30_000,
                    steps=3000, batch=8192, samples=32768, seed=0, log=print):
    """The learned warm start against the plain DP on batches of maps.

    Labels: the plain DP to residual `tol` on the training maps (all maps in
    one handle).  The network (8D state, phi, both clearances) is trained
    on the GPU by the library's MFMA kernels; delta is the largest
    underestimate over the inside pairs of the validation maps.  On the
    held-out test maps: the plain DP from phi and the warm start
    min(phi, net + delta), each run to residual `tol` with converged maps
    frozen; GPU seconds of both.  V* is the plain DP continued to residual
    vstar_tol (or vstar_cap sweeps).  The judge: the warm answer is as close
    to V* as the plain answer (or within 1e-3), and the warm start is
    nowhere more than 1e-3 below V* (a value below V* stays below it)."""
    out = dict(train_maps=len(train_seeds), validate_maps=len(validate_seeds),
               test_maps=len(test_seeds), tolerance=tol)
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    games = _batch(list(train_seeds) + list(validate_seeds), grid)
    n_train = len(train_seeds)
    with GameSolver([g[2] for g in games]) as s:
        label_s, _, ok, _ = _run_batch(s, tol, max_sweeps)
        fields = [s.read(k) for k in range(len(games))]
        inputs = [s.pair_inputs(0, f.size, k) for k, f in enumerate(fields)]
    out.update(label_gpu_seconds=label_s, labels_converged=bool(ok))
    xs, ys = [], []
    for (p, e, _), f, x in zip(games[:n_train], fields, inputs):
        inside = np.flatnonzero(dense_inside(p, e, EXIT_RADIUS).ravel())
        pick = np.sort(rng.choice(inside, size=min(samples, len(inside)), replace=False))
        xs.append(x[pick])
        ys.append(f.ravel()[pick])
    p0, e0, _ = games[0]
    t1 = time.perf_counter()
    net, info = train(np.concatenate(xs), np.concatenate(ys), p0, e0, steps=steps, batch=batch,
                      seed=seed, under_weight=4.0)
    out.update(training=info, train_wall_seconds=time.perf_counter() - t1)
    worst, rmse = 0.0, []
    for (p, e, _), f, x in zip(games[n_train:], fields[n_train:], inputs[n_train:]):
        keep = dense_inside(p, e, EXIT_RADIUS).ravel()
        miss = f.ravel()[keep] - net(x[keep])
        worst = max(worst, float(miss.max()))
        rmse.append(float(np.sqrt(np.mean(miss ** 2))))
    delta = max(0.0, worst)
    out.update(validation_rmse=rmse, delta=delta)
    log(f"[learned] labels {label_s:.1f} s GPU, training {out['train_wall_seconds']:.1f} s, "
        f"validation RMSE {np.round(rmse, 3).tolist()}, delta {delta:.3f}")
    held = _batch(list(test_seeds), grid)
    tables = [g[2] for g in held]
    masks = [dense_inside(p, e, EXIT_RADIUS) for p, e, _ in held]
    with GameSolver(tables) as ref:
        # V*: the plain DP continued to residual vstar_tol (or the cap)
        _, vstar_done, vstar_ok, _ = _run_batch(ref, vstar_tol, vstar_cap)
        vstar = [ref.read(k) for k in range(len(held))]
    with GameSolver(tables) as cold:
        cold_s, cold_done, cold_ok, cold_stops = _run_batch(cold, tol, max_sweeps)
        cold_fields = [cold.read(k) for k in range(len(held))]
    with GameSolver(tables) as warm:
        w0 = time.perf_counter()
        warm.warm_start(net, delta)
        start_s = time.perf_counter() - w0
        v0 = [warm.read(k) for k in range(len(held))]
        warm_s, warm_done, warm_ok, warm_stops = _run_batch(warm, tol, max_sweeps)
        warm_fields = [warm.read(k) for k in range(len(held))]

    def worst(fields, sign=None):
        out = []
        for f, v, keep in zip(fields, vstar, masks):
            d = (f - v)[keep] if sign is None else sign * (f - v)[keep]
            out.append(float(np.abs(d).max() if sign is None else max(d.max(), 0.0)))
        return out

    cold_err, warm_err = worst(cold_fields), worst(warm_fields)
    below = worst(v0, -1.0)
    out.update(
        vstar=dict(tolerance=vstar_tol, sweeps=vstar_done, converged=vstar_ok),
        cold=dict(gpu_seconds=cold_s, sweeps=cold_done, converged=cold_ok,
                  map_sweeps=cold_stops, max_abs_error_vs_vstar=cold_err),
        warm=dict(gpu_seconds=warm_s + start_s, start_seconds=start_s, sweeps=warm_done,
                  converged=warm_ok, map_sweeps=warm_stops, max_abs_error_vs_vstar=warm_err,
                  start_below_vstar_max=below,
                  start_mean_gap_to_vstar=[float((z - v)[keep].mean())
                                           for z, v, keep in zip(v0, vstar, masks)]),
        # The plain DP judges the warm answer: it must be as close to V* as
        # the plain answer at the same residual (or within 1e-3), and the
        # warm start must not lock in an underestimate beyond 1e-3.
        judge_ok=bool(all(w <= max(c, 1e-3) for w, c in zip(warm_err, cold_err))
                      and max(below) <= 1e-3),
        speedup=cold_s / (warm_s + start_s) if warm_s + start_s > 0 else None,
        total_wall_seconds=time.perf_counter() - t0)
    log(f"[learned] held-out: cold {cold_s:.2f} s / {cold_done} sweeps, warm "
        f"{warm_s + start_s:.2f} s / {warm_done} sweeps; error vs V*: cold "
        f"{max(cold_err):.2e}, warm {max(warm_err):.2e}; judge {out['judge_ok']}")
    net.close()
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
