"""Benchmark of the game value-function solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the configuration the headline metric is quoted on): the 5D
free-space reversible-car game, RelativeGrid(65, 36, 15, 15, 8.0) = 34,222,500
nodes, A3 limits (configs/reversible_free_a3.json) with capture radius 0.5 m,
ENO2 + local Lax-Friedrichs + RK2, FP32 local arithmetic, CFL 0.8.  A step is
one full RK2 step of the whole grid; the default K = 810 steps is the whole
3.0 s horizon.  The field is resident in HBM before the timed region; the
timed region is exactly K steps bracketed by device synchronisation.

Prints ONE JSON line: node-updates/s (whole job), the roofline of the
dominant kernel (the RK2 blend stage) from per-launch hipEvents recorded on
the library's own stream, and the oracle CPU baseline timed on a bounded
sample of the same workload on this host's cores.
"""

from __future__ import annotations

import argparse
import json
import os
from pathlib import Path
import sys
import time

ROOT = Path(__file__).resolve().parent
for p in (ROOT / "pursuit-evasion-core_amd", ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

# OpenMP placement of the CPU baselines, set before the oracle's runtime
# starts: one thread per physical core, packed.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")


_HOST_CORES = None


def host_cores():
    """(threads for the CPU baselines, physical cores this process may run
    on, logical CPUs).  The threads are the physical cores, capped by
    OMP_NUM_THREADS when the environment sets it (a shared host may grant
    each GPU a share of its cores that way).  Read once, at the first call:
    with OMP_PROC_BIND=close the first OpenMP region binds the calling
    thread to one core, and the affinity mask read after it has one CPU."""
    global _HOST_CORES
    if _HOST_CORES is None:
        _HOST_CORES = _read_host_cores()
    return _HOST_CORES


def _read_host_cores():
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else \
        list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            cores.add((base.joinpath("physical_package_id").read_text().strip(),
                       base.joinpath("core_id").read_text().strip()))
        except OSError:
            cores.add(("cpu", str(c)))
    physical = len(cores)
    threads = physical
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        threads = min(threads, int(cap))
    return threads, physical, len(cpus)


def cores_note():
    threads, physical, logical = host_cores()
    return (f"{threads} OpenMP threads (OMP_PROC_BIND=close, OMP_PLACES=cores) on a host "
            f"share of {physical} physical cores / {logical} logical CPUs"
            + ("" if threads == physical else
               f"; capped at OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')}"))

METRIC = ("node-updates/s + time-to-converge at max|ΔV|≤1e-3: "
          "MLP warm-start vs plain DP")
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per node of each RK2 stage launch:
#   stage 1: read a, write b                      = 8 B
#   stage 2: read b, read a (previous), write a   = 12 B
STAGE_BYTES = (8, 12)


def c1_config():
    from pe_amd.reversible.free_dp import RelativeGrid, native_config
    from pe_amd.reversible.model import ReversibleConfig, VehicleLimits

    config = ReversibleConfig(
        name="reversible_free_a3",
        pursuer=VehicleLimits(-1.0, 2.0, 3.0, 1.0),
        evader=VehicleLimits(-0.5, 1.0, 3.0, 2.0),
        capture_radius=0.5,
    )
    grid = RelativeGrid(65, 36, 15, 15, 8.0)
    return config, grid, native_config(config, grid, order=2, cfl=0.8)


def load_traffic():
    """Memory-side bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary: FETCH_SIZE and WRITE_SIZE scaled by a copy
    kernel's reading (bytes the L2 requested from the memory fabric; hits
    in the memory-side cache are not separable by any listed counter).
    Returns (bytes, note) or (None, None)."""
    path = ROOT / "profiles" / "hji5d_blend_stage_c1_traffic.json"
    if not path.exists():
        return None, None
    try:
        data = json.loads(path.read_text())
        return float(data["hbm_bytes_per_launch"]), data.get("traffic_kind")
    except Exception:
        return None, None


def cpu_baseline(cfg_native, sample_steps: int):
    """The oracle (C/OpenMP restatement, FP64 local arithmetic) on this host."""
    import oracle

    threads = host_cores()[0]
    ocfg = oracle.HJI5DCfg(
        cfg_native.positions, cfg_native.headings, cfg_native.p_speeds,
        cfg_native.e_speeds, cfg_native.order, threads, 0, cfg_native.extent,
        cfg_native.cfl, cfg_native.radius, cfg_native.p_min, cfg_native.p_max,
        cfg_native.e_min, cfg_native.e_max, cfg_native.p_acceleration,
        cfg_native.e_acceleration, cfg_native.p_curvature, cfg_native.e_curvature,
    )
    field, arrival = oracle.hji5d_init(ocfg)
    dt = oracle.hji5d_cfl_limit(ocfg)
    seconds = oracle.hji5d_time_steps(ocfg, field, arrival, dt, sample_steps)
    n = field.size
    return {
        "value": n * sample_steps / seconds,
        "unit": "node-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample_steps} RK2 steps of the same 34,222,500-node grid from phi "
                  f"({seconds:.2f} s, FP64 local arithmetic, float storage, OpenMP)",
        "host": cores_note(),
    }


def world8d_cpu_baseline(grid_dims, sample_steps: int):
    """cpu_baseline leg of the dense 8D line: the oracle (C/OpenMP, bit-exact
    with the reference executable) timed for a few RK2 steps of the grid."""
    import oracle

    threads = host_cores()[0]
    n, h, np_, ne = grid_dims
    ocfg = oracle.World8DCfg(n, h, np_, ne, 2, 1, threads, 3.5, 3.0, 0.3, -1.0, 2.0, -0.5, 1.0,
                             2.0, 1.0, 1.0, 2.0, 0.8)
    dt = oracle.world8d_cfl_limit(ocfg)
    field, arrival = oracle.world8d_init(ocfg)
    t0 = time.perf_counter()
    oracle.world8d_solve(ocfg, sample_steps * dt, sample_steps * dt, field=field,
                         arrival=arrival, history=False)
    seconds = time.perf_counter() - t0
    return {"value": field.size * sample_steps / seconds, "unit": "node-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"{sample_steps} RK2 steps ({seconds:.2f} s, float arithmetic, OpenMP)",
            "host": cores_note()}


def run_world8d(grid_dims, steps: int, warmup: int, cpu_steps: int = 0):
    """Dense 8D square game (default ReversibleConfig, RK2, linear speed
    boundary, CFL 0.8): CFL-limited steps of the whole grid."""
    from pe_amd import _native
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import WorldGrid, native_config

    grid = WorldGrid(*grid_dims)
    with _native.World8D(native_config(ReversibleConfig(), grid, order=2)) as w:
        dt = w.stats()["max_internal_dt"]
        w.advance_to(warmup * dt)
        s0 = w.stats()
        w.advance_to((warmup + steps) * dt)
        s1 = w.stats()
    seconds = s1["update_seconds"] - s0["update_seconds"]
    taken = s1["internal_steps"] - s0["internal_steps"]
    out = {"grid": list(grid_dims), "nodes": grid.states, "steps": taken,
           "ms_per_step": seconds / taken * 1e3,
           "node_updates_per_s": grid.states * taken / seconds,
           "stage_bytes_per_node": list(STAGE_BYTES),
           "hbm_fraction_algorithmic": grid.states * sum(STAGE_BYTES) * taken / seconds
                                       / (HBM_PEAK_GBS * 1e9)}
    if cpu_steps:
        out["cpu_baseline"] = world8d_cpu_baseline(grid_dims, cpu_steps)
    return out


def run_reader_c1(cfg, n_queries: int = 1_000_000, cpu_queries: int = 20_000):
    """5D table reader on the C1 history: the 31 frames (0, 0.1, ..., 3.0 s)
    of the C1 solve resident on the device, then batches of seeded queries
    [x, y, psi, vP, vE] over the grid (field at seeded budgets; first-crossing
    value).  Rates include the state upload and the result download."""
    import numpy as np
    from pe_amd import _native

    config, grid, _ = c1_config()
    axes = grid.axes(config)
    times = np.round(np.arange(31) * 0.1, 10)
    with _native.HJI5D(cfg) as h:
        hist = np.empty((31, grid.states), np.float32)
        h.read(hist[0], False)
        for k in range(1, 31):
            h.advance_to(times[k])
            h.read(hist[k], False)
    t0 = time.perf_counter()
    table = _native.RelTable(hist, axes, times)
    upload = time.perf_counter() - t0
    rng = np.random.default_rng(7)
    states = np.column_stack([rng.uniform(-8, 8, n_queries), rng.uniform(-8, 8, n_queries),
                             rng.uniform(-np.pi, np.pi, n_queries),
                             rng.uniform(-1, 2, n_queries), rng.uniform(-0.5, 1, n_queries)])
    budgets = rng.uniform(0, 3, n_queries)
    p, e = config.pursuer, config.evader
    limits = [p.speed_min, p.speed_max, e.speed_min, e.speed_max]
    table.field(states[:1000], budgets[:1000])
    t0 = time.perf_counter()
    table.field(states, budgets)
    field_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    values = table.value(states, config.capture_radius, limits)
    value_s = time.perf_counter() - t0
    table.close()
    out = {"history": "C1, 31 frames x 34,222,500 nodes (4.24 GB)", "queries": n_queries,
           "upload_s": upload, "field_queries_per_s": n_queries / field_s,
           "value_queries_per_s": n_queries / value_s,
           "resolved_fraction": float(np.isfinite(values).mean())}
    if cpu_queries:
        # cpu_baseline leg: the reader's numpy formulas (oracle restatement),
        # one thread, on a bounded sample of the same queries.
        import rtable_oracle

        t0 = time.perf_counter()
        for k in range(0, cpu_queries, 250_000):  # bounded host memory
            end = min(cpu_queries, k + 250_000)
            rtable_oracle.field_np(hist, times, axes, states[k:end], budgets[k:end])
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_queries / cpu_s, "unit": "field queries/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{cpu_queries} field queries, numpy ({cpu_s:.2f} s)"}
    return out


def sl_algorithmic_bytes(pairs, cars):
    """Algorithmic HBM bytes of one 8D sweep: each stored pair's value read
    and written (8 B) plus both cars' successor tables read once (int32 ids
    and float32 weights, 16 per state and command)."""
    tables = sum(len(g.states) * g.allowed.shape[1] * 16 * 8 for g in cars)
    return 8 * pairs + tables


def load_counters(name):
    """A committed rocprofv3 counter summary under profiles/ (JSON), or None."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    try:
        return json.loads(path.read_text())
    except Exception:
        return None


def sl_windows(s, sweeps, variants):
    """Per-launch P-side and E-side kernel times (ms) of `sweeps` profiled
    sweeps under each variant's flags, in order: {name: (p_ms, e_ms)}; a
    variant with no P-side launch (the fused sweep) reports p_ms 0."""
    s.set_profiling(True)
    out = {}
    for name, fl in variants.items():
        b = s.stats()
        s.sweep(sweeps, fl)
        a = s.stats()
        out[name] = ((a["p_ms"] - b["p_ms"]) / max(1, a["p_launches"] - b["p_launches"]),
                     (a["e_ms"] - b["e_ms"]) / max(1, a["e_launches"] - b["e_launches"]))
    s.set_profiling(False)
    return out


def sl_roofline(win, st0, sweep_seconds, cars, counters):
    """Roofline object of an 8D sweep line: HBM against the algorithmic
    bytes of the whole sweep (live timing), the sweep kernel's time, and
    the E side's VALU / LDS utilisation from the committed counter
    summary."""
    nbytes = sl_algorithmic_bytes(st0["pairs"], cars)
    achieved = nbytes / sweep_seconds / 1e9
    out = {"bound": "valu",
           "limiter": "dependent LDS gathers and FMAs of the E-side min-max (lane "
                      "divergence from the early exit); HBM is not the bound",
           "hbm": {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(achieved / HBM_PEAK_GBS, 4),
                   "algorithmic_bytes_per_sweep": nbytes},
           "fused_sweep_kernel_ms": win["fused"][1],
           "separate_p_side_ms": win["separate"][0],
           "separate_e_side_ms": win["separate"][1]}
    if counters:
        out["counters"] = counters
    return out


def c2_cpu_baseline(p, e, tables, v0, sweeps: int):
    """cpu_baseline leg of the C2 line: the oracle's FP64 OpenMP Jacobi
    sweep (the restatement of the reference's backup) timed for a few sweeps
    of the same map from the same initial field."""
    import numpy as np
    import oracle
    from pe_amd.local_game.grid_game import dense_inside

    threads = host_cores()[0]
    inside = dense_inside(p, e, tables["exit_value"])
    V = np.asarray(v0, np.float64)
    t0 = time.perf_counter()
    for _ in range(sweeps):
        V, _ = oracle.sl_sweep(V, p, e, inside, threads=threads)
    seconds = time.perf_counter() - t0
    legal = len(p.states) * len(e.states)
    return {"value": legal * sweeps / seconds, "unit": "legal-pair updates/s",
            "cores": threads, "kind": "port",
            "sample": f"{sweeps} FP64 Jacobi sweeps of the same map from the payoff "
                      f"({seconds:.2f} s, OpenMP)",
            "host": cores_note()}


def run_c2(args):
    """8D contact game C2: sweeps of the compact min-max backup."""
    import numpy as np
    from pe_amd import _native
    from pe_amd.local_game.grid_game import GameSolver
    from pe_amd.local_game.reference import benchmark_game

    t0 = time.perf_counter()
    config, arena, p, e, tables = benchmark_game(args.seed)
    build_s = time.perf_counter() - t0
    flags = _native.PE_SL_EARLY_EXIT | _native.PE_SL_ORDERED | _native.PE_SL_GRAPH
    with GameSolver(tables) as s:
        st0 = s.stats()
        v0 = s.read()
        if args.warmup:
            s.sweep(args.warmup, flags)
        t1 = time.perf_counter()
        s.sweep(args.steps, flags)
        wall = time.perf_counter() - t1
        base = flags & ~_native.PE_SL_GRAPH
        win = sl_windows(s, min(args.steps, 20), {
            "fused": base,
            "separate": base | _native.PE_SL_SEPARATE_P_SIDE,
            "stencil": base | _native.PE_SL_STENCIL_P_SIDE,
            "global": base | _native.PE_SL_GLOBAL_GATHER,
            "wide": base | _native.PE_SL_WIDE_IDS,
            "reference": _native.PE_SL_EARLY_EXIT})
        out = {
            "workload": "C2 8D contact game, 11x11 positions (+-3 m) x 12 headings x 5 speeds "
                        "per car, 30 random rectangles (seed %d)" % args.seed,
            "legal_pairs": st0["pairs"], "active_pairs": st0["active_pairs"],
            "dense_pairs": (11 * 11 * 12 * 5) ** 2,
            "table_build_s": build_s, "sweeps": args.steps, "ms_per_sweep": wall / args.steps * 1e3,
            "legal_pair_updates_per_s": st0["pairs"] * args.steps / wall,
            "active_pair_updates_per_s": st0["active_pairs"] * args.steps / wall,
            "e_side_lds_threads": st0["lds_e_side"],
            # The default sweep is one kernel: each E-side block computes its
            # P state's values into LDS from V (no G buffer).
            "fused_sweep_kernel_ms": win["fused"][1],
            # The same sweep with the P side as its own launch writing G.
            "p_side_ms": win["separate"][0], "e_side_ms": win["separate"][1],
            # That P side summing each command's stencil in vertex order
            # instead of adding each union row to all commands' sums.
            "p_side_stencil_ms": win["stencil"][0],
            "e_side_global_gather_ms": win["global"][1],
            "e_side_unpacked_tables_ms": win["wide"][1],
            # The reference's early exit in index order, without the stored
            # command order (separate P side).
            "e_side_reference_order_ms": win["reference"][1],
            "residual_after": s.stats()["last_residual"],
        }
        out["roofline"] = sl_roofline(win, st0, wall / args.steps, (p, e),
                                      load_counters("sl8d_e_side_c2_counters.json"))
        if args.to_residual:
            s.initialize()
            t2 = time.perf_counter()
            done, ok = s.run(args.to_residual, args.max_sweeps, 16, flags)
            out.update(to_residual=args.to_residual, sweeps_to_residual=done, converged=ok,
                       seconds_to_residual=time.perf_counter() - t2)
            v32 = s.read()
            # Accuracy at that point: the FP64 mode (the oracle's arithmetic)
            # after the same sweeps.
            with GameSolver(tables, precision=1) as s64:
                s64.sweep(done, flags)
                v64 = s64.read()
            dv = np.abs(v32 - v64)
            flips = int(((v32 <= 0) != (v64 <= 0)).sum())
            out.update(max_abs_dv_vs_fp64=float(dv.max()), mean_abs_dv_vs_fp64=float(dv.mean()),
                       capture_flips_vs_fp64=flips)
        if getattr(args, "vstar_sweeps", 0):
            out.update(time_to_vstar(s, flags, args.vstar_sweeps))
    if getattr(args, "cpu_sweeps", 0):
        out["cpu_baseline"] = c2_cpu_baseline(p, e, tables, v0, args.cpu_sweeps)
    return out


def time_to_vstar(s, flags, cap, tol=1e-3, every=64):
    """Sweeps and seconds from the payoff until max|V - V*| <= tol, where V*
    is the same solver's field at residual 1e-6 (or after `cap` sweeps).  The
    plain iterates decrease monotonically to V*, so the first check below
    tol is the time to converge at that accuracy."""
    import numpy as np

    s.initialize()
    done, ok = s.run(1e-6, cap, every, flags)
    vstar = s.read()
    res = s.stats()["last_residual"]
    s.initialize()
    t0 = s.stats()["update_seconds"]
    k, hit = 0, None
    while k < done:
        s.sweep(every, flags)
        k += every
        if np.abs(s.read() - vstar).max() <= tol:
            hit = (k, s.stats()["update_seconds"] - t0)
            break
    return {"vstar_sweeps": done, "vstar_residual": res, "vstar_converged_1e-6": ok,
            "sweeps_to_1e-3_of_vstar": hit[0] if hit else None,
            "gpu_seconds_to_1e-3_of_vstar": hit[1] if hit else None}


def run_c4(args):
    """C4: many seeded maps (9x9 positions, 8 headings, 4 speeds per car)
    solved together, one launch per kernel for all maps, sweeps replayed
    from a captured hipGraph."""
    from pe_amd import _native
    from pe_amd.local_game.grid_game import GameSolver
    from pe_amd.local_game.reference import benchmark_games

    t0 = time.perf_counter()
    # All maps' geometry in one launch per phase, then each map's stencils.
    games = benchmark_games(range(1000, 1000 + args.maps), positions=9, headings=8, speeds=4)
    tables = [g[4] for g in games]
    build_s = time.perf_counter() - t0
    print(f"[bench] C4 tables for {args.maps} maps ({build_s:.1f} s)", file=sys.stderr,
          flush=True)
    flags = _native.PE_SL_EARLY_EXIT | _native.PE_SL_ORDERED | _native.PE_SL_GRAPH
    with GameSolver(tables) as s:
        st0 = s.stats()
        if args.warmup:
            s.sweep(args.warmup, flags)
        t1 = time.perf_counter()
        s.sweep(args.steps, flags)
        wall = time.perf_counter() - t1
        res = s.trace()[-1]
        base = flags & ~_native.PE_SL_GRAPH
        win = sl_windows(s, 20, {
            "fused": base,
            "separate": base | _native.PE_SL_SEPARATE_P_SIDE,
            "stencil": base | _native.PE_SL_STENCIL_P_SIDE,
            "reference": _native.PE_SL_EARLY_EXIT})
        out = {
            "workload": f"C4 {args.maps} seeded 30-rectangle maps, 9x9 positions x 8 headings x "
                        "4 speeds per car, batched, hipGraph-replayed sweeps",
            "maps": args.maps, "legal_pairs": st0["pairs"], "active_pairs": st0["active_pairs"],
            "dense_pairs_per_map": (9 * 9 * 8 * 4) ** 2, "table_build_s": build_s,
            "sweeps": args.steps, "ms_per_sweep": wall / args.steps * 1e3,
            "legal_pair_updates_per_s": st0["pairs"] * args.steps / wall,
            "active_pair_updates_per_s": st0["active_pairs"] * args.steps / wall,
            "max_residual_after": float(res.max()),
            "e_side_lds_threads": st0["lds_e_side"],
            # The default sweep is one kernel: each E-side block computes its
            # P state's values into LDS from V (no G buffer).
            "fused_sweep_kernel_ms": win["fused"][1],
            # The same sweep with the P side as its own launch writing G.
            "p_side_ms": win["separate"][0], "e_side_ms": win["separate"][1],
            # That P side summing each command's stencil in vertex order.
            "p_side_stencil_ms": win["stencil"][0],
            # The reference's early exit in index order, without the stored
            # command order (separate P side).
            "e_side_reference_order_ms": win["reference"][1],
        }
        cars = [g[2] for g in games] + [g[3] for g in games]
        out["roofline"] = sl_roofline(win, st0, wall / args.steps, cars,
                                      load_counters("sl8d_e_side_c4_counters.json"))
        if args.to_residual:
            import numpy as np

            s.initialize()
            # Each map stops being swept at the first check (every 16 sweeps)
            # where its own residual is at most the tolerance, as if solved
            # alone; --no-freeze sweeps every map until the last converges.
            freeze = not args.no_freeze
            print(f"[bench] C4 to residual {args.to_residual} (freeze {freeze})",
                  file=sys.stderr, flush=True)
            t2 = time.perf_counter()
            done, ok = s.run(args.to_residual, args.max_sweeps, 16,
                             flags | (_native.PE_SL_FREEZE if freeze else 0))
            seconds = time.perf_counter() - t2
            checks = s.trace()[15::16]
            stops = [16 * (int(np.argmax(checks[:, k] <= args.to_residual)) + 1)
                     for k in range(args.maps)]
            print(f"[bench] C4 {done} sweeps, {seconds:.1f} s", file=sys.stderr, flush=True)
            out.update(to_residual=args.to_residual, freeze_converged_maps=freeze,
                       sweeps_to_residual=done, converged=ok, seconds_to_residual=seconds,
                       map_sweeps_to_residual={"min": min(stops), "median": int(np.median(stops)),
                                               "max": max(stops), "total": int(sum(stops))})
    if getattr(args, "cpu_sweeps", 0):
        # cpu_baseline leg: the oracle's FP64 OpenMP sweep of the first eight
        # maps from the payoff, in legal-pair updates per second.
        import numpy as np
        import oracle
        from pe_amd.local_game.grid_game import dense_inside

        threads = host_cores()[0]
        pairs = seconds = 0.0
        for _, arena, p, e, _ in games[:8]:
            inside = dense_inside(p, e, 3.0)
            V = oracle.sl_init(p.states, e.states, inside, 3.0, arena.pursuer, arena.evader,
                               threads=threads)
            t = time.perf_counter()
            for _ in range(args.cpu_sweeps):
                V, _ = oracle.sl_sweep(V, p, e, inside, threads=threads)
            seconds += time.perf_counter() - t
            pairs += len(p.states) * len(e.states) * args.cpu_sweeps
        out["cpu_baseline"] = {"value": pairs / seconds, "unit": "legal-pair updates/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{args.cpu_sweeps} FP64 Jacobi sweeps of each of maps "
                                         f"1000-1007 from the payoff ({seconds:.2f} s, OpenMP)",
                               "host": cores_note()}
    return out


def run_learned(args):
    """The learned warm start against the plain DP on C4-grid maps: network
    trained here on the GPU, held-out maps solved both ways."""
    from pe_amd.learned.warm_start import compare_batched

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    grid = dict(positions=9, headings=8, speeds=4)
    out = compare_batched(range(0, args.learn_train), range(100, 104),
                          range(1000, 1000 + args.learn_test), grid=grid,
                          steps=args.learn_steps, under_weight=args.learn_under_weight,
                          check_seeds=None if args.no_c4 else range(1000, 1000 + args.maps),
                          log=log)
    out["workload"] = (f"learned warm start vs plain DP: network trained on maps "
                       f"0-{args.learn_train - 1}, delta from maps 100-103, held-out maps "
                       f"1000-{999 + args.learn_test} (the first maps of the C4 batch); C4 "
                       "grid (9x9 positions x 8 headings x 4 speeds per car); each run swept "
                       "until every map is within 1e-3 of V* and at residual 1e-4; then the "
                       "whole C4 batch warm-started and judged against the plain DP")
    return out


def run_secondary(args, cfg):
    """The lines beside the headline: C1 to residual 1e-4, the 8D contact
    game (C2, C4), the learned warm start, the dense 8D square game and the
    5D reader.  A line that raises is recorded with its error, and no
    further line runs, so the headline line is still printed."""
    from pe_amd import _native

    secondary = {}

    def c1_to_residual():
        # The plain DP's time to residual 1e-4 of the C1 solve from phi.
        with _native.HJI5D(cfg) as h:
            t0 = time.perf_counter()
            ok = h.run_to_residual(1e-4, 30000, 16)
            wall = time.perf_counter() - t0
            st = h.stats()
        return {"tolerance": 1e-4, "converged": ok, "steps": st["internal_steps"],
                "solve_time_s": st["time"], "seconds": wall,
                "last_residual": st["last_residual"]}

    def c1_plane_ab():
        # The two-plane blend kernel (default for an even heading count)
        # against the one-plane kernel (PE_HJI5D_PSI_SINGLE) in one process:
        # both handles start from phi, then take the same steps in
        # alternating chunks, so the clock state is shared.
        import numpy as np

        single_cfg = type(cfg).from_buffer_copy(cfg)
        single_cfg.flags |= _native.PE_HJI5D_PSI_SINGLE
        chunk, rounds = 40, 4
        with _native.HJI5D(cfg) as pair, _native.HJI5D(single_cfg) as single:
            dt = pair.stats()["cfl_dt_limit"]
            handles = {"two_planes": pair, "one_plane": single}
            for h in handles.values():
                h.step(dt, 5)
            times = {k: [] for k in handles}
            for _ in range(rounds):
                for k, h in handles.items():
                    t0 = time.perf_counter()
                    h.step(dt, chunk)
                    times[k].append((time.perf_counter() - t0) / chunk * 1e3)
            out = {"chunk_steps": chunk, "rounds": rounds}
            for k, h in handles.items():
                h.set_profiling(True)
                h.step(dt, 20)
                st = h.stats()
                avg = [st["stage_ms"][j] / max(1, st["stage_launches"][j]) for j in range(2)]
                out[k] = {"ms_per_step_chunks": [round(t, 4) for t in times[k]],
                          "ms_per_step_median": round(float(np.median(times[k])), 4),
                          "stage1_launch_ms": round(avg[0], 5),
                          "blend_launch_ms": round(avg[1], 5)}
            f_pair, a_pair = pair.read()
            f_single, a_single = single.read()
            out["bit_identical"] = bool(np.array_equal(f_pair.view(np.uint32),
                                                       f_single.view(np.uint32))
                                        and np.array_equal(a_pair.view(np.uint32),
                                                           a_single.view(np.uint32)))
        out["speedup_two_planes"] = round(out["one_plane"]["ms_per_step_median"]
                                          / out["two_planes"]["ms_per_step_median"], 4)
        return out

    def c2_line():
        c2 = argparse.Namespace(seed=0, warmup=10, steps=200, to_residual=1e-4,
                                max_sweeps=20000, cpu_sweeps=0 if args.no_cpu_baseline else 48,
                                vstar_sweeps=args.vstar_sweeps)
        return run_c2(c2)

    def c4_line():
        c4 = argparse.Namespace(maps=128, warmup=10, steps=100, to_residual=1e-4,
                                max_sweeps=20000, no_freeze=False,
                                cpu_sweeps=0 if args.no_cpu_baseline else 100)
        return run_c4(c4)

    def learned_line():
        learned = run_learned(args)
        if "contact_game_c4" in secondary:
            # the plain DP's time to 1e-3 of V* on the first maps of the C4
            # batch, from the learned line's cold run
            cold = learned["cold"]
            secondary["contact_game_c4"]["to_1e-3_of_vstar"] = {
                "maps": f"1000-{999 + args.learn_test} (batched, no freezing)",
                "vstar": learned["vstar"],
                "sweeps": cold["to_vstar"], "to_residual_1e-4": cold["to_residual"]}
        return learned

    lines = [("hji5d_c1_to_residual", "C1 from phi to per-step residual 1e-4", c1_to_residual),
             ("hji5d_c1_plane_ab", "C1 two-plane against one-plane kernel", c1_plane_ab),
             ("contact_game_c2", "8D contact game (seed 0)", c2_line)]
    if not args.no_c4:
        lines.append(("contact_game_c4", "8D contact game, 128 maps (C4)", c4_line))
    if not args.no_learned:
        lines.append(("learned_warm_start_c4_grid", "learned warm start vs plain DP",
                      learned_line))
    lines += [("world8d_default", "dense 8D square game",
               lambda: run_world8d((9, 8, 7, 7), 40, 2, 0 if args.no_cpu_baseline else 40)),
              ("world8d_13x12x7x7", "dense 8D square game, 13x12x7x7",
               lambda: run_world8d((13, 12, 7, 7), 20, 2)),
              ("reader_c1", "5D table reader",
               lambda: run_reader_c1(cfg, cpu_queries=0 if args.no_cpu_baseline
                                     else 1_000_000))]
    for key, label, fn in lines:
        print(f"[bench] secondary: {label}", file=sys.stderr, flush=True)
        try:
            secondary[key] = fn()
        except Exception as exc:  # recorded; the remaining lines are skipped
            secondary[key] = {"error": f"{type(exc).__name__}: {exc}"}
            secondary["skipped_after_error"] = [k for k, _, _ in lines[
                [k for k, _, _ in lines].index(key) + 1:]]
            print(f"[bench] secondary line {key} failed: {exc}", file=sys.stderr, flush=True)
            break
    return secondary


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=None,
                    help="steps of the per-launch event pass (default: --steps)")
    ap.add_argument("--cpu-sample-steps", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the 8D contact-game line after the 5D headline")
    ap.add_argument("--workload", choices=["c1", "c2", "c4", "learned"], default="c1")
    ap.add_argument("--maps", type=int, default=128)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--to-residual", type=float, default=None)
    ap.add_argument("--max-sweeps", type=int, default=20000)
    ap.add_argument("--no-c4", action="store_true",
                    help="leave the 128-map C4 line out of the secondary object")
    ap.add_argument("--no-freeze", action="store_true",
                    help="C4: sweep every map until the last one converges")
    ap.add_argument("--no-learned", action="store_true",
                    help="leave the learned warm-start comparison out")
    ap.add_argument("--learn-train", type=int, default=16)
    ap.add_argument("--learn-test", type=int, default=8)
    ap.add_argument("--learn-steps", type=int, default=3000)
    ap.add_argument("--learn-under-weight", type=float, default=4.0,
                    help="learned line: loss weight of the network's underestimates")
    ap.add_argument("--vstar-sweeps", type=int, default=16384,
                    help="C2: cap of the run to V* for the time to 1e-3 of V* (0: skip)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="after the run, copy this many bytes once (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the 5D "
              "grid would shard by vP slabs with a +-2-plane halo exchange per "
              "stage).  Run with --gpus 1.", file=sys.stderr)
        return 2
    host_cores()   # before any OpenMP region narrows the affinity mask
    from pe_amd import _native

    info = _native.device_info()
    if args.workload == "learned":
        print(json.dumps(run_learned(args)), flush=True)
        return 0
    if args.workload in ("c2", "c4"):
        args.steps = args.steps or 200
        out = run_c2(args) if args.workload == "c2" else run_c4(args)
        if args.copy_probe:
            out["copy_probe_gbps"] = _native.copy_bandwidth(args.copy_probe, 1)
        print(json.dumps(out), flush=True)
        return 0
    args.steps = args.steps or 810
    config, grid, cfg = c1_config()
    n = grid.states
    print(f"[bench] {info['name']}: C1 5D HJI {n} nodes, warmup {args.warmup}, "
          f"steps {args.steps}", file=sys.stderr, flush=True)
    with _native.HJI5D(cfg) as h:
        dt = h.stats()["cfl_dt_limit"]
        if args.warmup:
            h.step(dt, args.warmup)
        before = h.stats()["update_seconds"]
        t0 = time.perf_counter()
        h.step(dt, args.steps)  # returns after the device has finished
        wall = time.perf_counter() - t0
        device_seconds = h.stats()["update_seconds"] - before
        print(f"[bench] timed {args.steps} steps: {wall:.4f} s wall, "
              f"{device_seconds:.4f} s device", file=sys.stderr, flush=True)
        residual = h.stats()["last_residual"]
    # Roofline pass: the same steps again from phi on a fresh handle, with
    # per-launch hipEvents on the library's stream, so that the per-launch
    # average covers the same part of the solve as the timed run (the step
    # cost drifts as the field evolves).
    profile_steps = args.steps if args.profile_steps is None else max(
        1, min(args.profile_steps, args.steps))
    with _native.HJI5D(cfg) as h:
        if args.warmup:
            h.step(dt, args.warmup)
        h.set_profiling(True)
        h.step(dt, profile_steps)
        st = h.stats()
    avg_ms = [st["stage_ms"][k] / max(1, st["stage_launches"][k]) for k in range(2)]
    achieved = [STAGE_BYTES[k] * n / (avg_ms[k] * 1e-3) / 1e9 for k in range(2)]
    traffic, traffic_kind = load_traffic()
    roofline = {
        "bound": "hbm",
        "limiter": "vector-memory instruction issue and load latency: the achieved "
                   "rate is set by the stage's neighbour loads per node (16 for the "
                   "two-plane blend stage), not by HBM bandwidth",
        "traffic_kind": traffic_kind,
        "kernel": "hji5d_stage_fast<true,2,false,true> (RK2 blend stage, two psi planes per thread)",
        "achieved": round(achieved[1], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved[1] / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": STAGE_BYTES[1] * n,
        "avg_launch_ms": round(avg_ms[1], 5),
        "stage1": {"achieved": round(achieved[0], 1), "avg_launch_ms": round(avg_ms[0], 5),
                   "algorithmic_bytes_per_launch": STAGE_BYTES[0] * n},
    }
    value = n * args.steps / wall
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "node-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "5D free-space reversible-car HJI (ENO2/LF/RK2), "
                        "RelativeGrid(65,36,15,15,8.0), A3 limits, capture radius 0.5 m, "
                        "CFL 0.8; one step = one RK2 step of all nodes",
            "nodes": n,
            "dt": dt,
            "horizon_s": dt * args.steps,
            "device_seconds": device_seconds,
            "last_step_residual": residual,
        },
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_sample_steps)
    if args.copy_probe:
        out["copy_probe_gbps"] = _native.copy_bandwidth(args.copy_probe, 1)
    if not args.no_secondary:
        out["secondary"] = run_secondary(args, cfg)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
