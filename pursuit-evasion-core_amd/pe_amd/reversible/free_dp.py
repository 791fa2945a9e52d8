"""Reversible 5D free-space game: solve on the GPU behind the reference API.

Drop-in for `reversible.free_dp.solve(config, RelativeGrid(...), out, ...)`
(reference reversible/free_dp.py:131-276).  The same grid type, validation,
guards, exceptions and output directory layout:

    meta.json      schema "reversible_relative5d_v1" (+ timing fields)
    fields.f32     [frames, vP, vE, psi, y, x] float32 history (history=True)
    final.f32      final field, float32
    arrival.f32    first zero-crossing time per node (NaN = unresolved)
    timing.json    states, internal_steps, frames, cfl_dt_limit, ...
    progress.jsonl {"time","steps","resolved","elapsed_seconds"} lines
    solver.log     budget/steps/resolved lines

The update runs in libpe_amd.so (HIP kernels for gfx950) instead of a
spawned C++ executable; `threads` is accepted and validated for API
compatibility only.  Extra keywords: `local_precision` ("float32", the
default, or "float64" for the reference's double-precision local arithmetic),
`exit_radius` and `capture_footprints` (the local-exit clamp and rectangle
capture of reversible/gpu_relative.py:67-169), and `residual_tol` /
`max_extra_steps` to keep stepping past the horizon until the per-step
residual max|dV| falls below the tolerance.
"""

from __future__ import annotations

from dataclasses import asdict, dataclass
import hashlib
import json
import math
from pathlib import Path
import resource
import shutil
import time

import numpy as np

from .. import _native


@dataclass(frozen=True)
class RelativeGrid:
    """Grid of the relative state (the reference's RelativeGrid,
    reversible/free_dp.py:23-72): `positions` nodes per planar offset over
    [-extent, extent], `headings` relative headings from -pi without the
    repeated endpoint, and the two speed ranges.  Storage order [vP, vE, psi,
    y, x], x fastest."""

    positions: int = 41
    headings: int = 24
    p_speeds: int = 7
    e_speeds: int = 7
    extent: float = 4.0

    def __post_init__(self):
        least = dict(positions=3, headings=4, p_speeds=3, e_speeds=3)
        for name, low in least.items():
            value = getattr(self, name)
            whole = not isinstance(value, bool) and int(value) == value
            if not whole or value < low:
                raise ValueError(f"Invalid {name}: {value}")
        if not (np.isfinite(self.extent) and self.extent > 0):
            raise ValueError("extent must be finite and positive")

    @property
    def shape(self):
        return (self.p_speeds, self.e_speeds, self.headings, self.positions, self.positions)

    @property
    def states(self) -> int:
        return math.prod(self.shape)

    def axes(self, config):
        offset = np.linspace(-self.extent, self.extent, self.positions)
        speeds = [np.linspace(lim.speed_min, lim.speed_max, count)
                  for lim, count in ((config.pursuer, self.p_speeds),
                                     (config.evader, self.e_speeds))]
        heading = np.linspace(-np.pi, np.pi, self.headings, endpoint=False)
        return (speeds[0], speeds[1], heading, offset, offset.copy())

    def to_dict(self):
        return asdict(self)


def output_times(horizon, output_dt):
    """Saved budgets: multiples of output_dt up to the horizon, the last
    one clipped to it (free_dp.py:75-77)."""
    frames = math.ceil(horizon / output_dt - 1e-10) + 1
    return np.minimum(np.arange(frames) * output_dt, horizon)


_LIMIT_KEYS = ("speed_min", "speed_max", "acceleration", "max_curvature")


def physics_signature(config):
    """The parameters the free-space solution depends on: capture radius,
    both cars' limits and a non-default capture rule (free_dp.py:80-96)."""
    def limits(car):
        return {key: float(getattr(car, key)) for key in _LIMIT_KEYS}

    signature = dict(capture_radius=float(config.capture_radius),
                     pursuer=limits(config.pursuer), evader=limits(config.evader))
    if config.capture_rule != "center_distance":
        signature["capture_rule"] = config.capture_rule
    return signature


def _validate(config, grid, horizon, output_dt, order, cfl, threads):
    """Option checks with the reference's messages and exceptions
    (free_dp.py:99-128)."""
    cars = (config.pursuer, config.evader)
    if any(c.speed_min >= c.speed_max or c.acceleration <= 0 or c.max_curvature <= 0
           for c in cars):
        raise ValueError("Invalid vehicle limits")
    numbers = [horizon, output_dt, cfl, config.capture_radius, grid.extent]
    numbers += [getattr(c, key) for c in cars for key in _LIMIT_KEYS]
    if not (np.isfinite(numbers).all() and horizon > 0 and output_dt > 0 and 0 < cfl <= 1):
        raise ValueError("Invalid numeric solve options")
    if not (0 < config.capture_radius < grid.extent and order in (1, 2) and threads >= 1):
        raise ValueError("Invalid radius, order or thread count")


def native_config(config, grid, *, order=2, cfl=0.8, local_precision="float32",
                  exit_radius=None, capture_footprints=None):
    """The C-ABI configuration struct of a solve."""
    if local_precision not in ("float32", "float64"):
        raise ValueError("local_precision must be float32 or float64")
    p, e = config.pursuer, config.evader
    c = _native.HJI5DConfig()
    c.positions, c.headings = grid.positions, grid.headings
    c.p_speeds, c.e_speeds = grid.p_speeds, grid.e_speeds
    c.order = order
    c.precision = 1 if local_precision == "float64" else 0
    c.extent, c.cfl, c.radius = grid.extent, cfl, config.capture_radius
    c.p_min, c.p_max, c.e_min, c.e_max = p.speed_min, p.speed_max, e.speed_min, e.speed_max
    c.p_acceleration, c.e_acceleration = p.acceleration, e.acceleration
    c.p_curvature, c.e_curvature = p.max_curvature, e.max_curvature
    c.exit_radius = float(exit_radius) if exit_radius is not None else 0.0
    if capture_footprints is not None:
        pf, ef = capture_footprints
        c.capture = 1
        c.p_length, c.p_width = pf.length, pf.width
        c.e_length, c.e_width = ef.length, ef.width
    return c


def solve(
    config,
    grid,
    output,
    *,
    horizon=2.0,
    output_dt=0.1,
    order=2,
    cfl=0.8,
    threads=4,
    history=True,
    max_working_gb=8.0,
    disk_reserve_gb=10.0,
    local_precision="float32",
    exit_radius=None,
    capture_footprints=None,
    residual_tol=None,
    max_extra_steps=100_000,
):
    """Compute a new field directory, refusing overwrite and excessive
    allocation (reference free_dp.py:131-276)."""
    solve_started = time.perf_counter()
    if config.capture_rule == "rectangle_contact" and capture_footprints is None:
        raise ValueError("Rectangle capture requires explicit P/E footprints")
    if config.capture_rule == "center_distance" and capture_footprints is not None:
        raise ValueError("Declare rectangle_contact in the configuration")
    if exit_radius is not None and not 0 < exit_radius < grid.extent:
        raise ValueError("Local exit radius must lie inside the numerical grid")
    _validate(config, grid, horizon, output_dt, order, cfl, threads)
    if (
        not np.isfinite([max_working_gb, disk_reserve_gb]).all()
        or max_working_gb <= 0
        or disk_reserve_gb < 0
    ):
        raise ValueError("Invalid memory/disk guards")
    if residual_tol is not None and not (residual_tol > 0 and max_extra_steps >= 1):
        raise ValueError("Invalid residual options")
    output = Path(output).resolve()
    if output.exists() and any(output.iterdir()):
        raise FileExistsError(f"Preserve existing output: {output}")
    times = output_times(horizon, output_dt)
    working = grid.states * 12
    storage = grid.states * 4 * (2 + (len(times) if history else 0))
    if working > max_working_gb * 1e9:
        raise MemoryError(
            f"Core arrays need {working / 1e9:.3f} GB, above guard {max_working_gb}"
        )
    output.parent.mkdir(parents=True, exist_ok=True)
    if storage + disk_reserve_gb * 1e9 > shutil.disk_usage(output.parent).free:
        raise OSError(
            f"Output needs {storage / 1e9:.3f} GB while retaining {disk_reserve_gb} GB disk reserve"
        )
    cfg = native_config(
        config, grid, order=order, cfl=cfl, local_precision=local_precision,
        exit_radius=exit_radius, capture_footprints=capture_footprints,
    )
    output.mkdir(exist_ok=True)
    library = _native.LIB_PATH
    digest = hashlib.sha256(library.read_bytes()).hexdigest() if library.exists() else None
    meta = {
        "schema": "reversible_relative5d_v1",
        "backend": "hip_gfx950",
        "local_precision": local_precision,
        "config": config.to_dict(),
        "physics": physics_signature(config),
        "grid": grid.to_dict(),
        "shape": grid.shape,
        "state_order": ["x", "y", "psi", "vP", "vE"],
        "axis_order": ["vP", "vE", "psi", "y", "x"],
        "axes": [a.tolist() for a in grid.axes(config)],
        "horizon": horizon,
        "output_dt": output_dt,
        "times": times.tolist(),
        "order": order,
        "cfl": cfl,
        "threads": threads,
        "history_saved": bool(history),
        "dtype": "float32",
        "target": "physical_capture_only",
        "avoid": None,
        "tube_projection": "min over every internal update",
        "position_boundary": "one-sided linear extrapolation; outside queries unresolved",
        "speed_boundary": "acceleration restricted to inward or zero at speed limits",
        "heading_boundary": "periodic",
        "exit_radius": exit_radius,
        "capture_footprints": None
        if capture_footprints is None
        else [asdict(f) for f in capture_footprints],
        "library": str(library),
        "library_sha256": digest,
        "compile_seconds": 0.0,
        "command": None,
        "predicted_working_bytes": working,
        "predicted_output_bytes": storage,
        "status": "running",
    }
    (output / "meta.json").write_text(json.dumps(meta, indent=2) + "\n")
    started = time.perf_counter()
    log = (output / "solver.log").open("w")
    trace = (output / "progress.jsonl").open("w")
    try:
        timing = _run(cfg, grid, output, times, history, residual_tol,
                      max_extra_steps, started, log, trace)
        meta.update(status="complete", returncode=0)
        meta.update(timing)
        if meta["frames"] != len(times):
            meta["status"] = "failed"
            meta["failure"] = "Output time count mismatch"
    except (RuntimeError, MemoryError) as exc:
        log.write(f"{exc}\n")
        meta.update(status="failed", returncode=1, failure=str(exc))
    finally:
        log.close()
        trace.close()
    meta["subprocess_seconds"] = time.perf_counter() - started
    meta["total_wall_seconds"] = time.perf_counter() - solve_started
    (output / "meta.json").write_text(json.dumps(meta, indent=2) + "\n")
    if meta["status"] != "complete":
        raise RuntimeError(f"Relative solve failed: {output / 'solver.log'}")
    return meta


def _run(cfg, grid, output, times, history, residual_tol, max_extra_steps,
         started, log, trace):
    """The time loop of native_relative.cpp:173-233 driven through the C ABI."""
    t_alloc = time.perf_counter()
    solver = _native.HJI5D(cfg)
    allocation_seconds = time.perf_counter() - t_alloc
    n = grid.states
    write_seconds = scan_seconds = 0.0
    frames = 0
    buf = np.empty(n, np.float32)
    hist = (output / "fields.f32").open("wb") if history else None
    try:
        def save():
            nonlocal write_seconds, frames
            before = time.perf_counter()
            if hist is not None:
                solver.read(buf, False)
                buf.tofile(hist)
            write_seconds += time.perf_counter() - before
            frames += 1

        save()
        outputs = len(times) - 1
        for out in range(1, outputs + 1):
            solver.advance_to(float(times[out]))
            save()
            if out % 5 == 0 or out == outputs:
                before = time.perf_counter()
                resolved = solver.resolved()
                scan_seconds += time.perf_counter() - before
                st = solver.stats()
                elapsed = time.perf_counter() - started
                trace.write(json.dumps({"time": st["time"], "steps": st["internal_steps"],
                                        "resolved": resolved,
                                        "elapsed_seconds": elapsed}) + "\n")
                trace.flush()
                log.write(f"budget={st['time']:.12g} steps={st['internal_steps']} "
                          f"resolved={resolved} elapsed={elapsed:.6g}\n")
                log.flush()
        converged = None
        horizon_steps = solver.stats()["internal_steps"]
        if residual_tol is not None:
            converged = solver.run_to_residual(residual_tol, max_extra_steps)
        before = time.perf_counter()
        field, arrival = solver.read()
        field.tofile(output / "final.f32")
        arrival.tofile(output / "arrival.f32")
        final_write = time.perf_counter() - before
        st = solver.stats()
        residuals = solver.trace()
    finally:
        if hist is not None:
            hist.close()
        solver.close()
    timing = {
        "states": n,
        "working_array_bytes": 12 * n,
        "peak_rss_kib": resource.getrusage(resource.RUSAGE_SELF).ru_maxrss,
        "cfl_dt_limit": st["cfl_dt_limit"],
        "internal_steps": int(horizon_steps),
        "frames": frames,
        "allocation_seconds": allocation_seconds,
        "update_seconds": st["update_seconds"],
        "history_write_seconds": write_seconds,
        "final_write_seconds": final_write,
        "summary_scan_seconds": scan_seconds,
        "elapsed_seconds": time.perf_counter() - started,
        "maximum_increase": st["maximum_increase"],
        "horizon_residual": float(residuals[horizon_steps - 1]) if horizon_steps else None,
    }
    if residual_tol is not None:
        timing.update(
            residual_tol=residual_tol,
            residual_converged=bool(converged),
            residual_steps=int(st["internal_steps"]),
            residual_time=st["time"],
            final_residual=st["last_residual"],
        )
    (output / "timing.json").write_text(json.dumps(timing) + "\n")
    np.save(output / "residual_trace.npy", residuals)
    return timing


def main():
    import argparse

    from .model import ReversibleConfig

    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("output")
    parser.add_argument("--config")
    parser.add_argument("--positions", type=int, default=41)
    parser.add_argument("--headings", type=int, default=24)
    parser.add_argument("--p-speeds", type=int, default=7)
    parser.add_argument("--e-speeds", type=int, default=7)
    parser.add_argument("--extent", type=float, default=4)
    parser.add_argument("--horizon", type=float, default=2)
    parser.add_argument("--output-dt", type=float, default=0.1)
    parser.add_argument("--threads", type=int, default=4)
    parser.add_argument("--order", type=int, choices=[1, 2], default=2)
    parser.add_argument("--no-history", action="store_true")
    parser.add_argument("--max-working-gb", type=float, default=8)
    parser.add_argument("--local-precision", choices=["float32", "float64"], default="float32")
    parser.add_argument("--residual-tol", type=float)
    args = parser.parse_args()
    config = (
        ReversibleConfig.from_dict(json.loads(Path(args.config).read_text()))
        if args.config
        else ReversibleConfig()
    )
    grid = RelativeGrid(args.positions, args.headings, args.p_speeds, args.e_speeds, args.extent)
    result = solve(
        config, grid, args.output, horizon=args.horizon, output_dt=args.output_dt,
        threads=args.threads, order=args.order, history=not args.no_history,
        max_working_gb=args.max_working_gb, local_precision=args.local_precision,
        residual_tol=args.residual_tol,
    )
    print(json.dumps({k: result[k] for k in ("status", "states", "update_seconds", "peak_rss_kib")},
                     indent=2))


if __name__ == "__main__":
    main()
