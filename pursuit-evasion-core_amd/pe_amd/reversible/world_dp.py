"""Dense 8D square game on the GPU: the drop-in for reversible/world_dp.py.

State [xP, yP, thP, vP, xE, yE, thE, vE] on a dense grid, row-major with vE
fastest, float32 fields.  The numerical scheme is the reference executable's
(reversible/native_world.cpp:32-314), evaluated by the pe_world8d kernels of
libpe_amd.so in the same float operation order: ENO1/ENO2 with a non-strict
selection, local Lax-Friedrichs dissipation, RK1/RK2, tube projection
max(min(value, previous), avoid) on completed updates, linear arrival times.

`solve` keeps the reference's options, validation, resource guards, output
files (meta.json, fields.f32, final.f32, arrival.f32, probe_fields.f64,
solver.log, dependencies.json) and meta keys (world_dp.py:177-426).  The
solver runs in this process on the device instead of a spawned executable;
`threads` is validated and recorded for compatibility only.
"""

from __future__ import annotations

from dataclasses import asdict, dataclass
import hashlib
import json
import math
import os
from pathlib import Path
import resource
import time

import numpy as np

from .. import _native
from .model import ReversibleConfig

STATE_ORDER = ["xP", "yP", "thetaP", "vP", "xE", "yE", "thetaE", "vE"]
PERIODIC_AXES = (2, 6)
TUBE_PROJECTION = "min over completed numerical updates, then max with avoid"


# Least allowed size of each grid dimension.
_LEAST = dict(positions=5, headings=4, p_speeds=3, e_speeds=3)


@dataclass(frozen=True)
class WorldGrid:
    """Dense grid of the square game: `positions` nodes per coordinate over
    [-extent, extent], `headings` periodic headings from -pi, and the speed
    ranges of the two cars (the reference's WorldGrid,
    reversible/world_dp.py:24-88)."""

    positions: int = 9
    headings: int = 8
    p_speeds: int = 7
    e_speeds: int = 7
    extent: float = 3.5

    def __post_init__(self):
        for name, least in _LEAST.items():
            value = getattr(self, name)
            integral = isinstance(value, (int, np.integer)) and not isinstance(value, bool)
            if not integral or value < least:
                raise ValueError(f"{name} must be an integer >= {least}.")
        if not (np.isfinite(self.extent) and self.extent > 0):
            raise ValueError("extent must be finite and positive.")

    @property
    def shape(self):
        car = (self.positions, self.positions, self.headings)
        return car + (self.p_speeds,) + car + (self.e_speeds,)

    @property
    def states(self):
        return math.prod(self.shape)

    def axes(self, config):
        side = np.linspace(-self.extent, self.extent, self.positions)
        heading = -np.pi + np.arange(self.headings) * 2 * np.pi / self.headings
        speed = [np.linspace(lim.speed_min, lim.speed_max, count)
                 for lim, count in ((config.pursuer, self.p_speeds),
                                    (config.evader, self.e_speeds))]
        return (side, side, heading, speed[0], side, side, heading, speed[1])

    def spacings(self, config):
        return tuple(float(axis[1] - axis[0]) for axis in self.axes(config))

    def to_dict(self, config):
        axes = self.axes(config)
        wall = config.arena_half_width
        offsets = [float(np.abs(axes[0] - w).min()) for w in (-wall, wall)]
        info = asdict(self)
        info.update(shape=list(self.shape), states=self.states,
                    spacings=list(self.spacings(config)), periodic_axes=list(PERIODIC_AXES),
                    state_order=STATE_ORDER,
                    axis_bounds=[[float(axis[0]), float(axis[-1])] for axis in axes],
                    physical_wall_grid_offsets=offsets,
                    physical_walls_on_nodes=max(offsets) < 1e-8)
        return info


def output_times(horizon, output_dt):
    """Saved budgets: multiples of output_dt below the horizon, then the
    horizon itself (reversible/world_dp.py:91-98)."""
    finite = np.isfinite(horizon) and np.isfinite(output_dt)
    if not finite or horizon <= 0 or output_dt <= 0:
        raise ValueError("Positive finite horizon and output_dt are required.")
    frames = math.ceil(horizon / output_dt) + 1
    return np.unique(np.minimum(np.arange(frames) * output_dt, horizon))


def _top_speed(limits):
    return max(abs(limits.speed_min), abs(limits.speed_max))


def global_rate_bound(config, grid):
    """Largest rate of the state over one grid spacing, summed over the
    axes: translation of both cars in x and y, turning, and the two
    throttles.  The CFL step is cfl / bound (reversible/world_dp.py:101-112)."""
    dx, _, dh, dvp, _, _, _, dve = grid.spacings(config)
    p, e = config.pursuer, config.evader
    vp, ve = _top_speed(p), _top_speed(e)
    translation = 2 * (vp + ve) / dx
    turning = (vp * p.max_curvature + ve * e.max_curvature) / dh
    return translation + turning + p.acceleration / dvp + e.acceleration / dve


def surfaces(config, grid, dtype=np.float32):
    """Target and avoid fields on the grid (reversible/world_dp.py:115-135):
    the target is min(centre distance - capture radius, evader wall
    clearance), the avoid field the pursuer's wall violation + 1e-6.  Dense
    arrays of the full grid: small grids only."""
    if grid.extent <= config.arena_half_width:
        raise ValueError("World grid needs an exterior padding band.")
    axes = grid.axes(config)
    wall = config.arena_half_width

    def coordinate(k):
        shape = [1] * 8
        shape[k] = -1
        return np.asarray(axes[k], dtype=dtype).reshape(shape)

    xp, yp, xe, ye = map(coordinate, (0, 1, 4, 5))
    capture = np.hypot(xp - xe, yp - ye) - config.capture_radius
    evader_in = wall - np.maximum(abs(xe), abs(ye))
    pursuer_out = np.maximum(abs(xp), abs(yp)) - wall + 1e-6
    full = grid.shape
    return (np.broadcast_to(np.minimum(capture, evader_in), full).copy(),
            np.broadcast_to(pursuer_out, full).copy())


def hamiltonian(states, gradients, config):
    """Upper value of the game's Hamiltonian at states [..., 8]: the pursuer
    minimises, the evader maximises gradient . f over their commands.
    Translation does not depend on the commands; turning contributes
    -|vP| kP |g_thP| and +|vE| kE |g_thE|; each throttle picks the better of
    full braking and full acceleration, an acceleration that would leave the
    speed interval being replaced by 0 (reversible/world_dp.py:138-152)."""
    s, g = np.asarray(states), np.asarray(gradients)
    p, e = config.pursuer, config.evader

    def drift(base):
        heading, speed = s[..., base + 2], s[..., base + 3]
        return speed * (np.cos(heading) * g[..., base] + np.sin(heading) * g[..., base + 1])

    value = drift(0) + drift(4)
    value = value - abs(s[..., 3]) * p.max_curvature * abs(g[..., 2])
    value = value + abs(s[..., 7]) * e.max_curvature * abs(g[..., 6])
    for axis, lim, choose in ((3, p, np.minimum), (7, e, np.maximum)):
        v, grad = s[..., axis], g[..., axis]
        brake = np.where(v <= lim.speed_min + 1e-10, 0, -lim.acceleration)
        push = np.where(v >= lim.speed_max - 1e-10, 0, lim.acceleration)
        value = value + choose(brake * grad, push * grad)
    return value


def native_config(config, grid, *, order=2, cfl=0.8, speed_boundary="linear"):
    """The C-ABI configuration of a solve (the executable's flags,
    reversible/world_dp.py:262-292)."""
    p, e = config.pursuer, config.evader
    c = _native.World8DConfig()
    c.positions, c.headings = grid.positions, grid.headings
    c.p_speeds, c.e_speeds = grid.p_speeds, grid.e_speeds
    c.order = order
    c.linear_speed_boundary = int(speed_boundary == "linear")
    c.extent, c.arena, c.radius = grid.extent, config.arena_half_width, config.capture_radius
    c.p_min, c.p_max, c.e_min, c.e_max = p.speed_min, p.speed_max, e.speed_min, e.speed_max
    c.p_accel, c.e_accel = p.acceleration, e.acceleration
    c.p_curvature, c.e_curvature = p.max_curvature, e.max_curvature
    c.cfl = cfl
    return c


def _library_digest():
    path = _native.LIB_PATH
    return hashlib.sha256(path.read_bytes()).hexdigest() if path.exists() else None


def _check_options(config, grid, order, cfl, threads, speed_boundary):
    """The solve's option rules, in the order the reference checks them."""
    if not isinstance(config, ReversibleConfig):
        raise TypeError("Use the reversible physical model.")
    rules = (
        (config.point_vehicles,
         "This square backend uses points; use the footprint-map backend."),
        (not getattr(config, "obstacles", ()),
         "This native 8D backend currently supports the square boundary only."),
        (grid.extent > config.arena_half_width,
         "Numerical domain must include exterior padding."),
        (speed_boundary in ("linear", "legacy_signed"),
         "Speed boundary must be linear or legacy_signed."),
        (order in (1, 2) and 0 < cfl <= 1 and isinstance(threads, int) and threads >= 1,
         "Require order1/2, CFL in(0,1], positive integer threads."),
    )
    for ok, message in rules:
        if not ok:
            raise ValueError(message)


def _check_parent(parent, config, grid, order, cfl, speed_boundary, history, horizon):
    """Continuation checks of reversible/world_dp.py:225-255."""
    if parent.get("status") != "complete" or parent.get("backend") != "native_world8d":
        raise ValueError("Continuation requires a completed native world table.")
    if parent["config"] != config.to_dict():
        raise ValueError("Continuation physics/map must match the parent.")
    keys = ("positions", "headings", "p_speeds", "e_speeds", "extent")
    if any(parent["grid"][k] != getattr(grid, k) for k in keys):
        raise ValueError("Continuation grid must match the parent.")
    if (parent.get("speed_boundary", "legacy_signed") != speed_boundary
            or parent["order"] != order or parent["cfl"] != cfl):
        raise ValueError("Continuation numerical scheme must match the parent.")
    if history and not parent["history_saved"]:
        raise ValueError("A continued query history requires the retained parent history.")
    start = float(parent["horizon"])
    if horizon <= start:
        raise ValueError("Continuation horizon must exceed the parent's horizon.")
    return start


def solve(
    config,
    grid,
    output,
    *,
    horizon=0.5,
    output_dt=0.25,
    order=2,
    cfl=0.8,
    threads=8,
    history=False,
    max_core_gb=12.0,
    max_output_gb=30.0,
    probes=None,
    speed_boundary="linear",
    resume_from=None,
    write_final=True,
    write_arrival=True,
):
    """Solve the fixed square on the GPU (reversible/world_dp.py:177-426).

    probes: optional (N, 8) states whose multilinearly interpolated values
    are stored at every output time in probe_fields.f64; they do not
    influence the computation.
    """
    _check_options(config, grid, order, cfl, threads, speed_boundary)
    parent = parent_path = None
    start_time = 0.0
    if resume_from is not None:
        parent_path = Path(resume_from).resolve()
        parent = json.loads((parent_path / "meta.json").read_text())
        start_time = _check_parent(parent, config, grid, order, cfl, speed_boundary,
                                   history, horizon)
    if not history and not write_final:
        raise ValueError("Retain history or a final field.")
    times = output_times(horizon - start_time, output_dt) + start_time
    if parent is not None:
        times = times[1:]  # the parent holds the shared boundary frame
    core_bytes = grid.states * 12
    output_bytes = grid.states * 4 * (
        int(write_final) + int(write_arrival) + (len(times) if history else 0))
    if not np.isfinite([max_core_gb, max_output_gb]).all() or min(max_core_gb,
                                                                   max_output_gb) <= 0:
        raise ValueError("Positive explicit resource limits are required.")
    if core_bytes > max_core_gb * 1e9 or output_bytes > max_output_gb * 1e9:
        raise MemoryError(
            f"Requested core {core_bytes/1e9:.3f}GB, outputs {output_bytes/1e9:.3f}GB "
            "exceed limits.")
    output = Path(output).resolve()
    if (output / "meta.json").exists():
        raise FileExistsError(f"Preserve existing run: {output}")
    output.mkdir(parents=True, exist_ok=True)
    cfg = native_config(config, grid, order=order, cfl=cfl, speed_boundary=speed_boundary)

    initial_path = initial_offset = None
    if parent is not None:
        if parent.get("final_source", "final.f32") == "history_last":
            initial_path = parent_path / "fields.f32"
            initial_offset = (len(parent["times"]) - 1) * grid.states * 4
        else:
            initial_path = parent_path / "final.f32"
            initial_offset = 0
        if initial_path.stat().st_size < initial_offset + grid.states * 4:
            raise ValueError("Parent continuation field is incomplete.")
    probe_count = 0
    stencil = None
    if probes is not None:
        from .grid_stencil import interpolation_stencil

        queries = np.asarray(probes, dtype=float).reshape(-1, 8)
        ii, ww, valid, _ = interpolation_stencil(queries, grid.axes(config), PERIODIC_AXES)
        if not valid.all():
            raise ValueError("All probe states must lie in the numerical domain.")
        probe_count = len(queries)
        stencil = (ii.astype(np.uint64), ww.astype(np.float64))
        stencil[0].tofile(output / "probe_indices.u64")
        stencil[1].tofile(output / "probe_weights.f64")
        queries.tofile(output / "probe_states.f64")

    digest = _library_digest()
    meta = dict(
        backend="native_world8d",
        device_backend="hip_gfx950",
        status="running",
        config=config.to_dict(),
        grid=grid.to_dict(config),
        horizon=horizon,
        output_dt=output_dt,
        times=times.tolist(),
        order=order,
        cfl=cfl,
        threads=threads,
        history_saved=history,
        dtype="float32",
        layout="C",
        target="capture_or_evader_boundary_fault",
        avoid="pursuer_boundary_fault",
        value_meaning="numerical_time_to_P_success_not_unconditionally_physical_capture",
        tube_projection=TUBE_PROJECTION,
        source_sha256=digest,
        library=str(_native.LIB_PATH),
        compile_seconds=0.0,
        command=None,
        predicted_core_bytes=core_bytes,
        predicted_output_bytes=output_bytes,
        probe_count=probe_count,
        peak_rss_scope="solving process",
        speed_boundary=speed_boundary,
        start_time=start_time,
        final_source="final.f32" if write_final else "history_last",
        arrival_retained=bool(write_arrival),
    )
    if parent is not None:
        meta["parent_table"] = os.path.relpath(parent_path, output)
        meta["parent_source_sha256"] = parent["source_sha256"]
        meta["continuation_note"] = ("Parent signed history is referenced, not copied; "
                                     "output times contain only new frames.")
        required = ["meta.json", "fields.f32"]
        if parent.get("final_source", "final.f32") != "history_last":
            required.append("final.f32")
        (output / "dependencies.json").write_text(json.dumps({
            "parent_table": meta["parent_table"],
            "required_parent_files_for_queries": required,
            "recursive_parent_dependencies": bool(parent.get("parent_table")),
            "instruction": "Retain the parent history while this continuation is in use.",
        }, indent=2) + "\n")
    path = output / "meta.json"
    path.write_text(json.dumps(meta, indent=2) + "\n")
    started = time.perf_counter()
    with (output / "solver.log").open("w") as log:
        try:
            timing = _run(cfg, grid, output, horizon, output_dt, start_time, history,
                          write_final, write_arrival, stencil, initial_path, initial_offset,
                          parent_path, log)
            returncode = 0
        except (RuntimeError, MemoryError, ValueError, OSError) as exc:
            log.write(f"{exc}\n")
            timing, returncode = None, 2
    meta.update(status="complete" if returncode == 0 else "failed", returncode=returncode,
                subprocess_seconds=time.perf_counter() - started)
    if timing is not None:
        meta.update(timing)
        if parent is not None:
            meta["cumulative_update_seconds"] = (
                parent.get("cumulative_update_seconds", parent["update_seconds"])
                + meta["update_seconds"])
            meta["cumulative_internal_steps"] = (
                parent.get("cumulative_internal_steps", parent["internal_steps"])
                + meta["internal_steps"])
    path.write_text(json.dumps(meta, indent=2) + "\n")
    if returncode:
        raise RuntimeError(f"8D solve failed; inspect internal log at {output}")
    return meta


def _run(cfg, grid, output, horizon, output_dt, start_time, history, write_final,
         write_arrival, stencil, initial_path, initial_offset, parent_path, log):
    """The executable's run() (native_world.cpp:115-366) through the C ABI."""
    begin = time.perf_counter()
    solver = _native.World8D(cfg)
    n = grid.states
    fields = probes_out = None
    write_seconds = probe_seconds = 0.0
    frames = 0
    try:
        if initial_path is not None:
            field = np.fromfile(initial_path, dtype=np.float32, count=n,
                                offset=initial_offset)
            if field.size != n:
                raise RuntimeError("Cannot load complete continuation field")
            arrival = None
            first = parent_path / "arrival.f32"
            if first.exists():
                arrival = np.fromfile(first, dtype=np.float32, count=n)
                if arrival.size != n:
                    raise RuntimeError("Cannot load continuation arrival diagnostic")
            solver.set_field(field, arrival, start_time)
        if stencil is not None:
            solver.set_probes(*stencil)
        initialization = time.perf_counter() - begin
        if history:
            fields = (output / "fields.f32").open("wb")
        if stencil is not None:
            probes_out = (output / "probe_fields.f64").open("wb")
        buf = np.empty(n, np.float32)

        def save():
            nonlocal write_seconds, probe_seconds, frames
            t = time.perf_counter()
            if fields is not None:
                solver.read(buf, False)
                buf.tofile(fields)
            write_seconds += time.perf_counter() - t
            if probes_out is not None:
                t = time.perf_counter()
                solver.probe().tofile(probes_out)
                probe_seconds += time.perf_counter() - t
            frames += 1

        if start_time == 0:
            save()
        count = int(math.ceil((horizon - start_time) / output_dt - 1e-10))
        for frame in range(1, count + 1):
            solver.advance_to(min(horizon, start_time + frame * output_dt))
            save()
            st = solver.stats()
            log.write(f"budget {st['time']:.10g} steps {st['internal_steps']} "
                      f"updates_s {st['update_seconds']:.10g}\n")
            log.flush()
        t = time.perf_counter()
        bad = solver.count_nonfinite()
        validation = time.perf_counter() - t
        if bad:
            raise RuntimeError("Nonfinite values in the final signed field")
        t = time.perf_counter()
        final, arrival = solver.read(bool(write_final), bool(write_arrival))
        if write_final:
            final.tofile(output / "final.f32")
        if write_arrival:
            arrival.tofile(output / "arrival.f32")
        write_seconds += time.perf_counter() - t
        st = solver.stats()
    finally:
        for f in (fields, probes_out):
            if f is not None:
                f.close()
        solver.close()
    return {
        "states": n,
        "internal_steps": st["internal_steps"],
        "frames": frames,
        "update_seconds": st["update_seconds"],
        "initialization_seconds": initialization,
        "write_seconds": write_seconds,
        "validation_seconds": validation,
        "probe_seconds": probe_seconds,
        "process_seconds": time.perf_counter() - begin,
        "core_array_bytes": n * 12,
        "peak_rss_bytes": resource.getrusage(resource.RUSAGE_SELF).ru_maxrss * 1024,
        "max_internal_dt": st["max_internal_dt"],
        "arithmetic": "float32 fields and local arithmetic",
    }


def main(argv=None):
    """Command line of reversible/world_dp.py:429-463 (same flags, plus
    --speed-boundary)."""
    import argparse

    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("output")
    parser.add_argument("--positions", type=int, default=9)
    parser.add_argument("--headings", type=int, default=8)
    parser.add_argument("--speeds", type=int, default=7)
    parser.add_argument("--extent", type=float, default=3.5)
    parser.add_argument("--horizon", type=float, default=0.5)
    parser.add_argument("--output-dt", type=float, default=0.25)
    parser.add_argument("--threads", type=int, default=8)
    parser.add_argument("--history", action="store_true")
    parser.add_argument("--order", type=int, choices=[1, 2], default=2)
    parser.add_argument("--max-core-gb", type=float, default=12)
    parser.add_argument("--max-output-gb", type=float, default=30)
    parser.add_argument("--speed-boundary", choices=["linear", "legacy_signed"],
                        default="linear")
    a = parser.parse_args(argv)
    m = solve(
        ReversibleConfig(),
        WorldGrid(a.positions, a.headings, a.speeds, a.speeds, a.extent),
        a.output,
        horizon=a.horizon,
        output_dt=a.output_dt,
        threads=a.threads,
        history=a.history,
        order=a.order,
        max_core_gb=a.max_core_gb,
        max_output_gb=a.max_output_gb,
        speed_boundary=a.speed_boundary,
    )
    print(json.dumps({k: m[k] for k in ["status", "states", "update_seconds", "peak_rss_bytes",
                                        "internal_steps"]}, indent=2))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
