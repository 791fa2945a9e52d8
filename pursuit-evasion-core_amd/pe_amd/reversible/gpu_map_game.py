"""Footprint-map 8D game on the GPU: the drop-in for
reversible/gpu_map_game.py.

The first-fault-loses game with oriented footprints and box obstacles: the
target is capture or an evader collision, the avoid set a pursuer collision,
so the field is a P-success reach-avoid value, not a capture time.  The
solver is the dense 8D square-game kernel of libpe_amd.so in its
footprint-map mode (`pe_world8d_set_clearance`): per-pose clearance grids
for both cars, the linear speed boundary, float heading spacing 2 pi / nh,
the CFL step cfl / global_rate_bound.  Options, validation, meta.json,
labels.npz, fields.f32, final.f32 and progress.jsonl follow
gpu_map_game.py:78-298.
"""

from __future__ import annotations

import hashlib
import json
from pathlib import Path
import shutil
import time

import numpy as np

from .. import _native
from .footprint_map import FootprintMap
from .world_dp import WorldGrid, global_rate_bound, native_config, output_times


def clearance_grids(geometry, grid, config):
    """Clearance of each car's footprint at every (x, y, heading) node
    (gpu_map_game.py:78-84)."""
    axes = grid.axes(config)
    x, y, h = np.meshgrid(axes[0], axes[1], axes[2], indexing="ij")
    poses = np.stack([x, y, h], axis=-1)
    return tuple(np.asarray(geometry.clearance(poses, role), np.float32)
                 for role in ("pursuer", "evader"))


def node_states(indices, grid, config):
    """States of flat node indices (gpu_map_game.py:87-89)."""
    q = np.unravel_index(np.asarray(indices, np.int64), grid.shape)
    return np.stack([axis[i] for axis, i in zip(grid.axes(config), q)], axis=-1)


def solve(config, grid, geometry, output, *, horizon=1.5, output_dt=0.5, order=2, cfl=0.7,
          query_indices=None, query_states=None, save_final=False, save_history=False,
          max_working_gb=7.6):
    """Solve and sample the footprint-map game (gpu_map_game.py:92-298)."""
    if config.capture_rule != "center_distance" or config.boundary_rule != "first_fault_loses":
        raise ValueError("Legacy map HJI uses center-distance capture and first-fault-loses")
    if config.point_vehicles:
        raise ValueError("Footprint HJI requires point_vehicles=False.")
    if not isinstance(geometry, FootprintMap) or not isinstance(grid, WorldGrid):
        raise TypeError("Use explicit FootprintMap and WorldGrid")
    if order not in (1, 2) or not 0 < cfl <= 1 or grid.extent <= geometry.half_width:
        raise ValueError("Require order1/2, CFL<=1 and spatial exterior padding")
    output = Path(output)
    output.mkdir(parents=True, exist_ok=True)
    if (output / "meta.json").exists():
        raise FileExistsError("Preserve existing solve")
    n = grid.states
    working = n * 12
    if working > max_working_gb * 1e9:
        raise MemoryError(f"Need {working/1e9:.3f} GB core GPU arrays")
    idx = np.asarray([] if query_indices is None else query_indices, np.int64)
    if idx.ndim != 1 or ((idx < 0) | (idx >= n)).any():
        raise ValueError("Invalid query node indices")
    states = node_states(idx, grid, config)
    if query_states is not None:
        if query_indices is not None:
            raise ValueError("Choose node indices OR continuous query states")
        from .grid_stencil import interpolation_stencil

        states = np.asarray(query_states, float)
        if states.ndim != 2 or states.shape[1] != 8 or len(states) > 100000:
            raise ValueError("Expected at most100000 world states[n,8]")
        stencil, weights, valid, _ = interpolation_stencil(states, grid.axes(config), (2, 6))
        if not valid.all():
            raise ValueError("Query outside numerical domain")
    else:
        # Node queries as one-vertex stencils: exact node values.
        stencil = np.repeat(idx[:, None], 256, axis=1)
        weights = np.zeros((len(idx), 256))
        weights[:, 0] = 1.0
    times = output_times(horizon, output_dt)
    expected_bytes = 4 * n * (int(save_final) + len(times) * int(save_history))
    if shutil.disk_usage(output).free < expected_bytes + 5 * 10**9:
        raise OSError("Insufficient disk space for requested retained fields")
    started = time.perf_counter()
    library = _native.LIB_PATH
    meta = dict(
        schema="footprint_map_game8d_v1", status="running", config=config.to_dict(),
        geometry=geometry.to_dict(), grid=grid.to_dict(config), times=times.tolist(),
        target="capture_OR_evader_collision", avoid="pursuer_collision",
        game_rule="first_fault_loses; P collision wins numerical event ties",
        output_meaning="signed finite-budget P-success reach-avoid value, NOT capture time",
        core_array_bytes=working, order=order, cfl=cfl, local_precision="float32",
        state_order=["xP", "yP", "headingP", "vP", "xE", "yE", "headingE", "vE"],
        source_sha256=hashlib.sha256(library.read_bytes()).hexdigest(),
        device_backend="hip_gfx950", query_count=len(states),
        history_saved=bool(save_history))
    (output / "meta.json").write_text(json.dumps(meta, indent=2))
    meta["compile_seconds"] = 0.0
    pc, ec = clearance_grids(geometry, grid, config)
    cfg = native_config(config, grid, order=order, cfl=cfl, speed_boundary="linear")
    cfg.arena = geometry.half_width
    maxdt = cfl / global_rate_bound(config, grid)
    updates = 0.0
    steps = 0
    with _native.World8D(cfg) as w:
        w.set_clearance(pc, ec)
        w.set_step_limit(maxdt)
        if len(states):
            w.set_probes(stencil.astype(np.uint64), weights.astype(np.float64))
        buf = np.empty(n, np.float32)
        history = (output / "fields.f32").open("wb") if save_history else None

        def sample():
            return w.probe().astype(np.float32) if len(states) else np.zeros(0, np.float32)

        try:
            if history is not None:
                w.read(buf, False)
                buf.tofile(history)
            samples = [sample()]
            for goal in times[1:]:
                before = w.stats()
                w.advance_to(float(goal))
                after = w.stats()
                updates += after["update_seconds"] - before["update_seconds"]
                steps = after["internal_steps"]
                samples.append(sample())
                if history is not None:
                    w.read(buf, False)
                    buf.tofile(history)
                with (output / "progress.jsonl").open("a") as f:
                    f.write(json.dumps({"budget": after["time"], "steps": steps,
                                        "update_seconds": updates}) + "\n")
                print(f"budget={after['time']:.3f}s updates={updates:.2f}s states={n:,}",
                      flush=True)
        finally:
            if history is not None:
                history.close()
        if w.count_nonfinite():
            raise RuntimeError("Non-finite field")
        final, arrival = w.read(True, True)
    np.savez_compressed(
        output / "labels.npz", indices=idx, states=states, times=times,
        values=np.stack(samples),
        arrival=arrival[idx] if query_states is None else np.full(len(states), np.nan))
    if save_final:
        final.tofile(output / "final.f32")
    meta.update(status="complete", update_seconds=updates, steps=steps, max_internal_dt=maxdt,
                wall_seconds=time.perf_counter() - started, saved_final=bool(save_final),
                output_bytes=sum(x.stat().st_size for x in output.iterdir() if x.is_file()))
    (output / "meta.json").write_text(json.dumps(meta, indent=2))
    return meta
