"""Contact-game reference solve on the GPU behind the reference API.

Drop-in for `local_game.reference.solve(config, arena, output, ...)`
(reference local_game/reference.py:138-369): same rule checks, spec
signature and resume, same output directory:

    meta.json             schema "local_contact_compact8d_v1"
    p_grid.npz e_grid.npz VehicleGrid tables of each car
    domain.npz            PairDomain (packed pair layout)
    queries.npz           sampled query states (sample_queries)
    labels.npz            field at the queries per saved budget
    fields.f16            signed-preserving float16 history (retain_history)
    checkpoint_NNNN.f32   latest packed field (atomic; removed on completion)

The backup runs device-resident in libpe_amd.so for the whole solve (no
per-sweep host staging).  Extra keywords: `positions` / `half_width` for an
explicit x = y = linspace(-half, half, positions) axis, `p_speeds` /
`e_speeds` (the reference fixes 7), `residual_tol` / `max_sweeps` to keep
sweeping after the horizon until max|dV| <= tol, `warm_start` (a dense
[p.n, e.n] field, applied as min(phi, warm_start)), and `flags`
(PE_SL_EARLY_EXIT | PE_SL_ORDERED).
"""

from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path
import shutil
import time

import numpy as np

from .. import _native
from ..reversible.map_game_field import first_crossing
from ..reversible.model import ReversibleConfig, wrap_angle
from .geometry import CircleArena
from .grid_game import GameSolver, PairDomain, VehicleGrid, game_tables, make_grid
from .native import GeometryEngine


def write_half_preserving_sign(values, file):
    """float16 history that never rounds a nonzero value to zero
    (reference.py:19-25)."""
    for start in range(0, len(values), 2_000_000):
        x = values[start: start + 2_000_000]
        y = x.astype(np.float16)
        y[(x > 0) & (y == 0)] = np.nextafter(np.float16(0), np.float16(1))
        y[(x < 0) & (y == 0)] = np.nextafter(np.float16(0), np.float16(-1))
        y.tofile(file)


def sample_queries(p, e, domain, arena, config, *, frames=192, per_frame=256, seed=1031):
    """Observation frames and query pair states for the labels
    (reference.py:28-124): per frame a pursuer pose with >= 30 evader states
    nearby, `per_frame` non-captured pairs, a jittered sensor pose, and a
    continuous perturbation of half the query states kept when legal."""
    rng = np.random.default_rng(seed)
    observation, pi, ei, frame_ids = [], [], [], []
    for f in range(frames):
        observer = 0
        own, other = (p, e) if observer == 0 else (e, p)
        for _ in range(1000):
            i = int(rng.integers(len(own.states)))
            pose = own.states[i, :3].copy()
            nearby = np.flatnonzero(
                np.linalg.norm(other.states[:, :2] - pose[:2], axis=1) <= domain.exit_radius - 0.1)
            if len(nearby) < 30:
                continue
            j = rng.choice(nearby, per_frame * 3, replace=True)
            same = np.flatnonzero(np.all(own.states[:, :3] == pose, axis=1))
            ii = rng.choice(same, len(j), replace=True)
            pids, eids = (ii, j) if observer == 0 else (j, ii)
            q = np.concatenate([p.states[pids], e.states[eids]], 1)
            keep = arena.capture_value(q) > 0.03
            pids, eids = pids[keep][:per_frame], eids[keep][:per_frame]
            if len(pids) < per_frame:
                continue
            sensor = pose.copy()
            for _ in range(20):
                jitter = 0.11 if observer == 0 else 0.06
                sensor = pose + np.array([rng.uniform(-jitter, jitter),
                                          rng.uniform(-jitter, jitter),
                                          rng.uniform(-0.11, 0.11)])
                role = "pursuer" if observer == 0 else "evader"
                if arena.clearance(np.r_[sensor, 0], role) > 1e-4:
                    break
            else:
                sensor = pose.copy()
            observation.append([*sensor, observer])
            pi.extend(pids)
            ei.extend(eids)
            frame_ids.extend([f] * per_frame)
            break
        else:
            raise ValueError("Could not sample a legal observation/query group")
    pi = np.array(pi, np.int32)
    ei = np.array(ei, np.int32)
    local = domain.lookup[domain.p_position[pi], ei]
    assert (local >= 0).all()
    compact = domain.offsets[pi] + local
    states = np.concatenate([p.states[pi], e.states[ei]], 1)
    proposal = states.copy()
    continuous = rng.random(len(states)) < 0.5
    for k, limits in ((0, config.pursuer), (4, config.evader)):
        proposal[:, k: k + 2] += rng.uniform(-0.12, 0.12, (len(states), 2))
        proposal[:, k + 2] = wrap_angle(proposal[:, k + 2] + rng.uniform(-0.15, 0.15, len(states)))
        proposal[:, k + 3] = np.clip(proposal[:, k + 3] + rng.uniform(-0.1, 0.1, len(states)),
                                     limits.speed_min, limits.speed_max)
    pe_ = GeometryEngine(arena, config.pursuer, "pursuer", control_dt=config.control_dt)
    ee_ = GeometryEngine(arena, config.evader, "evader", control_dt=config.control_dt)
    legal = pe_.backup_safe(proposal[:, :4]) & ee_.backup_safe(proposal[:, 4:])
    legal &= (arena.capture_value(proposal) > 0.03) & (
        np.linalg.norm(proposal[:, :2] - proposal[:, 4:6], axis=1) < domain.exit_radius - 0.02)
    legal &= p.stencil(proposal[:, :4])[2] & e.stencil(proposal[:, 4:])[2]
    continuous &= legal
    states[continuous] = proposal[continuous]
    return dict(frames=np.array(observation), frame_ids=np.array(frame_ids, np.int32),
                p_indices=pi, e_indices=ei, original_compact_indices=compact,
                states=states, continuous=continuous)


def build_grids(config, arena, *, headings=16, position_step=0.3, positions=None,
                half_width=None, p_speeds=7, e_speeds=7):
    dt = config.control_dt
    kw = dict(headings=headings, position_step=position_step, positions=positions,
              half_width=half_width, dt=dt)
    p = make_grid(arena, config.pursuer, "pursuer",
                  speeds=np.linspace(config.pursuer.speed_min, config.pursuer.speed_max, p_speeds), **kw)
    e = make_grid(arena, config.evader, "evader",
                  speeds=np.linspace(config.evader.speed_min, config.evader.speed_max, e_speeds), **kw)
    return p, e


def solve(
    config,
    arena,
    output,
    *,
    headings=16,
    position_step=0.3,
    horizon=2.0,
    output_dt=0.25,
    exit_radius=3.0,
    retain_history=False,
    frames=192,
    per_frame=256,
    seed=1031,
    max_gb=7.6,
    positions=None,
    half_width=None,
    p_speeds=7,
    e_speeds=7,
    residual_tol=None,
    max_sweeps=100_000,
    warm_start=None,
    flags=_native.PE_SL_EARLY_EXIT | _native.PE_SL_ORDERED,
):
    """Streaming contact-game solve with resumable atomic checkpoints."""
    if (config.capture_rule != "rectangle_contact"
            or config.boundary_rule != "collision_constrained" or config.point_vehicles):
        raise ValueError("Use the explicit finite-body collision-constrained contact game")
    output = Path(output)
    output.mkdir(parents=True, exist_ok=True)
    dt = config.control_dt
    steps = round(horizon / dt)
    every = round(output_dt / dt)
    if min(steps, every) < 1 or not np.isclose(steps * dt, horizon) or not np.isclose(every * dt, output_dt):
        raise ValueError("Budgets must be integer control periods")
    spec = dict(config=config.to_dict(), arena=arena.to_dict(), headings=headings,
                position_step=position_step, horizon=horizon, output_dt=output_dt,
                exit_radius=exit_radius, retain_history=retain_history, frames=frames,
                per_frame=per_frame, seed=seed, query_version=3)
    if positions is not None or half_width is not None or (p_speeds, e_speeds) != (7, 7):
        spec.update(positions=positions, half_width=half_width, p_speeds=p_speeds,
                    e_speeds=e_speeds)
    signature = hashlib.sha256(json.dumps(spec, sort_keys=True).encode()).hexdigest()
    mp = output / "meta.json"
    started = time.perf_counter()
    resume = mp.exists()
    if resume:
        meta = json.loads(mp.read_text())
        if meta["signature"] != signature:
            raise ValueError("Existing output belongs to another model/setup")
        if meta["status"] == "complete":
            return meta
        p = VehicleGrid.load(output / "p_grid.npz")
        e = VehicleGrid.load(output / "e_grid.npz")
        domain = PairDomain.load(output / "domain.npz")
        queries = dict(np.load(output / "queries.npz"))
    else:
        p, e = build_grids(config, arena, headings=headings, position_step=position_step,
                           positions=positions, half_width=half_width,
                           p_speeds=p_speeds, e_speeds=e_speeds)
        domain = PairDomain.build(p, e, exit_radius)
        p.save(output / "p_grid.npz")
        e.save(output / "e_grid.npz")
        domain.save(output / "domain.npz")
        queries = sample_queries(p, e, domain, arena, config, frames=frames,
                                 per_frame=per_frame, seed=seed)
        np.savez_compressed(output / "queries.npz", **queries)
        meta = dict(
            schema="local_contact_compact8d_v1", status="running", signature=signature,
            spec=spec, p_states=len(p.states), e_states=len(e.states),
            pair_states=domain.states, unpruned_pair_states=len(p.states) * len(e.states),
            history_dtype="float16_sign_preserved", completed_step=0, completed_outputs=0,
            times=[], update_seconds=0.0, backend="hip_gfx950",
            interpolation="multilinear, omit invalid single-vehicle vertices and renormalize",
            action_constraint="continuous obstacle-free command plus one collision-free "
                              "constant-curvature braking backup",
            objective="capture by budget before center separation exceeds local radius; "
                      "range exit is local failure, not a physical capture",
            numerical_game="finite-grid semi-Lagrangian minP maxE, exact factorization of "
                           "the product interpolation; no dimension reduction",
        )
        mp.write_text(json.dumps(meta, indent=2))
    p.bind_geometry(arena, config.pursuer, "pursuer", dt)
    e.bind_geometry(arena, config.evader, "evader", dt)
    expected = (len(range(0, steps + 1, every)) * 2 if retain_history else 0) * domain.states \
        + 8 * domain.states
    already = sum(f.stat().st_size for f in output.iterdir() if f.is_file())
    if shutil.disk_usage(output).free < max(0, expected - already) + 8 * 10**9:
        raise OSError("Insufficient disk space for history/checkpoint transaction plus 8GB reserve")
    need = len(p.states) * len(e.states) * (4 * 2 + 1 + 4 * len(p.controls))
    if need > max_gb * 1e9:
        raise MemoryError(f"{need / 1e9:.2f}GB required; guard {max_gb}GB")
    tables = game_tables(p, e, arena, exit_radius)
    solver = GameSolver(tables)
    pi, pw, pvalid, _, _ = p.stencil(queries["states"][:, :4])
    ei, ew, evalid, _, _ = e.stencil(queries["states"][:, 4:])
    if not (pvalid & evalid).all():
        raise ValueError("Generated query has no numerical support")
    try:
        checkpoint = output / meta.get("checkpoint", "missing")
        if resume and meta["completed_outputs"] and checkpoint.exists():
            packed = np.fromfile(checkpoint, np.float32, count=domain.states)
            solver.initialize(domain.unpack(packed, exit_radius, len(e.states)))
            prior = np.load(output / "labels.npz")
            labels = list(prior["values"][: meta["completed_outputs"]])
            times = list(prior["times"][: meta["completed_outputs"]])
            start = meta["completed_step"]
        else:
            solver.initialize()
            if warm_start is not None:
                phi = solver.read()
                solver.initialize(np.minimum(phi, np.asarray(warm_start, np.float32)))
            labels, times, start = [], [], 0
        history = output / "fields.f16"
        if retain_history:
            with history.open("ab" if history.exists() else "wb") as f:
                f.truncate(meta["completed_outputs"] * domain.states * 2)

        def save_output(step):
            field = solver.read()
            labels.append(arena.capture_value(queries["states"]).astype(np.float32)
                          if step == 0 else solver.query(pi, pw, ei, ew))
            times.append(step * dt)
            packed = domain.pack(field)
            if retain_history:
                with history.open("ab") as file:
                    write_half_preserving_sign(packed, file)
            np.savez_compressed(output / "labels.npz", times=np.array(times), values=np.stack(labels))
            name = f"checkpoint_{step:04d}.f32"
            tmp = output / (name + ".partial")
            packed.tofile(tmp)
            os.replace(tmp, output / name)
            old = meta.get("checkpoint")
            meta.update(completed_step=step, completed_outputs=len(times), times=times,
                        checkpoint=name, gpu_working_bytes=need,
                        host_working_bytes=domain.states * 4)
            temporary = mp.with_suffix(".tmp")
            temporary.write_text(json.dumps(meta, indent=2))
            os.replace(temporary, mp)
            if old and old != name and (output / old).exists():
                (output / old).unlink()
            print(f'budget {step * dt:.2f}s; states {domain.states:,}; '
                  f'updates {meta["update_seconds"]:.1f}s', flush=True)

        if not labels:
            save_output(0)
        for step in range(start + 1, steps + 1):
            before = solver.stats()["update_seconds"]
            solver.sweep(1, flags)
            meta["update_seconds"] += solver.stats()["update_seconds"] - before
            if step % every == 0 or step == steps:
                save_output(step)
        if residual_tol is not None:
            before = solver.stats()["update_seconds"]
            done, converged = solver.run(residual_tol, max_sweeps, 8, flags)
            meta.update(residual_tol=residual_tol, residual_sweeps=done,
                        residual_converged=converged,
                        final_residual=solver.stats()["last_residual"],
                        residual_seconds=solver.stats()["update_seconds"] - before)
            np.asarray(domain.pack(solver.read()), np.float32).tofile(output / "converged.f32")
        meta["residual_trace"] = solver.trace()[:, 0].tolist()
    finally:
        solver.close()
    meta.update(status="complete", wall_seconds_this_process=time.perf_counter() - started)
    mp.write_text(json.dumps(meta, indent=2))
    (output / meta["checkpoint"]).unlink()
    meta.pop("checkpoint")
    meta["saved_bytes"] = sum(f.stat().st_size for f in output.iterdir() if f.is_file())
    mp.write_text(json.dumps(meta, indent=2))
    return meta


def contact_config(config=None):
    """A3 limits with the finite-body contact rules."""
    from dataclasses import replace

    base = config or ReversibleConfig.from_dict({
        "name": "reversible_free_a3",
        "pursuer": {"speed_min": -1.0, "speed_max": 2.0, "acceleration": 3.0, "max_curvature": 1.0},
        "evader": {"speed_min": -0.5, "speed_max": 1.0, "acceleration": 3.0, "max_curvature": 2.0},
        "capture_radius": 0.3, "control_dt": 0.05, "arena_half_width": 3.0,
        "boundary_rule": "first_fault_loses", "point_vehicles": True})
    return replace(base, point_vehicles=False, capture_rule="rectangle_contact",
                   boundary_rule="collision_constrained")


def benchmark_game(seed, *, positions=11, headings=12, speeds=5, half_width=3.0,
                   exit_radius=3.0, rectangles=30):
    """The seeded random-rectangle contact game of the benchmark configs:
    two A3 cars with 0.5 x 0.3 m footprints, x = y = linspace(-3, 3,
    positions), `headings` headings, `speeds` speeds per car, 30 random
    rectangles rasterised at 0.05 m.  Returns (config, arena, p, e, tables)."""
    from .geometry import Footprint, random_rectangle_arena

    config = contact_config()
    body = Footprint(0.5, 0.3)
    arena = random_rectangle_arena(seed, pursuer=body, evader=body, count=rectangles,
                                   half_width=half_width)
    p, e = build_grids(config, arena, headings=headings, positions=positions,
                       half_width=half_width, p_speeds=speeds, e_speeds=speeds)
    return config, arena, p, e, game_tables(p, e, arena, exit_radius)


def benchmark_games(seeds, *, positions=11, headings=12, speeds=5, half_width=3.0,
                    exit_radius=3.0, rectangles=30):
    """benchmark_game for several seeds, with the geometry of all maps in one
    GPU launch per phase (make_grids); equal to benchmark_game seed by
    seed.  Returns a list of (config, arena, p, e, tables)."""
    from .geometry import Footprint, random_rectangle_arena
    from .grid_game import make_grids

    config = contact_config()
    body = Footprint(0.5, 0.3)
    arenas = [random_rectangle_arena(seed, pursuer=body, evader=body, count=rectangles,
                                     half_width=half_width) for seed in seeds]
    kw = dict(headings=headings, positions=positions, half_width=half_width,
              dt=config.control_dt)
    ps = make_grids(arenas, config.pursuer, "pursuer", speeds=np.linspace(
        config.pursuer.speed_min, config.pursuer.speed_max, speeds), **kw)
    es = make_grids(arenas, config.evader, "evader", speeds=np.linspace(
        config.evader.speed_min, config.evader.speed_max, speeds), **kw)
    return [(config, a, p, e, game_tables(p, e, a, exit_radius))
            for a, p, e in zip(arenas, ps, es)]


class ContactReference:
    """Reader of a completed solve with retained history (reference.py:372-454).

    The float16 history stays on the GPU; `fields` sums the 256-term product
    stencil of every query at every saved budget in one launch
    (pe_history_query).  Stencils, the payoff and the final masking are
    computed on the host as in the reference."""

    def __init__(self, folder):
        self.folder = Path(folder)
        self.meta = json.loads((self.folder / "meta.json").read_text())
        if self.meta["status"] != "complete" or not self.meta["spec"]["retain_history"]:
            raise ValueError("A complete signed history is required")
        if not (self.folder / "fields.f16").exists():
            raise ValueError("History was not retained; regenerate into a new output folder")
        self.config = ReversibleConfig.from_dict(self.meta["spec"]["config"])
        self.arena = CircleArena.from_dict(self.meta["spec"]["arena"])
        self.p = VehicleGrid.load(self.folder / "p_grid.npz")
        self.e = VehicleGrid.load(self.folder / "e_grid.npz")
        self.domain = PairDomain.load(self.folder / "domain.npz")
        dt = self.config.control_dt
        self.p.bind_geometry(self.arena, self.config.pursuer, "pursuer", dt)
        self.e.bind_geometry(self.arena, self.config.evader, "evader", dt)
        self.times = np.array(self.meta["times"])
        self.horizon = self.times[-1]
        if (self.folder / "fields.f16").stat().st_size != len(self.times) * self.domain.states * 2:
            raise ValueError("Incomplete history length")
        self.values = np.memmap(self.folder / "fields.f16", np.float16, mode="r",
                                shape=(len(self.times), self.domain.states))
        self._history = None

    def history(self):
        """The device-resident history (uploaded on first use)."""
        if self._history is None:
            self._history = _native.History(np.asarray(self.values))
        return self._history

    def stencil(self, s):
        """Packed pair ids [n,16,16], product weights, stored-pair mask,
        validity, inside-radius mask and the query batch shape."""
        s = np.asarray(s, float)
        shape = s.shape[:-1]
        flat = s.reshape(-1, 8)
        pi, pw, pv, _, _ = self.p.stencil(flat[:, :4])
        ei, ew, ev, _, _ = self.e.stencil(flat[:, 4:])
        rank = self.domain.lookup[self.domain.p_position[pi][:, :, None], ei[:, None, :]]
        ids = self.domain.offsets[pi][:, :, None] + np.maximum(rank, 0)
        weights = pw[:, :, None] * ew[:, None, :]
        valid = pv & ev
        inside = np.linalg.norm(flat[:, :2] - flat[:, 4:6], axis=1) <= self.domain.exit_radius
        return ids, weights, rank >= 0, valid, inside, shape

    def fields(self, s):
        """Values at every saved budget [budgets, n], the valid-and-inside
        mask and the batch shape."""
        ids, w, kept, valid, inside, shape = self.stencil(s)
        n = len(ids)
        values = self.history().query(ids.reshape(n, -1), w.reshape(n, -1),
                                      kept.reshape(n, -1), self.domain.exit_radius)
        phi = self.arena.capture_value(np.asarray(s, float).reshape(-1, 8))
        values = np.minimum(values, phi[None])
        values = np.where(inside[None], values, self.domain.exit_radius)
        values[:, ~valid] = np.nan
        values[0, valid & inside] = phi[valid & inside]
        return values, valid & inside, shape

    def field(self, s, budget=None):
        values, valid, shape = self.fields(s)
        budget = self.horizon if budget is None else budget
        if not 0 <= budget <= self.horizon:
            raise ValueError("Budget outside reference")
        i = min(np.searchsorted(self.times, budget, side="right") - 1, len(self.times) - 2)
        f = (budget - self.times[i]) / (self.times[i + 1] - self.times[i])
        return ((1 - f) * values[i] + f * values[i + 1]).reshape(shape)

    def time_value(self, s):
        values, valid, shape = self.fields(s)
        return first_crossing(self.times, values, valid).reshape(shape)

    def close(self):
        if self._history is not None:
            self._history.close()
            self._history = None
