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
    """Append values as float16, keeping the sign of every nonzero value: a
    value that would round to zero becomes the smallest subnormal of its
    sign (reference.py:19-25)."""
    tiny = np.float16(2.0 ** -24)
    for start in range(0, len(values), 1 << 21):
        x = np.asarray(values[start: start + (1 << 21)])
        half = x.astype(np.float16)
        lost = (half == 0) & (x != 0)
        half[lost] = np.where(x[lost] > 0, tiny, -tiny)
        half.tofile(file)


def _pick_frame(rng, own, other, domain, arena, per_frame, pose_ok):
    """One observation frame: a pursuer pose with enough evader states in
    range and `per_frame` non-captured pairs at that pose; None if the draw
    fails (the caller redraws)."""
    i = int(rng.integers(len(own.states)))
    if not pose_ok[i]:
        return None
    pose = own.states[i, :3]
    near = np.flatnonzero(np.hypot(*(other.states[:, :2] - pose[:2]).T)
                          <= domain.exit_radius - 0.1)
    same = np.flatnonzero((own.states[:, :3] == pose).all(1))
    e_ids = near[rng.integers(len(near), size=3 * per_frame)]
    p_ids = same[rng.integers(len(same), size=3 * per_frame)]
    free = arena.capture_value(np.hstack([own.states[p_ids], other.states[e_ids]])) > 0.03
    if free.sum() < per_frame:
        return None
    return pose, p_ids[free][:per_frame], e_ids[free][:per_frame]


def _sensor_pose(rng, pose, arena, role, jitter):
    """The observed pose: a jittered copy of `pose` whose footprint is clear
    (the first of 20 draws), else the pose itself."""
    draws = pose + rng.uniform(-1, 1, (20, 3)) * np.array([jitter, jitter, 0.11])
    clear = arena.clearance(np.hstack([draws, np.zeros((20, 1))]), role) > 1e-4
    return draws[int(np.argmax(clear))] if clear.any() else pose.copy()


def _perturb(rng, states, config, domain, arena, p, e):
    """Continuous query states: half the pairs moved off the grid nodes
    (position +-0.12 m, heading +-0.15 rad, speed +-0.1 m/s, clipped to the
    limits); a move is kept when both cars stay legal and interpolable, the
    pair uncaptured and inside the exit radius."""
    m = len(states)
    moved = states.copy()
    for col, limits in ((0, config.pursuer), (4, config.evader)):
        moved[:, col: col + 2] += rng.uniform(-0.12, 0.12, (m, 2))
        moved[:, col + 2] = wrap_angle(moved[:, col + 2] + rng.uniform(-0.15, 0.15, m))
        moved[:, col + 3] = np.clip(moved[:, col + 3] + rng.uniform(-0.1, 0.1, m),
                                    limits.speed_min, limits.speed_max)
    chosen = rng.random(m) < 0.5
    engines = [GeometryEngine(arena, lim, role, control_dt=config.control_dt)
               for lim, role in ((config.pursuer, "pursuer"), (config.evader, "evader"))]
    keep = engines[0].backup_safe(moved[:, :4]) & engines[1].backup_safe(moved[:, 4:])
    keep &= arena.capture_value(moved) > 0.03
    keep &= np.hypot(*(moved[:, :2] - moved[:, 4:6]).T) < domain.exit_radius - 0.02
    keep &= p.stencil(moved[:, :4])[2] & e.stencil(moved[:, 4:])[2]
    chosen &= keep
    out = states.copy()
    out[chosen] = moved[chosen]
    return out, chosen


def sample_queries(p, e, domain, arena, config, *, frames=192, per_frame=256, seed=1031):
    """Observation frames and query pair states for the labels, with the
    reference's distribution (reference.py:28-124; its draws are not
    reproduced): per frame a pursuer pose drawn uniformly over the pursuer
    states that has at least 30 evader states within exit_radius - 0.1,
    `per_frame` non-captured (payoff > 0.03) pairs at that pose drawn from
    3 * per_frame uniform candidates, a sensor pose jittered by up to
    0.11 m / 0.11 rad with a clear footprint, and a continuous perturbation
    of half the query states kept when legal."""
    rng = np.random.default_rng(seed)
    in_range = (np.hypot(p.states[:, None, 0] - e.states[None, :, 0],
                         p.states[:, None, 1] - e.states[None, :, 1])
                <= domain.exit_radius - 0.1).sum(1)
    pose_ok = in_range >= 30
    sensors, p_all, e_all = [], [], []
    for _ in range(frames):
        for _ in range(1000):
            picked = _pick_frame(rng, p, e, domain, arena, per_frame, pose_ok)
            if picked is not None:
                break
        else:
            raise ValueError("Could not sample a legal observation/query group")
        pose, p_ids, e_ids = picked
        sensors.append([*_sensor_pose(rng, pose, arena, "pursuer", 0.11), 0])
        p_all.append(p_ids)
        e_all.append(e_ids)
    pi = np.concatenate(p_all).astype(np.int32)
    ei = np.concatenate(e_all).astype(np.int32)
    rank = domain.lookup[domain.p_position[pi], ei]
    if (rank < 0).any():
        raise RuntimeError("a query pair lies outside the pair domain")
    nodes = np.hstack([p.states[pi], e.states[ei]])
    states, continuous = _perturb(rng, nodes, config, domain, arena, p, e)
    return dict(frames=np.array(sensors), frame_ids=np.repeat(np.arange(frames, dtype=np.int32),
                                                              per_frame),
                p_indices=pi, e_indices=ei, original_compact_indices=domain.offsets[pi] + rank,
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


def _solve_spec(config, arena, **options):
    """The solve's spec (what the signature covers) and its sha256."""
    spec = dict(config=config.to_dict(), arena=arena.to_dict(), query_version=3)
    extra = ("positions", "half_width", "p_speeds", "e_speeds")
    custom = any(options[k] is not None for k in extra[:2]) or \
        (options["p_speeds"], options["e_speeds"]) != (7, 7)
    spec.update({k: v for k, v in options.items() if k not in extra or custom})
    return spec, hashlib.sha256(json.dumps(spec, sort_keys=True).encode()).hexdigest()


def _atomic_json(path, data):
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=2))
    os.replace(tmp, path)


_META_TEXT = dict(
    schema="local_contact_compact8d_v1",
    history_dtype="float16_sign_preserved",
    backend="hip_gfx950",
    interpolation="multilinear, omit invalid single-vehicle vertices and renormalize",
    action_constraint="continuous obstacle-free command plus one collision-free "
                      "constant-curvature braking backup",
    objective="capture by budget before center separation exceeds local radius; "
              "range exit is local failure, not a physical capture",
    numerical_game="finite-grid semi-Lagrangian minP maxE, exact factorization of "
                   "the product interpolation; no dimension reduction",
)


class _ContactSolve:
    """One contact solve writing into `output`: tables and queries (built, or
    loaded when resuming), the device solver, and the output records."""

    def __init__(self, config, arena, output, spec, signature, opts):
        self.config, self.arena, self.output = config, arena, Path(output)
        self.opts = opts
        self.dt = config.control_dt
        self.meta_path = self.output / "meta.json"
        self.resumed = self.meta_path.exists()
        if self.resumed:
            self.meta = json.loads(self.meta_path.read_text())
            if self.meta["signature"] != signature:
                raise ValueError("Existing output belongs to another model/setup")
            if self.meta["status"] == "complete":
                return
            self.p = VehicleGrid.load(self.output / "p_grid.npz")
            self.e = VehicleGrid.load(self.output / "e_grid.npz")
            self.domain = PairDomain.load(self.output / "domain.npz")
            self.queries = dict(np.load(self.output / "queries.npz"))
        else:
            self._prepare(spec, signature)

    def _prepare(self, spec, signature):
        o = self.opts
        self.p, self.e = build_grids(self.config, self.arena, headings=o["headings"],
                                     position_step=o["position_step"],
                                     positions=o["positions"], half_width=o["half_width"],
                                     p_speeds=o["p_speeds"], e_speeds=o["e_speeds"])
        self.domain = PairDomain.build(self.p, self.e, o["exit_radius"])
        for name, obj in (("p_grid", self.p), ("e_grid", self.e), ("domain", self.domain)):
            obj.save(self.output / f"{name}.npz")
        self.queries = sample_queries(self.p, self.e, self.domain, self.arena, self.config,
                                      frames=o["frames"], per_frame=o["per_frame"],
                                      seed=o["seed"])
        np.savez_compressed(self.output / "queries.npz", **self.queries)
        self.meta = dict(_META_TEXT, status="running", signature=signature, spec=spec,
                         p_states=len(self.p.states), e_states=len(self.e.states),
                         pair_states=self.domain.states,
                         unpruned_pair_states=len(self.p.states) * len(self.e.states),
                         completed_step=0, completed_outputs=0, times=[],
                         update_seconds=0.0)
        self.meta_path.write_text(json.dumps(self.meta, indent=2))

    def check_resources(self, frames_kept):
        """Disk for the history and checkpoint plus an 8 GB reserve; device
        memory of the value fields and the work buffer under max_gb."""
        states = self.domain.states
        history = (2 * frames_kept * states) if self.opts["retain_history"] else 0
        present = sum(f.stat().st_size for f in self.output.iterdir() if f.is_file())
        if shutil.disk_usage(self.output).free < max(0, history + 8 * states - present) + 8e9:
            raise OSError("Insufficient disk space for history/checkpoint transaction plus "
                          "8GB reserve")
        pairs = len(self.p.states) * len(self.e.states)
        self.need = pairs * (9 + 4 * len(self.p.controls))
        if self.need > self.opts["max_gb"] * 1e9:
            raise MemoryError(f"{self.need / 1e9:.2f}GB required; guard {self.opts['max_gb']}GB")

    def open(self):
        for grid, limits, role in ((self.p, self.config.pursuer, "pursuer"),
                                   (self.e, self.config.evader, "evader")):
            grid.bind_geometry(self.arena, limits, role, self.dt)
        self.solver = GameSolver(game_tables(self.p, self.e, self.arena,
                                             self.opts["exit_radius"]))
        q = self.queries["states"]
        ps, es = self.p.stencil(q[:, :4]), self.e.stencil(q[:, 4:])
        if not (ps[2] & es[2]).all():
            raise ValueError("Generated query has no numerical support")
        self.query_stencils = (ps[0], ps[1], es[0], es[1])

    def start(self, warm_start):
        """Initial field and the records already written; returns the first
        step still to run."""
        meta, exit_value = self.meta, self.opts["exit_radius"]
        checkpoint = self.output / meta.get("checkpoint", "missing")
        if self.resumed and meta["completed_outputs"] and checkpoint.exists():
            packed = np.fromfile(checkpoint, np.float32, count=self.domain.states)
            self.solver.initialize(self.domain.unpack(packed, exit_value, len(self.e.states)))
            done = meta["completed_outputs"]
            with np.load(self.output / "labels.npz") as prior:
                self.labels = list(prior["values"][:done])
                self.times = list(prior["times"][:done])
            first = meta["completed_step"] + 1
        else:
            self.solver.initialize()
            if warm_start is not None:
                phi = self.solver.read()
                self.solver.initialize(np.minimum(phi, np.asarray(warm_start, np.float32)))
            self.labels, self.times, first = [], [], 1
        self.history = self.output / "fields.f16"
        if self.opts["retain_history"]:
            with self.history.open("ab" if self.history.exists() else "wb") as f:
                f.truncate(meta["completed_outputs"] * self.domain.states * 2)
        if not self.labels:
            self.record(0)
        return first

    def record(self, step):
        """Labels, history frame and an atomically replaced checkpoint of
        the field after `step` sweeps; meta.json is rewritten last."""
        field = self.solver.read()
        if step == 0:
            label = self.arena.capture_value(self.queries["states"]).astype(np.float32)
        else:
            label = self.solver.query(*self.query_stencils)
        self.labels.append(label)
        self.times.append(step * self.dt)
        packed = self.domain.pack(field)
        if self.opts["retain_history"]:
            with self.history.open("ab") as f:
                write_half_preserving_sign(packed, f)
        np.savez_compressed(self.output / "labels.npz", times=np.array(self.times),
                            values=np.stack(self.labels))
        name = f"checkpoint_{step:04d}.f32"
        partial = self.output / (name + ".partial")
        packed.tofile(partial)
        os.replace(partial, self.output / name)
        previous = self.meta.get("checkpoint")
        self.meta.update(completed_step=step, completed_outputs=len(self.times),
                         times=self.times, checkpoint=name, gpu_working_bytes=self.need,
                         host_working_bytes=self.domain.states * 4)
        _atomic_json(self.meta_path, self.meta)
        if previous not in (None, name) and (self.output / previous).exists():
            (self.output / previous).unlink()
        print(f"budget {step * self.dt:.2f}s; states {self.domain.states:,}; "
              f"updates {self.meta['update_seconds']:.1f}s", flush=True)

    def sweep(self, count, flags):
        before = self.solver.stats()["update_seconds"]
        self.solver.sweep(count, flags)
        self.meta["update_seconds"] += self.solver.stats()["update_seconds"] - before

    def to_residual(self, tol, max_sweeps, flags):
        before = self.solver.stats()["update_seconds"]
        done, converged = self.solver.run(tol, max_sweeps, 8, flags)
        st = self.solver.stats()
        self.meta.update(residual_tol=tol, residual_sweeps=done, residual_converged=converged,
                         final_residual=st["last_residual"],
                         residual_seconds=st["update_seconds"] - before)
        np.asarray(self.domain.pack(self.solver.read()), np.float32).tofile(
            self.output / "converged.f32")

    def finish(self, started):
        self.meta.update(status="complete",
                         wall_seconds_this_process=time.perf_counter() - started)
        self.meta_path.write_text(json.dumps(self.meta, indent=2))
        (self.output / self.meta.pop("checkpoint")).unlink()
        self.meta["saved_bytes"] = sum(f.stat().st_size for f in self.output.iterdir()
                                       if f.is_file())
        self.meta_path.write_text(json.dumps(self.meta, indent=2))
        return self.meta


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
    """Contact-game solve over `horizon` with a record every `output_dt`,
    resumable from its last atomic checkpoint (the reference's solve,
    reference.py:138-369)."""
    if (config.capture_rule != "rectangle_contact"
            or config.boundary_rule != "collision_constrained" or config.point_vehicles):
        raise ValueError("Use the explicit finite-body collision-constrained contact game")
    dt = config.control_dt
    steps, every = round(horizon / dt), round(output_dt / dt)
    whole = np.isclose(steps * dt, horizon) and np.isclose(every * dt, output_dt)
    if min(steps, every) < 1 or not whole:
        raise ValueError("Budgets must be integer control periods")
    output = Path(output)
    output.mkdir(parents=True, exist_ok=True)
    started = time.perf_counter()
    spec, signature = _solve_spec(
        config, arena, headings=headings, position_step=position_step, horizon=horizon,
        output_dt=output_dt, exit_radius=exit_radius, retain_history=retain_history,
        frames=frames, per_frame=per_frame, seed=seed, positions=positions,
        half_width=half_width, p_speeds=p_speeds, e_speeds=e_speeds)
    opts = dict(headings=headings, position_step=position_step, exit_radius=exit_radius,
                retain_history=retain_history, frames=frames, per_frame=per_frame, seed=seed,
                max_gb=max_gb, positions=positions, half_width=half_width, p_speeds=p_speeds,
                e_speeds=e_speeds)
    run = _ContactSolve(config, arena, output, spec, signature, opts)
    if run.resumed and run.meta["status"] == "complete":
        return run.meta
    run.check_resources(len(range(0, steps + 1, every)))
    run.open()
    try:
        step = run.start(warm_start)
        while step <= steps:
            run.sweep(1, flags)
            if step % every == 0 or step == steps:
                run.record(step)
            step += 1
        if residual_tol is not None:
            run.to_residual(residual_tol, max_sweeps, flags)
        run.meta["residual_trace"] = run.solver.trace()[:, 0].tolist()
    finally:
        run.solver.close()
    return run.finish(started)


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
