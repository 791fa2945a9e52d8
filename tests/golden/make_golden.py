"""Generate the golden vectors for the 5D relative HJI and the dense 8D world
game paths.

Runs the reference's own `relative5d` executable, compiled unmodified from
reversible/native_relative.cpp into oracle/_ref by oracle/ref.mk, on

  * C0: configs/reversible_free_a3.json, RelativeGrid(17, 12, 7, 7, 4.0),
    horizon 0.2, output_dt 0.1, order 2, cfl 0.8;
  * the reference unit-test grid RelativeGrid(9, 8, 7, 7, 1.0) with the
    default ReversibleConfig, horizon 0.06, output_dt 0.02, orders 1 and 2
    (tests/test_reversible_free_dp.py:97-129);

and the reference's `world8d` executable (reversible/native_world.cpp, also
compiled unmodified by oracle/ref.mk) on small dense 8D grids with the
default ReversibleConfig (arena 3 m, extent 3.5 m), RK2 with the linear
speed boundary and RK1 with the legacy signed one; and stores the output
files (fields, final, arrival, timing numbers) as compressed .npz fixtures
next to this script (the first world case also with 24 probe stencils).
Needs /root/reference, so it only runs in the build container:

    python tests/golden/make_golden.py [file-name prefix]
"""

from __future__ import annotations

import json
from pathlib import Path
import subprocess
import sys
import tempfile

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]

# argv flag values (reversible/native_relative.cpp:236-259).
A3 = dict(
    p_min=-1.0, p_max=2.0, e_min=-0.5, e_max=1.0,
    p_acceleration=3.0, e_acceleration=3.0, p_curvature=1.0, e_curvature=2.0,
    radius=0.3,
)
DEFAULT = dict(
    p_min=-1.0, p_max=2.0, e_min=-0.5, e_max=1.0,
    p_acceleration=2.0, e_acceleration=1.0, p_curvature=1.0, e_curvature=2.0,
    radius=0.3,
)
CASES = {
    "c0_a3_17x12x7x7": dict(
        physics=A3, positions=17, headings=12, p_speeds=7, e_speeds=7,
        extent=4.0, horizon=0.2, output_dt=0.1, order=2, cfl=0.8,
    ),
    "unit_9x8x7x7_order1": dict(
        physics=DEFAULT, positions=9, headings=8, p_speeds=7, e_speeds=7,
        extent=1.0, horizon=0.06, output_dt=0.02, order=1, cfl=0.8,
    ),
    "unit_9x8x7x7_order2": dict(
        physics=DEFAULT, positions=9, headings=8, p_speeds=7, e_speeds=7,
        extent=1.0, horizon=0.06, output_dt=0.02, order=2, cfl=0.8,
    ),
}


def run_case(binary: Path, case: dict, out: Path) -> dict:
    flags = dict(case["physics"])
    flags.update(
        {k: case[k] for k in ("positions", "headings", "p_speeds", "e_speeds",
                              "extent", "horizon", "output_dt", "order", "cfl")}
    )
    argv = [str(binary)]
    for k, v in flags.items():
        argv += ["--" + k.replace("_", "-"), repr(v)]
    argv += ["--threads", "4", "--history", "1", "--output", str(out)]
    subprocess.run(argv, check=True, stdout=subprocess.DEVNULL)
    shape = (case["p_speeds"], case["e_speeds"], case["headings"],
             case["positions"], case["positions"])
    n = int(np.prod(shape))
    timing = json.loads((out / "timing.json").read_text())
    fields = np.fromfile(out / "fields.f32", np.float32).reshape(-1, n)
    return dict(
        fields=fields,
        final=np.fromfile(out / "final.f32", np.float32),
        arrival=np.fromfile(out / "arrival.f32", np.float32),
        shape=np.array(shape),
        cfl_dt_limit=timing["cfl_dt_limit"],
        internal_steps=timing["internal_steps"],
        frames=timing["frames"],
        maximum_increase=timing["maximum_increase"],
        argv=np.array(argv[1:]),
        case=json.dumps(case),
    )


# argv flag values of reversible/native_world.cpp:77-104 (default config).
WORLD = dict(p_min=-1.0, p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0, e_accel=1.0,
             p_curvature=1.0, e_curvature=2.0, radius=0.3, arena=3.0, extent=3.5)
WORLD_CASES = {
    "world8d_5x4x3x3_order2_linear": dict(positions=5, headings=4, p_speeds=3, e_speeds=3,
                                          horizon=0.4, output_dt=0.1, order=2, cfl=0.8,
                                          linear_speed_boundary=1, history=1,
                                          probes=24),
    "world8d_5x4x3x3_order1_legacy": dict(positions=5, headings=4, p_speeds=3, e_speeds=3,
                                          horizon=0.4, output_dt=0.1, order=1, cfl=0.8,
                                          linear_speed_boundary=0, history=1),
    "world8d_6x6x4x3_order2_linear": dict(positions=6, headings=6, p_speeds=4, e_speeds=3,
                                          horizon=0.3, output_dt=0.1, order=2, cfl=0.7,
                                          linear_speed_boundary=1, history=0),
}


def probe_states(count: int) -> np.ndarray:
    """Probe states: the two of tests/test_reversible_world_dp.py:127-129,
    then seeded uniform states over the numerical domain (5D speeds inside
    the default limits)."""
    rng = np.random.default_rng(2024)
    fixed = np.array([[-1, 0, 0, 0.5, 0, 1, np.pi, 0.25], [0, 0, 0, 0.5, 0.1, 0.1, 0, 0.25]])
    lo = np.array([-3.5, -3.5, -4.0, -1.0, -3.5, -3.5, -4.0, -0.5])
    hi = np.array([3.5, 3.5, 4.0, 2.0, 3.5, 3.5, 4.0, 1.0])
    return np.concatenate([fixed, rng.uniform(lo, hi, size=(count - 2, 8))])


def run_world(binary: Path, case: dict, out: Path) -> dict:
    case = dict(case)
    probes = case.pop("probes", 0)
    flags = dict(WORLD)
    flags.update(case)
    argv = [str(binary)]
    for k, v in flags.items():
        argv += ["--" + k.replace("_", "-"), repr(v)]
    argv += ["--threads", "4", "--output", str(out)]
    extra = {}
    if probes:
        # Stencils from the oracle's uniform-grid stencil (inputs of the
        # executable; its summation is what these vectors pin).
        sys.path.insert(0, str(ROOT / "pursuit-evasion-core_amd"))
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle
        from pe_amd.reversible.grid_stencil import axis_spec
        from pe_amd.reversible.model import ReversibleConfig
        from pe_amd.reversible.world_dp import PERIODIC_AXES, WorldGrid

        grid = WorldGrid(case["positions"], case["headings"], case["p_speeds"],
                         case["e_speeds"], WORLD["extent"])
        states = probe_states(probes)
        n, lo, hi, step, per, mask = axis_spec(grid.axes(ReversibleConfig()), PERIODIC_AXES)
        ii, ww, valid = oracle.grid_stencil(states, n, lo, hi, step, per, mask, 1e-10)
        assert valid.all()
        ii.astype(np.uint64).tofile(out.parent / "probe_indices.u64")
        ww.astype(np.float64).tofile(out.parent / "probe_weights.f64")
        argv += ["--probe-count", str(probes),
                 "--probe-indices", str(out.parent / "probe_indices.u64"),
                 "--probe-weights", str(out.parent / "probe_weights.f64")]
        extra = dict(probe_states=states, probe_index=ii.astype(np.uint64),
                     probe_weight=ww.astype(np.float64))
    subprocess.run(argv, check=True, stdout=subprocess.DEVNULL)
    n = case["positions"] ** 4 * case["headings"] ** 2 * case["p_speeds"] * case["e_speeds"]
    timing = json.loads((out / "native_timing.json").read_text())
    data = dict(
        final=np.fromfile(out / "final.f32", np.float32),
        arrival=np.fromfile(out / "arrival.f32", np.float32),
        states=timing["states"], internal_steps=timing["internal_steps"],
        frames=timing["frames"], max_internal_dt=timing["max_internal_dt"],
        argv=np.array(argv[1:]), case=json.dumps(flags),
    )
    if case["history"]:
        data["fields"] = np.fromfile(out / "fields.f32", np.float32).reshape(-1, n)
    if probes:
        data.update(extra)
        data["probe_fields"] = np.fromfile(out / "probe_fields.f64",
                                           np.float64).reshape(-1, probes)
    return data


def main() -> int:
    """Writes every fixture, or those whose file name starts with argv[1]."""
    only = sys.argv[1] if len(sys.argv) > 1 else ""
    subprocess.run(["make", "-f", str(ROOT / "oracle" / "ref.mk")], check=True,
                   stdout=subprocess.DEVNULL)
    binary = ROOT / "oracle" / "_ref" / "relative5d"
    for name, case in CASES.items():
        if not f"hji5d_{name}".startswith(only):
            continue
        with tempfile.TemporaryDirectory() as tmp:
            out = Path(tmp) / "out"
            out.mkdir()
            data = run_case(binary, case, out)
        if name.startswith("c0"):
            # C0 is large; the first frame is phi (recomputed by the tests)
            # and the last equals final.f32.
            data["fields"] = data["fields"][1:-1]
        np.savez_compressed(HERE / f"hji5d_{name}.npz", **data)
        print(name, data["internal_steps"], "steps", data["frames"], "frames")
    world = ROOT / "oracle" / "_ref" / "world8d"
    for name, case in WORLD_CASES.items():
        if not name.startswith(only):
            continue
        with tempfile.TemporaryDirectory() as tmp:
            out = Path(tmp) / "out"
            out.mkdir()
            data = run_world(world, case, out)
        np.savez_compressed(HERE / f"{name}.npz", **data)
        print(name, data["internal_steps"], "steps", data["frames"], "frames")
    return 0


if __name__ == "__main__":
    sys.exit(main())
