"""Generate the golden vectors for the 5D relative HJI path.

Runs the reference's own `relative5d` executable, compiled unmodified from
reversible/native_relative.cpp into oracle/_ref by oracle/ref.mk, on

  * C0: configs/reversible_free_a3.json, RelativeGrid(17, 12, 7, 7, 4.0),
    horizon 0.2, output_dt 0.1, order 2, cfl 0.8;
  * the reference unit-test grid RelativeGrid(9, 8, 7, 7, 1.0) with the
    default ReversibleConfig, horizon 0.06, output_dt 0.02, orders 1 and 2
    (tests/test_reversible_free_dp.py:97-129);

and stores its output files (fields, final, arrival, timing.json numbers) as
compressed .npz fixtures next to this script.  Needs /root/reference, so it
only runs in the build container:

    python tests/golden/make_golden.py
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


def main() -> int:
    subprocess.run(["make", "-f", str(ROOT / "oracle" / "ref.mk")], check=True,
                   stdout=subprocess.DEVNULL)
    binary = ROOT / "oracle" / "_ref" / "relative5d"
    for name, case in CASES.items():
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
    return 0


if __name__ == "__main__":
    sys.exit(main())
