"""GPU parity of the 5D relative HJI path (through the C ABI).

* float64 local arithmetic: bit-identical to the reference binary's golden
  vectors (fields, final, arrival) on C0 and the reference unit grid.
* float32 local arithmetic (the performance path): against the same golden
  vectors and the FP64-storage oracle.  Tolerance: max|dV| <= 1e-3 and
  mean|dV| <= 1e-5.  The scheme's ENO2 selection (strict "<") turns rounding
  into O(h * second difference) jumps at exact ties, 4e-4 at most on C0 even
  between float and double storage of the CPU reference itself.  Capture-set
  membership (V <= 0) must agree except at nodes whose reference value lies
  within that tolerance of zero; those nodes are printed.
* C1 at full size (34,222,500 nodes, 810 RK2 steps): size-independent
  properties (monotone in time, V <= phi, arrival set == {V <= 0}, mirror
  symmetry) plus a bounded-sample comparison with the oracle.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from pe_amd import _native
from pe_amd.reversible.free_dp import RelativeGrid, native_config, solve
from pe_amd.reversible.model import ReversibleConfig, VehicleLimits

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
TOL_MAX, TOL_MEAN = 1e-3, 1e-5


def config_from_case(case, radius=None):
    ph = case["physics"]
    return ReversibleConfig(
        pursuer=VehicleLimits(ph["p_min"], ph["p_max"], ph["p_acceleration"], ph["p_curvature"]),
        evader=VehicleLimits(ph["e_min"], ph["e_max"], ph["e_acceleration"], ph["e_curvature"]),
        capture_radius=ph["radius"] if radius is None else radius,
    )


def load_case(name):
    z = np.load(GOLDEN / f"hji5d_{name}.npz")
    case = json.loads(str(z["case"]))
    grid = RelativeGrid(case["positions"], case["headings"], case["p_speeds"],
                        case["e_speeds"], case["extent"])
    return z, case, grid


def run(tmp_path, name, precision):
    z, case, grid = load_case(name)
    print(f"[gpu] 5D solve {name} {precision}", flush=True)
    meta = solve(config_from_case(case), grid, tmp_path / name, horizon=case["horizon"],
                 output_dt=case["output_dt"], order=case["order"], cfl=case["cfl"],
                 local_precision=precision)
    out = tmp_path / name
    n = grid.states
    fields = np.fromfile(out / "fields.f32", np.float32).reshape(-1, n)
    final = np.fromfile(out / "final.f32", np.float32)
    arrival = np.fromfile(out / "arrival.f32", np.float32)
    return z, case, meta, fields, final, arrival


def u32(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", ["c0_a3_17x12x7x7", "unit_9x8x7x7_order1", "unit_9x8x7x7_order2"])
def test_float64_local_is_bit_exact_with_reference_binary(gpu, tmp_path, name):
    z, case, meta, fields, final, arrival = run(tmp_path, name, "float64")
    assert meta["status"] == "complete"
    assert meta["internal_steps"] == int(z["internal_steps"])
    assert meta["frames"] == int(z["frames"])
    assert meta["maximum_increase"] == float(z["maximum_increase"])
    np.testing.assert_array_equal(u32(final), u32(z["final"]))
    stored = fields[1:-1] if name.startswith("c0") else fields
    np.testing.assert_array_equal(u32(stored), u32(z["fields"]))
    np.testing.assert_array_equal(np.isnan(arrival), np.isnan(z["arrival"]))
    np.testing.assert_array_equal(u32(np.nan_to_num(arrival)), u32(np.nan_to_num(z["arrival"])))


def capture_disagreement(v, ref, tol):
    """Nodes whose capture membership (V <= 0) differs; all must lie within
    `tol` of the boundary in the reference."""
    diff = np.flatnonzero((v <= 0) != (ref <= 0))
    near = np.abs(ref[diff]) <= tol
    return diff, near


@pytest.mark.parametrize("name", ["c0_a3_17x12x7x7", "unit_9x8x7x7_order2"])
def test_float32_local_within_tolerance_of_reference_binary(gpu, tmp_path, name):
    z, case, meta, fields, final, arrival = run(tmp_path, name, "float32")
    ref = z["final"].astype(np.float64)
    d = np.abs(final - ref)
    print(f"[{name}] float32 vs reference: max|dV|={d.max():.3e} mean|dV|={d.mean():.3e}")
    assert d.max() <= TOL_MAX and d.mean() <= TOL_MEAN
    diff, near = capture_disagreement(final, ref, TOL_MAX)
    print(f"[{name}] capture-set disagreements: {len(diff)} (listed): {diff.tolist()[:50]}")
    assert near.all()
    assert meta["maximum_increase"] <= 1e-6
    assert np.diff(fields, axis=0).max() <= 1e-6


def a3(radius=0.3):
    return ReversibleConfig(
        name="reversible_free_a3",
        pursuer=VehicleLimits(-1.0, 2.0, 3.0, 1.0),
        evader=VehicleLimits(-0.5, 1.0, 3.0, 2.0),
        capture_radius=radius,
    )


def test_float32_against_fp64_storage_oracle_mid_grid(gpu, tmp_path, oracle_lib):
    """A grid larger than the golden ones: GPU FP32 vs the FP64-storage
    oracle (double everywhere), states the FP32 error."""
    grid = RelativeGrid(33, 24, 9, 9, 4.0)
    cfg = oracle_lib.HJI5DCfg(33, 24, 9, 9, 2, 16, 1, 4.0, 0.8, 0.3,
                              -1.0, 2.0, -0.5, 1.0, 3.0, 3.0, 1.0, 2.0)
    ref = oracle_lib.hji5d_solve(cfg, 0.5, 0.25, history=False)
    print("[gpu] 5D mid grid float32", flush=True)
    meta = solve(a3(), grid, tmp_path / "mid", horizon=0.5, output_dt=0.25, history=False)
    final = np.fromfile(tmp_path / "mid" / "final.f32", np.float32)
    d = np.abs(final - ref["final"])
    print(f"[33x24x9x9] float32 vs fp64 oracle: max|dV|={d.max():.3e} mean|dV|={d.mean():.3e}")
    assert meta["internal_steps"] == ref["internal_steps"]
    assert d.max() <= TOL_MAX and d.mean() <= TOL_MEAN
    diff, near = capture_disagreement(final, ref["final"], TOL_MAX)
    print(f"capture-set disagreements: {len(diff)}: {diff.tolist()[:50]}")
    assert near.all()


C1 = RelativeGrid(65, 36, 15, 15, 8.0)


def test_c1_full_size_properties(gpu, oracle_lib):
    """34.2M nodes, horizon 3.0 s (810 RK2 steps): FP32 local arithmetic
    against FP64 local arithmetic on the same GPU kernel (the FP64 mode is
    bit-identical to the oracle; see the sample test below), plus
    size-independent properties.  The scheme is mirror-symmetric only up to
    ENO2 tie-breaking (strict "<"), which drifts over 810 steps: the mean
    asymmetry is bounded, its maximum is printed."""
    print("[gpu] C1 34.2M nodes, 810 steps, float32 and float64", flush=True)
    fields, asym_max = {}, {}
    for precision in ("float64", "float32"):
        cfg = native_config(a3(0.5), C1, order=2, cfl=0.8, local_precision=precision)
        with _native.HJI5D(cfg) as h:
            phi, _ = h.read()
            for k in range(1, 31):
                h.advance_to(min(3.0, k * 0.1))
            st = h.stats()
            final, arrival = h.read()
            trace = h.trace()
        assert st["internal_steps"] == 810
        assert st["time"] == pytest.approx(3.0, abs=1e-12)
        assert st["maximum_increase"] <= 1e-6
        assert (final <= phi + 1e-6).all()
        assert (final[phi <= 0] <= 0).all()
        np.testing.assert_array_equal(np.isfinite(arrival), final <= 0)
        assert len(trace) == 810 and np.all(trace >= 0)
        fields[precision] = final
        f = final.reshape(C1.shape)
        nh = C1.headings
        asym = np.abs(f - f[:, :, (-np.arange(nh)) % nh][:, :, :, ::-1, :])
        print(f"C1 {precision}: mirror asymmetry max={asym.max():.3e} mean={asym.mean():.3e}; "
              f"captured fraction {np.mean(final <= 0):.4f}; last residual {trace[-1]:.3e}")
        assert asym.mean() < 1e-4
        asym_max[precision] = float(asym.max())
    d = np.abs(fields["float32"].astype(np.float64) - fields["float64"])
    flips = np.flatnonzero((fields["float32"] <= 0) != (fields["float64"] <= 0))
    print(f"C1 float32 vs float64 local arithmetic after 810 steps: max|dV|={d.max():.3e} "
          f"mean|dV|={d.mean():.3e}; capture-set flips {len(flips)} "
          f"(|V64| max {np.abs(fields['float64'][flips]).max() if len(flips) else 0:.2e})")
    # every flip by node: storage index (vP, vE, psi, y, x), both values;
    # each lies within one grid cell of the FP64 capture boundary (a grid
    # neighbour along some axis, psi periodic, is on the other side)
    f64 = fields["float64"].reshape(C1.shape)

    def on_boundary(node):
        inside = f64[node] <= 0
        for axis in range(5):
            for step in (-1, 1):
                nb = list(node)
                nb[axis] += step
                if axis == 2:
                    nb[2] %= C1.headings
                elif not 0 <= nb[axis] < C1.shape[axis]:
                    continue
                if (f64[tuple(nb)] <= 0) != inside:
                    return True
        return False

    for flat in flips:
        node = tuple(int(i) for i in np.unravel_index(flat, C1.shape))
        print(f"  flip {node}: V32={fields['float32'][flat]:+.3e} "
              f"V64={fields['float64'][flat]:+.3e}, on the FP64 boundary: {on_boundary(node)}")
        assert on_boundary(node)
    # The FP32 error is bounded by the scheme's own tie-breaking ambiguity
    # (the FP64 field's mirror asymmetry) and is small on average.
    assert d.max() <= 2 * asym_max["float64"] + 1e-3 and d.mean() <= 1e-4
    assert len(flips) == 0 or np.abs(fields["float64"][flips]).max() <= 1e-3


def test_c1_bounded_sample_against_oracle(gpu, oracle_lib):
    """Three full RK2 steps of C1 from phi: GPU FP32 vs the oracle (FP64
    local arithmetic, float storage) on all 34.2M nodes."""
    ocfg = oracle_lib.HJI5DCfg(65, 36, 15, 15, 2, 16, 0, 8.0, 0.8, 0.5,
                               -1.0, 2.0, -0.5, 1.0, 3.0, 3.0, 1.0, 2.0)
    field, arrival = oracle_lib.hji5d_init(ocfg)
    dt = oracle_lib.hji5d_cfl_limit(ocfg)
    oracle_lib.hji5d_time_steps(ocfg, field, arrival, dt, 3)
    cfg = native_config(a3(0.5), C1, order=2, cfl=0.8)
    print("[gpu] C1 3 steps vs oracle", flush=True)
    with _native.HJI5D(cfg) as h:
        assert h.stats()["cfl_dt_limit"] == pytest.approx(dt, rel=1e-15)
        h.step(dt, 3)
        gpu_field, gpu_arrival = h.read()
    d = np.abs(gpu_field.astype(np.float64) - field)
    print(f"C1 3 steps: max|dV|={d.max():.3e} mean|dV|={d.mean():.3e}")
    assert d.max() <= TOL_MAX and d.mean() <= 1e-6
    np.testing.assert_array_equal(np.isfinite(gpu_arrival), np.isfinite(arrival))


def test_float64_local_matches_oracle_bitwise_on_c1_sample(gpu, oracle_lib):
    ocfg = oracle_lib.HJI5DCfg(65, 36, 15, 15, 2, 16, 0, 8.0, 0.8, 0.5,
                               -1.0, 2.0, -0.5, 1.0, 3.0, 3.0, 1.0, 2.0)
    field, arrival = oracle_lib.hji5d_init(ocfg)
    dt = oracle_lib.hji5d_cfl_limit(ocfg)
    oracle_lib.hji5d_time_steps(ocfg, field, arrival, dt, 2)
    cfg = native_config(a3(0.5), C1, order=2, cfl=0.8, local_precision="float64")
    print("[gpu] C1 2 steps float64 vs oracle", flush=True)
    with _native.HJI5D(cfg) as h:
        h.step(dt, 2)
        gpu_field, _ = h.read()
    np.testing.assert_array_equal(u32(gpu_field), u32(field))


def test_residual_iteration_past_horizon(gpu):
    cfg = native_config(a3(), RelativeGrid(17, 12, 7, 7, 4.0), order=2, cfl=0.8)
    print("[gpu] residual iteration", flush=True)
    with _native.HJI5D(cfg) as h:
        h.advance_to(0.2)
        converged = h.run_to_residual(1e-4, 20000, 32)
        st = h.stats()
        trace = h.trace()
    assert converged and st["last_residual"] <= 1e-4
    assert len(trace) == st["internal_steps"]


def test_exit_radius_and_rectangle_capture_initialisation(gpu):
    from pe_amd.local_game.geometry import Footprint, rectangle_separation

    grid = RelativeGrid(17, 12, 7, 7, 4.0)
    rect = ReversibleConfig(point_vehicles=False, capture_rule="rectangle_contact",
                            boundary_rule="collision_constrained")
    feet = (Footprint(0.5, 0.3), Footprint(0.5, 0.3))
    cfg = native_config(rect, grid, exit_radius=3.0, capture_footprints=feet)
    print("[gpu] rectangle capture + exit radius", flush=True)
    with _native.HJI5D(cfg) as h:
        phi, arrival = h.read()
        h.advance_to(0.1)
        after, _ = h.read()
    _, _, psi, y, x = np.meshgrid(*grid.axes(rect), indexing="ij")
    e = np.stack([x, y, psi], -1)
    expect = rectangle_separation(np.zeros(3), e, *feet)
    expect = np.where(x * x + y * y > 9.0, 3.0, expect).astype(np.float32).ravel()
    np.testing.assert_allclose(phi, expect, atol=1e-6)
    outside = (x * x + y * y > 9.0).ravel()
    assert (after[outside] == 3.0).all()
    assert (after <= phi + 1e-6).all()


def test_profiling_counts_launches(gpu):
    cfg = native_config(a3(), RelativeGrid(17, 12, 7, 7, 4.0))
    with _native.HJI5D(cfg) as h:
        h.set_profiling(True)
        h.step(0.01, 5)
        st = h.stats()
    assert st["stage_launches"] == [5, 5]
    assert all(ms > 0 for ms in st["stage_ms"])


@pytest.mark.parametrize("grid,steps,order", [(RelativeGrid(17, 12, 7, 7, 4.0), 18, 2),
                                              (RelativeGrid(33, 24, 9, 9, 4.0), 12, 2),
                                              (RelativeGrid(65, 36, 15, 15, 8.0), 3, 2),
                                              (RelativeGrid(33, 24, 9, 9, 4.0), 12, 1),
                                              (RelativeGrid(17, 7, 7, 7, 4.0), 18, 2),
                                              (RelativeGrid(17, 4, 7, 7, 4.0), 18, 2),
                                              (RelativeGrid(17, 6, 5, 8, 4.0), 12, 2),
                                              (RelativeGrid(17, 6, 5, 8, 4.0), 12, 1)])
def test_fp32_kernel_variants_are_bit_identical(gpu, grid, steps, order):
    """The FP32 kernel's launch variants do the same arithmetic on the same
    neighbour values: two psi planes per thread (even heading counts; with
    four, the planes at -2 and +2 coincide, as do -1 and +3) or one,
    the next vP node's neighbours requested early or not, and either block
    order.  Fields and arrival times must agree bit for bit (the generic
    template, a different operation order, is compared by tolerance in the
    oracle tests)."""
    single = _native.PE_HJI5D_PSI_SINGLE
    variants = {
        "default": 0,
        "psi_outer": _native.PE_HJI5D_PSI_OUTER,
        "single": single,
        "single_no_prefetch": single | _native.PE_HJI5D_NO_PREFETCH,
        "single_psi_outer": single | _native.PE_HJI5D_PSI_OUTER,
        "single_no_prefetch_psi_outer":
            single | _native.PE_HJI5D_NO_PREFETCH | _native.PE_HJI5D_PSI_OUTER,
    }
    out = {}
    for name, flags in variants.items():
        cfg = native_config(a3(0.5), grid, order=order, cfl=0.8)
        cfg.flags = flags
        print(f"[gpu] variant {name} on {grid.positions}x{grid.headings}x{grid.p_speeds}x"
              f"{grid.e_speeds}", flush=True)
        with _native.HJI5D(cfg) as h:
            dt = h.stats()["cfl_dt_limit"]
            h.step(dt, steps)
            out[name] = h.read()
    f0, a0 = out["default"]
    for name, (f, arr) in out.items():
        np.testing.assert_array_equal(f.view(np.uint32), f0.view(np.uint32), err_msg=name)
        np.testing.assert_array_equal(np.isnan(arr), np.isnan(a0), err_msg=name)
        np.testing.assert_array_equal(np.nan_to_num(arr).view(np.uint32),
                                      np.nan_to_num(a0).view(np.uint32), err_msg=name)


@pytest.mark.parametrize("radius", [2.3, float(np.nextafter(2.5, 0.0))])
def test_exit_radius_float64_bit_exact_with_oracle(gpu, oracle_lib, radius):
    """The local exit clamp (gpu_relative.py:126-131, 167-169) in FP64 local
    arithmetic: bit for bit with the oracle's exit mode over 8 RK2 steps on
    a 33-point grid.  2.3: the kernel test uses the literal 5.29, not
    r * r = 5.289999999999999.  nextafter(2.5, 0): r * r < 6.25 = the
    literal, so the nodes with x^2 + y^2 = 6.25 start at the exit value
    (the initial plane compares with r * r) and are then updated (the kernel
    compares with the literal)."""
    grid = RelativeGrid(33, 12, 7, 7, 4.0)
    ocfg = oracle_lib.HJI5DCfg(33, 12, 7, 7, 2, 16, 0, 4.0, 0.8, 0.3, -1.0, 2.0, -0.5, 1.0,
                               3.0, 3.0, 1.0, 2.0, radius)
    field, arrival = oracle_lib.hji5d_init(ocfg)
    dt = oracle_lib.hji5d_cfl_limit(ocfg)
    cfg = native_config(a3(0.3), grid, order=2, cfl=0.8, local_precision="float64",
                        exit_radius=radius)
    print(f"[gpu] exit radius {radius!r} float64 vs oracle", flush=True)
    with _native.HJI5D(cfg) as h:
        first, _ = h.read()
        np.testing.assert_array_equal(u32(first), u32(field))
        h.step(dt, 8)
        gpu_field, gpu_arrival = h.read()
    oracle_lib.hji5d_time_steps(ocfg, field, arrival, dt, 8)
    np.testing.assert_array_equal(u32(gpu_field), u32(field))
    np.testing.assert_array_equal(np.isnan(gpu_arrival), np.isnan(arrival))
    x = -4.0 + np.arange(33) * (8.0 / 32)
    d = (x[None, :] ** 2 + x[:, None] ** 2)
    on_circle = d == 6.25
    if 2.4 < radius < 2.5:
        assert on_circle.any() and (first.reshape(-1, 33, 33)[:, on_circle] == 2.5).all()


def test_exit_radius_float32_uses_float_arithmetic(gpu, oracle_lib):
    """FP32 mode: the exit test is the reference's float one (x = -W +
    ix * dx with W and dx = 2 W / (n - 1) in float, x*x + y*y > float
    literal of r^2), checked against numpy float32 arithmetic on a grid
    whose spacing is not a binary fraction; exit nodes hold the float
    literal of r; elsewhere the field is within the FP32 tolerance of the
    oracle's float64 exit mode (where the two exit sets agree)."""
    n, extent, radius = 37, 3.7, 2.9
    grid = RelativeGrid(n, 12, 7, 7, extent)
    cfg = native_config(a3(0.3), grid, order=2, cfl=0.8, exit_radius=radius)
    ocfg = oracle_lib.HJI5DCfg(n, 12, 7, 7, 2, 16, 0, extent, 0.8, 0.3, -1.0, 2.0, -0.5, 1.0,
                               3.0, 3.0, 1.0, 2.0, radius)
    field, arrival = oracle_lib.hji5d_init(ocfg)
    dt = oracle_lib.hji5d_cfl_limit(ocfg)
    print("[gpu] exit radius float32", flush=True)
    with _native.HJI5D(cfg) as h:
        h.step(dt, 6)
        got, _ = h.read()
    oracle_lib.hji5d_time_steps(ocfg, field, arrival, dt, 6)
    W = np.float32(extent)
    dx = (np.float32(2.0) * W) / np.float32(n - 1)
    xf = -W + np.arange(n, dtype=np.float32) * dx
    out32 = (xf[None, :] * xf[None, :] + xf[:, None] * xf[:, None]) > np.float32(
        f"{radius * radius:.9e}")
    x64 = -extent + np.arange(n) * (2 * extent / (n - 1))
    out64 = (x64[None, :] ** 2 + x64[:, None] ** 2) > float(f"{radius * radius:.9e}")
    g = got.reshape(-1, n, n)
    assert (g[:, out32] == np.float32(f"{radius:.9e}")).all()
    agree = out32 == out64
    print(f"exit sets: float32 {out32.sum()}, float64 {out64.sum()}, differ at "
          f"{np.argwhere(~agree).tolist()}")
    inside = agree & ~out32
    o = field.reshape(-1, n, n)
    d = np.abs(g[:, inside].astype(np.float64) - o[:, inside])
    assert d.max() <= 1e-3 and d.mean() <= 1e-5


def test_rectangle_capture_float64_bit_exact_with_oracle(gpu, oracle_lib):
    """Rectangle-contact capture (gpu_relative.py:156-160) changes only the
    initial plane, whose closed form the initialisation test checks; the
    time steps do not read it again.  From the GPU's own initial field,
    8 FP64 RK2 steps with the exit clamp at r = 2.9 equal the oracle's
    steps bit for bit, fields and arrival times."""
    from pe_amd.local_game.geometry import Footprint

    grid = RelativeGrid(33, 12, 7, 7, 4.0)
    rect = ReversibleConfig(point_vehicles=False, capture_rule="rectangle_contact",
                            boundary_rule="collision_constrained")
    feet = (Footprint(0.8, 0.5), Footprint(0.4, 0.25))
    cfg = native_config(rect, grid, order=2, cfl=0.8, local_precision="float64",
                        exit_radius=2.9, capture_footprints=feet)
    ocfg = oracle_lib.HJI5DCfg(
        cfg.positions, cfg.headings, cfg.p_speeds, cfg.e_speeds, cfg.order, 16, 0,
        cfg.extent, cfg.cfl, cfg.radius, cfg.p_min, cfg.p_max, cfg.e_min, cfg.e_max,
        cfg.p_acceleration, cfg.e_acceleration, cfg.p_curvature, cfg.e_curvature, 2.9)
    dt = oracle_lib.hji5d_cfl_limit(ocfg)
    print("[gpu] rectangle capture + exit radius 2.9, float64 vs oracle", flush=True)
    with _native.HJI5D(cfg) as h:
        assert h.stats()["cfl_dt_limit"] == pytest.approx(dt, rel=1e-14)
        phi, arrival0 = h.read()
        h.step(dt, 8)
        gpu_field, gpu_arrival = h.read()
    field, arrival = phi.copy(), arrival0.copy()
    np.testing.assert_array_equal(np.isnan(arrival), phi > 0)
    oracle_lib.hji5d_time_steps(ocfg, field, arrival, dt, 8)
    assert (field != phi).any()
    np.testing.assert_array_equal(u32(gpu_field), u32(field))
    np.testing.assert_array_equal(np.isnan(gpu_arrival), np.isnan(arrival))
    np.testing.assert_array_equal(u32(np.nan_to_num(gpu_arrival)), u32(np.nan_to_num(arrival)))


@pytest.mark.parametrize("feet", [None, "rect"])
def test_two_plane_kernel_with_exit_radius(gpu, feet):
    """The two-plane FP32 kernel (default for even heading counts) against
    the one-plane kernel with the exit clamp, with and without rectangle
    capture: fields and arrival times bit for bit."""
    from pe_amd.local_game.geometry import Footprint

    grid = RelativeGrid(17, 12, 7, 7, 4.0)
    if feet:
        config = ReversibleConfig(point_vehicles=False, capture_rule="rectangle_contact",
                                  boundary_rule="collision_constrained")
        fp = (Footprint(0.5, 0.3), Footprint(0.5, 0.3))
    else:
        config, fp = a3(0.3), None
    out = {}
    for name, flags in (("pair", 0), ("single", _native.PE_HJI5D_PSI_SINGLE)):
        cfg = native_config(config, grid, exit_radius=3.0, capture_footprints=fp)
        cfg.flags = flags
        print(f"[gpu] exit radius variants {name} {feet}", flush=True)
        with _native.HJI5D(cfg) as h:
            h.step(h.stats()["cfl_dt_limit"], 18)
            out[name] = h.read()
    (f0, a0), (f1, a1) = out["pair"], out["single"]
    np.testing.assert_array_equal(u32(f0), u32(f1))
    np.testing.assert_array_equal(np.isnan(a0), np.isnan(a1))
    np.testing.assert_array_equal(u32(np.nan_to_num(a0)), u32(np.nan_to_num(a1)))
