"""Dense 8D square game on the GPU (pe_world8d through the C ABI) against
the reference executable's golden vectors and the oracle, bit for bit, and
the drop-in `world_dp.solve` / `WorldTable` behaviour
(tests/test_reversible_world_dp.py)."""

import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["world8d_5x4x3x3_order2_linear", "world8d_5x4x3x3_order1_legacy",
         "world8d_6x6x4x3_order2_linear"]


def bits(a):
    return np.nan_to_num(np.asarray(a, np.float32)).view(np.uint32)


def native_cfg(case):
    from pe_amd import _native

    c = _native.World8DConfig()
    for k in ("positions", "headings", "p_speeds", "e_speeds", "order",
              "linear_speed_boundary", "extent", "arena", "radius", "p_min", "p_max", "e_min",
              "e_max", "p_accel", "e_accel", "p_curvature", "e_curvature", "cfl"):
        setattr(c, k, case[k])
    return c


def oracle_cfg(oracle_lib, case, threads=16):
    return oracle_lib.World8DCfg(
        case["positions"], case["headings"], case["p_speeds"], case["e_speeds"], case["order"],
        case["linear_speed_boundary"], threads, case["extent"], case["arena"], case["radius"],
        case["p_min"], case["p_max"], case["e_min"], case["e_max"], case["p_accel"],
        case["e_accel"], case["p_curvature"], case["e_curvature"], case["cfl"])


def run_frames(case, probes=None, flags=0):
    """Frames at 0, output_dt, ..., horizon, final, arrival, stats (and the
    probe values per frame) of the GPU solver, stepped as the executable's
    frame loop (native_world.cpp:317-333)."""
    from pe_amd import _native

    cfg = native_cfg(case)
    cfg.flags = flags
    with _native.World8D(cfg) as w:
        if probes is not None:
            w.set_probes(*probes)
        frames = [w.read(True, False)[0]]
        values = [w.probe()] if probes is not None else []
        count = int(np.ceil(case["horizon"] / case["output_dt"] - 1e-10))
        for frame in range(1, count + 1):
            w.advance_to(min(case["horizon"], frame * case["output_dt"]))
            frames.append(w.read(True, False)[0])
            if probes is not None:
                values.append(w.probe())
        final, arrival = w.read()
        assert w.count_nonfinite() == 0
        return np.stack(frames), final, arrival, w.stats(), values


@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_reference_world_binary(gpu, name):
    print(f"[world8d] golden {name}", flush=True)
    z = np.load(GOLDEN / f"{name}.npz")
    case = json.loads(str(z["case"]))
    probes = None
    if "probe_index" in z.files:
        probes = (z["probe_index"], z["probe_weight"])
    frames, final, arrival, st, values = run_frames(case, probes)
    assert st["internal_steps"] == int(z["internal_steps"])
    assert st["max_internal_dt"] == pytest.approx(float(z["max_internal_dt"]), rel=1e-12)
    np.testing.assert_array_equal(bits(final), bits(z["final"]))
    np.testing.assert_array_equal(np.isnan(arrival), np.isnan(z["arrival"]))
    np.testing.assert_array_equal(bits(arrival), bits(z["arrival"]))
    if case["history"]:
        np.testing.assert_array_equal(bits(frames), bits(z["fields"]))
    if probes is not None:
        np.testing.assert_array_equal(np.stack(values), z["probe_fields"])


@pytest.mark.parametrize("shape, order, linear, horizon", [
    ((7, 6, 5, 4), 2, 1, 0.25),
    ((8, 5, 4, 5), 1, 0, 0.2),
    ((9, 8, 3, 3), 2, 0, 0.15),
])
def test_gpu_matches_oracle_on_ragged_grids(gpu, oracle_lib, shape, order, linear, horizon):
    """Shapes whose slabs are not multiples of the 256-thread block, odd and
    even position counts, both orders and speed boundaries."""
    print(f"[world8d] oracle {shape} order {order} linear {linear}", flush=True)
    n, h, np_, ne = shape
    case = dict(positions=n, headings=h, p_speeds=np_, e_speeds=ne, order=order,
                linear_speed_boundary=linear, extent=3.5, arena=3.0, radius=0.3, p_min=-1.0,
                p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0, e_accel=1.0, p_curvature=1.0,
                e_curvature=2.0, cfl=0.8, horizon=horizon, output_dt=horizon / 2)
    frames, final, arrival, st, _ = run_frames(case)
    o_frames, o_final, o_arrival, o_steps, o_dt = oracle_lib.world8d_solve(
        oracle_cfg(oracle_lib, case), horizon, horizon / 2)
    assert st["internal_steps"] == o_steps
    np.testing.assert_array_equal(bits(frames), bits(o_frames))
    np.testing.assert_array_equal(bits(final), bits(o_final))
    np.testing.assert_array_equal(bits(arrival), bits(o_arrival))


@pytest.mark.parametrize("shape", [(7, 6, 5, 4), (9, 8, 3, 3), (6, 5, 3, 4)])
def test_kernel_variants_are_bit_identical(gpu, shape):
    """The LDS-slab kernel (256, 512 and 1024 threads per slab) and the
    thread-per-node kernel give the same bits, RK2 and RK1, both speed
    boundaries."""
    from pe_amd import _native

    print(f"[world8d] variants {shape}", flush=True)
    n, h, np_, ne = shape
    for order, linear in ((2, 1), (1, 0)):
        case = dict(positions=n, headings=h, p_speeds=np_, e_speeds=ne, order=order,
                    linear_speed_boundary=linear, extent=3.5, arena=3.0, radius=0.3,
                    p_min=-1.0, p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0, e_accel=1.0,
                    p_curvature=1.0, e_curvature=2.0, cfl=0.8, horizon=0.3, output_dt=0.1)
        runs = {}
        for flags in (0, _native.PE_WORLD8D_GENERIC, _native.PE_WORLD8D_SLAB_256,
                      _native.PE_WORLD8D_SLAB_1024):
            frames, final, arrival, st, _ = run_frames(case, flags=flags)
            runs[st["kernel"]] = (frames, arrival)
        assert sorted(runs) == [0, 256, 512, 1024]
        for k in (256, 512, 1024):
            np.testing.assert_array_equal(bits(runs[k][0]), bits(runs[0][0]))
            np.testing.assert_array_equal(bits(runs[k][1]), bits(runs[0][1]))


@pytest.mark.parametrize("shape", [(13, 12, 3, 7), (11, 16, 3, 5)])
def test_split_slabs_are_bit_identical(gpu, shape):
    """Slabs above 36 KiB are staged in parts of xE rows (4 and 2 parts
    here): two RK2 steps give the bits of the thread-per-node kernel."""
    from pe_amd import _native

    print(f"[world8d] split slabs {shape}", flush=True)
    n, h, np_, ne = shape
    case = dict(positions=n, headings=h, p_speeds=np_, e_speeds=ne, order=2,
                linear_speed_boundary=1, extent=3.5, arena=3.0, radius=0.3, p_min=-1.0,
                p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0, e_accel=1.0, p_curvature=1.0,
                e_curvature=2.0, cfl=0.8)
    out = {}
    for flags in (0, _native.PE_WORLD8D_GENERIC):
        cfg = native_cfg(case)
        cfg.flags = flags
        with _native.World8D(cfg) as w:
            w.advance_to(2 * w.stats()["max_internal_dt"])
            out[w.stats()["kernel"]] = w.read()
    assert sorted(out) == [0, 512]
    for k in (0, 1):
        np.testing.assert_array_equal(bits(out[512][k]), bits(out[0][k]))


def test_reciprocal_division_is_exact_for_the_grid_spacings(gpu):
    """The kernels divide by a spacing h or its square hh through two
    correction steps on RN(1/h) while the field stays in range; every float
    numerator with exponent in [-90, 70) of either sign (and +0), which
    covers the [2^-83, 2^64] the range flag guarantees, gives the IEEE
    quotient bit for bit, for the spacings of every grid tested here."""
    from pe_amd import _native

    print("[world8d] exhaustive division check", flush=True)
    divisors = set()
    for n, h, np_, ne, extent in ((9, 8, 7, 7, 3.5), (13, 12, 7, 7, 3.5), (5, 4, 3, 3, 3.5),
                                  (6, 6, 4, 3, 3.5), (7, 6, 5, 4, 3.5), (8, 5, 4, 5, 3.5),
                                  (9, 8, 3, 3, 3.5), (5, 4, 4, 4, 3.5), (6, 5, 3, 4, 3.5)):
        dx = np.float32(2) * np.float32(extent) / np.float32(n - 1)
        dh = np.float32(2 * np.pi / h)
        dvp = (np.float32(2.0) - np.float32(-1.0)) / np.float32(np_ - 1)
        dve = (np.float32(1.0) - np.float32(-0.5)) / np.float32(ne - 1)
        for d in (dx, dh, dvp, dve):
            divisors.update((float(d), float(np.float32(d) * np.float32(d))))
    divisors = np.array(sorted(divisors), np.float32)
    bad = _native.world8d_check_division(divisors, -90, 70)
    assert bad.tolist() == [0] * len(divisors), dict(zip(divisors.tolist(), bad.tolist()))


def test_out_of_range_field_takes_the_ieee_division(gpu, oracle_lib):
    """A continuation field holding a -0, a 1e-30 and a 1e30 value (outside
    the fast division's range) still gives the oracle's bits."""
    from pe_amd import _native

    print("[world8d] out-of-range continuation", flush=True)
    case = dict(positions=6, headings=4, p_speeds=3, e_speeds=4, order=2,
                linear_speed_boundary=1, extent=3.5, arena=3.0, radius=0.3, p_min=-1.0,
                p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0, e_accel=1.0, p_curvature=1.0,
                e_curvature=2.0, cfl=0.8)
    ocfg = oracle_cfg(oracle_lib, case)
    field, arrival = oracle_lib.world8d_init(ocfg)
    field[[5, 77, 901]] = np.array([-0.0, 1e-30, 1e30], np.float32)
    arrival[:] = np.nan
    with _native.World8D(native_cfg(case)) as w:
        w.set_field(field, arrival, 0.0)
        dt = w.stats()["max_internal_dt"]
        w.advance_to(2 * dt)
        final, arr = w.read()
    _, o_final, o_arrival, steps, _ = oracle_lib.world8d_solve(
        ocfg, 2 * dt, 2 * dt, field=field, arrival=arrival, history=False)
    assert steps == 2
    np.testing.assert_array_equal(bits(final), bits(o_final))
    np.testing.assert_array_equal(bits(arr), bits(o_arrival))


def test_gpu_default_grid_matches_oracle_and_invariants(gpu, oracle_lib):
    """WorldGrid() = 9 x 9 x 8 x 7 x 9 x 9 x 8 x 7 (20,575,296 nodes), RK2,
    linear speed boundary, 3 steps against the oracle bit for bit; then
    monotone frames, the avoid floor and the y-mirror symmetry of the game
    (up to ENO tie-breaking and the FP32 heading tables) over 0.5 s."""
    print("[world8d] default grid vs oracle", flush=True)
    case = dict(positions=9, headings=8, p_speeds=7, e_speeds=7, order=2,
                linear_speed_boundary=1, extent=3.5, arena=3.0, radius=0.3, p_min=-1.0,
                p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0, e_accel=1.0, p_curvature=1.0,
                e_curvature=2.0, cfl=0.8)
    from pe_amd import _native

    with _native.World8D(native_cfg(case)) as w:
        dt = w.stats()["max_internal_dt"]
        w.advance_to(3 * dt)
        final, arrival = w.read()
        assert w.stats()["internal_steps"] == 3
    _, o_final, o_arrival, o_steps, _ = oracle_lib.world8d_solve(
        oracle_cfg(oracle_lib, case), 3 * dt, 3 * dt, history=False)
    assert o_steps == 3
    np.testing.assert_array_equal(bits(final), bits(o_final))
    np.testing.assert_array_equal(bits(arrival), bits(o_arrival))
    del o_final, o_arrival

    print("[world8d] default grid to 0.5 s", flush=True)
    shape = (9, 9, 8, 7, 9, 9, 8, 7)
    with _native.World8D(native_cfg(case)) as w:
        previous = w.read(True, False)[0]
        for t in (0.125, 0.25, 0.375, 0.5):
            w.advance_to(t)
            field = w.read(True, False)[0]
            assert (field <= previous).all()
            previous = field
        final, arrival = w.read()
        assert w.count_nonfinite() == 0
    x = np.linspace(-3.5, 3.5, 9).astype(np.float32)
    avoid = np.maximum(abs(x)[:, None], abs(x)[None, :]) - np.float32(3.0) + np.float32(1e-6)
    f = final.reshape(shape)
    assert (f >= avoid[:, :, None, None, None, None, None, None] - 1e-6).all()
    captured = final <= 0
    assert captured.any() and np.isfinite(arrival[captured]).all()
    assert (arrival[captured] <= 0.5 + 1e-6).all() and np.isnan(arrival[~captured]).all()
    # y -> -y, theta -> -theta for both cars: indices y -> n-1-y,
    # theta_i -> theta_{(nh - i) mod nh}.
    heading = (-np.arange(8)) % 8
    # The non-strict ENO selection breaks ties toward the left neighbour, which
    # the mirror turns into the right one: asymmetry stays small on average.
    mirror = f[:, ::-1][:, :, heading][:, :, :, :, :, ::-1][:, :, :, :, :, :, heading]
    asym = np.abs(mirror - f)
    print(f"default grid mirror asymmetry max={asym.max():.3e} mean={asym.mean():.3e}")
    assert asym.mean() < 1e-4


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("speed_boundary", ["linear", "legacy_signed"])
def test_solve_tiny_update_matches_independent_numpy_and_probe_query(
        gpu, tmp_path, order, speed_boundary):
    """(tests/test_reversible_world_dp.py:113-167) through world_dp.solve; the
    saved probe values are the probe stencils applied to the saved frames."""
    from pe_amd.reversible.grid_stencil import interpolation_stencil
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import PERIODIC_AXES, WorldGrid, global_rate_bound, solve
    from world8d_numpy import one_step

    print(f"[world8d] solve order {order} {speed_boundary}", flush=True)
    config, grid, dt = ReversibleConfig(), WorldGrid(5, 4, 3, 3, 3.5), 0.01
    assert dt < 0.8 / global_rate_bound(config, grid)
    probe = np.array([[-1, 0, 0, 0.5, 0, 1, np.pi, 0.25], [0, 0, 0, 0.5, 0.1, 0.1, 0, 0.25]])
    out = tmp_path / "solve"
    meta = solve(config, grid, out, horizon=dt, output_dt=dt, order=order, threads=2,
                 history=True, probes=probe, speed_boundary=speed_boundary)
    initial, avoid, expected = one_step(grid, config, dt, order, speed_boundary)
    actual = np.memmap(out / "final.f32", dtype=np.float32, shape=grid.shape, mode="r")
    np.testing.assert_allclose(actual, expected, atol=2e-6, rtol=2e-6)
    assert meta["internal_steps"] == 1 and meta["status"] == "complete"
    assert np.max(actual - initial) < 2e-6
    assert np.min(actual - avoid) > -2e-6
    saved = np.fromfile(out / "probe_fields.f64", dtype=np.float64).reshape(2, 2)
    frames = np.fromfile(out / "fields.f32", dtype=np.float32).reshape(2, -1)
    ii, ww, valid, _ = interpolation_stencil(probe, grid.axes(config), PERIODIC_AXES)
    assert valid.all()
    for t in range(2):
        np.testing.assert_allclose((frames[t][ii].astype(np.float64) * ww).sum(1), saved[t],
                                   atol=1e-12)
    outside = probe[0].copy()
    outside[0] = 20
    assert not interpolation_stencil(outside, grid.axes(config), PERIODIC_AXES)[2].any()


@pytest.mark.parametrize("order", [1, 2])
def test_resume_reuses_parent_history_and_matches_monolithic_solve(gpu, tmp_path, order):
    """(tests/test_reversible_world_dp.py:229-287)"""
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import WorldGrid, solve

    print(f"[world8d] resume order {order}", flush=True)
    config, grid = ReversibleConfig(), WorldGrid(5, 4, 4, 4)
    parent, extension, monolithic = (tmp_path / k for k in ("parent", "extension", "mono"))
    solve(config, grid, parent, horizon=0.1, output_dt=0.05, order=order, history=True,
          threads=1)
    meta = solve(config, grid, extension, horizon=0.2, output_dt=0.05, order=order,
                 history=True, threads=1, resume_from=parent, write_final=False,
                 write_arrival=False)
    solve(config, grid, monolithic, horizon=0.2, output_dt=0.05, order=order, history=True,
          threads=1)
    assert meta["start_time"] == 0.1
    assert len(meta["times"]) == 2
    assert not (extension / "final.f32").exists()
    assert not (extension / "arrival.f32").exists()
    assert (extension / "fields.f32").stat().st_size == 2 * grid.states * 4
    full_meta = json.loads((monolithic / "meta.json").read_text())
    np.testing.assert_allclose(meta["times"], full_meta["times"][-2:], atol=1e-15)
    # the extension's frames are the monolithic solve's last two, bit for bit
    ext = np.fromfile(extension / "fields.f32", np.float32).reshape(2, -1)
    mono = np.fromfile(monolithic / "fields.f32", np.float32).reshape(-1, grid.states)
    np.testing.assert_array_equal(ext, mono[-2:])
    np.testing.assert_array_equal(ext[-1], np.fromfile(monolithic / "final.f32", np.float32))
    assert meta["cumulative_internal_steps"] == full_meta["internal_steps"]


def test_speed_limit_generator_on_gpu(gpu):
    """phi = 10 - xP - vP, RK1 one step: rate -2 with the inward speed slope,
    0 with the legacy extrapolation (tests/test_reversible_world_dp.py:204-226;
    the value -2 is the pursuer's best hold at top speed, heading 0)."""
    from pe_amd import _native
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import WorldGrid, native_config

    config, grid = ReversibleConfig(), WorldGrid(5, 4, 4, 4)
    mesh = np.broadcast_arrays(*np.meshgrid(*grid.axes(config), indexing="ij", sparse=True))
    phi = (10 - mesh[0] - mesh[3]).astype(np.float32)
    node = (2, 2, 2, 3, 2, 2, 2, 3)
    dt = 0.005
    for boundary, expected in (("linear", -2.0), ("legacy_signed", 0.0)):
        with _native.World8D(native_config(config, grid, order=1,
                                           speed_boundary=boundary)) as w:
            w.set_field(phi, None, 0.0)
            w.advance_to(dt)
            assert w.stats()["internal_steps"] == 1
            field = w.read(True, False)[0].reshape(grid.shape)
        assert (float(field[node]) - float(phi[node])) / dt == pytest.approx(expected, abs=2e-3)


def test_errors_are_reported(gpu, tmp_path):
    from pe_amd import _native

    case = dict(positions=5, headings=4, p_speeds=3, e_speeds=3, order=2,
                linear_speed_boundary=1, extent=3.5, arena=3.0, radius=0.3, p_min=-1.0,
                p_max=2.0, e_min=-0.5, e_max=1.0, p_accel=2.0, e_accel=1.0, p_curvature=1.0,
                e_curvature=2.0, cfl=0.8)
    bad = dict(case, extent=2.5)
    with pytest.raises(ValueError, match="Invalid arguments"):
        _native.World8D(native_cfg(bad))
    with _native.World8D(native_cfg(case)) as w:
        ii = np.zeros((1, 256), np.uint64)
        ww = np.full((1, 256), 1 / 256)
        ii[0, 3] = w.n
        with pytest.raises(ValueError, match="Bad probe stencil"):
            w.set_probes(ii, ww)
        ii[0, 3] = 0
        ww[0, 0] = 0.5
        with pytest.raises(ValueError, match="Bad probe weights"):
            w.set_probes(ii, ww)
        # A non-finite continuation field is reported by the final check.
        field = w.read(True, False)[0]
        field[7] = np.nan
        w.set_field(field, None, 0.0)
        w.advance_to(0.01)
        assert w.count_nonfinite() > 0


def test_grid_stencil_matches_oracle(gpu, oracle_lib):
    """pe_grid_stencil (probe stencils of this game and the 5D reader's
    stencil) against the oracle's uniform-grid stencil, bit for bit, with
    periodic wraps, states outside the grid and non-finite states."""
    from pe_amd.reversible.grid_stencil import axis_spec, interpolation_stencil
    from pe_amd.reversible.model import ReversibleConfig
    from pe_amd.reversible.world_dp import PERIODIC_AXES, WorldGrid

    rng = np.random.default_rng(11)
    axes = WorldGrid(6, 5, 4, 3).axes(ReversibleConfig())
    lo = np.array([a[0] for a in axes])
    hi = np.array([a[-1] for a in axes])
    s = rng.uniform(lo - 0.4, hi + 0.4, (5000, 8))
    s[::9, 2] += 40.0
    s[7, 5] = np.nan
    s[8, 6] = np.inf
    ii, ww, valid, batch = interpolation_stencil(s, axes, PERIODIC_AXES)
    n, lo_, hi_, step, per, mask = axis_spec(axes, PERIODIC_AXES)
    oi, ow, ov = oracle_lib.grid_stencil(s, n, lo_, hi_, step, per, mask, 1e-10)
    assert batch == (5000,) and 0 < valid.sum() < 5000
    np.testing.assert_array_equal(ii, oi)
    np.testing.assert_array_equal(ww.view(np.uint64), ow.view(np.uint64))
    np.testing.assert_array_equal(valid, ov)
