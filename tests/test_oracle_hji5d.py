"""Pins the 5D HJI oracle (oracle/hji5d.c) before it is trusted.

1. Golden vectors: the reference's own relative5d binary, compiled from
   reversible/native_relative.cpp (tests/golden/make_golden.py).  Bit-exact.
2. An independent formulation: one explicit Euler / RK2 step with the
   Hamiltonian computed by enumerating the 4 x 4 extreme control pairs over
   the relative dynamics (the check of
   tests/test_reversible_free_dp.py:16-129, restated here), not the closed
   absolute-value form the solver uses.
3. Invariants the mathematics fixes: fields never increase; the grid and the
   dynamics are mirror-symmetric under (y, psi) -> (-y, -psi); the CFL limit
   has a closed form; phi = hypot - r at t = 0.
Planted mistakes in the independent formulation (dropped dissipation, the
players' min/max roles swapped, a wrong sign on the heading sine) must fail
check 2, so the check has teeth.
"""

import json
from itertools import product
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["c0_a3_17x12x7x7", "unit_9x8x7x7_order1", "unit_9x8x7x7_order2"]


def golden(name):
    z = np.load(GOLDEN / f"hji5d_{name}.npz")
    return z, json.loads(str(z["case"]))


def cfg_from_case(oracle, case, threads=4, store_f64=False):
    ph = case["physics"]
    return oracle.HJI5DCfg(
        case["positions"], case["headings"], case["p_speeds"], case["e_speeds"],
        case["order"], threads, int(store_f64), case["extent"], case["cfl"],
        ph["radius"], ph["p_min"], ph["p_max"], ph["e_min"], ph["e_max"],
        ph["p_acceleration"], ph["e_acceleration"], ph["p_curvature"], ph["e_curvature"],
    )


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_binary_bit_for_bit(oracle_lib, name):
    z, case = golden(name)
    cfg = cfg_from_case(oracle_lib, case)
    r = oracle_lib.hji5d_solve(cfg, case["horizon"], case["output_dt"])
    assert r["internal_steps"] == int(z["internal_steps"])
    assert r["frames"] == int(z["frames"])
    assert r["cfl_dt_limit"] == pytest.approx(float(z["cfl_dt_limit"]), rel=1e-13)
    assert r["maximum_increase"] == float(z["maximum_increase"])
    np.testing.assert_array_equal(bits(r["final"]), bits(z["final"]))
    arr_ref, arr = z["arrival"], r["arrival"]
    np.testing.assert_array_equal(np.isnan(arr), np.isnan(arr_ref))
    np.testing.assert_array_equal(bits(np.nan_to_num(arr)), bits(np.nan_to_num(arr_ref)))
    if name.startswith("c0"):
        np.testing.assert_array_equal(bits(r["fields"][1:-1]), bits(z["fields"]))
    else:
        np.testing.assert_array_equal(bits(r["fields"]), bits(z["fields"]))


# ---------------------------------------------------------------- check 2

def _one_sided(f, axis, h, periodic, order):
    """ENO differences (minus, plus) along one axis with np.roll."""
    fm1, fp1 = np.roll(f, 1, axis), np.roll(f, -1, axis)
    minus, plus = (f - fm1) / h, (fp1 - f) / h
    raw_m, raw_p = minus.copy(), plus.copy()
    if order == 2:
        fm2, fp2 = np.roll(f, 2, axis), np.roll(f, -2, axis)
        mid = (fp1 - 2 * f + fm1) / h**2
        left = (f - 2 * fm1 + fm2) / h**2
        right = (fp2 - 2 * fp1 + f) / h**2
        if not periodic:
            n = f.shape[axis]
            idx = np.arange(n).reshape([-1 if k == axis else 1 for k in range(f.ndim)])
            left = np.where(idx < 2, mid, left)
            right = np.where(idx >= n - 2, mid, right)
        minus = minus + 0.5 * h * np.where(abs(left) < abs(mid), left, mid)
        plus = plus - 0.5 * h * np.where(abs(mid) < abs(right), mid, right)
    if not periodic:
        first = [slice(None)] * f.ndim
        last = [slice(None)] * f.ndim
        first[axis], last[axis] = 0, -1
        minus[tuple(first)] = plus[tuple(first)] = raw_p[tuple(first)]
        minus[tuple(last)] = plus[tuple(last)] = raw_m[tuple(last)]
    return minus, plus


def enumerated_step(field, axes, limits, dt, order, *, cast=True, plant=None):
    """Euler step f + dt * min(0, H) with H from explicit control enumeration.

    limits = (pmin, pmax, emin, emax, Ap, Ae, Kp, Ke).  Storage axes are
    [vP, vE, psi, y, x]; `plant` injects a deliberate mistake."""
    pmin, pmax, emin, emax, Ap, Ae, Kp, Ke = limits
    f = field.astype(np.float64)
    vp, ve, psi, y, x = np.meshgrid(*axes, indexing="ij")
    grads, jumps = [], []
    for k, ax in enumerate(axes):
        m, p = _one_sided(f, k, ax[1] - ax[0], k == 2, order)
        grads.append(0.5 * (m + p))
        jumps.append(p - m)
    g_vp, g_ve, g_psi, g_y, g_x = grads
    sin_psi = -np.sin(psi) if plant == "sin_sign" else np.sin(psi)
    outer, inner = (np.maximum, np.minimum) if plant == "swap_roles" else (np.minimum, np.maximum)

    def acc(v, a, lo, hi):
        return np.where(((v <= lo) & (a < 0)) | ((v >= hi) & (a > 0)), 0.0, a)

    best = None
    for a, k in product((-Ap, Ap), (-Kp, Kp)):
        worst = None
        for b, l in product((-Ae, Ae), (-Ke, Ke)):
            wp, we = vp * k, ve * l
            dot = (
                g_x * (ve * np.cos(psi) - vp + wp * y)
                + g_y * (ve * sin_psi - wp * x)
                + g_psi * (we - wp)
                + g_vp * acc(vp, a, pmin, pmax)
                + g_ve * acc(ve, b, emin, emax)
            )
            worst = dot if worst is None else inner(worst, dot)
        best = worst if best is None else outer(best, worst)
    wp, we = abs(vp) * Kp, abs(ve) * Ke
    alpha = [Ap, Ae, wp + we, abs(ve * np.sin(psi)) + wp * abs(x),
             abs(ve * np.cos(psi) - vp) + wp * abs(y)]
    if plant != "no_dissipation":
        best = best + sum(0.5 * a_ * j for a_, j in zip(alpha, jumps))
    out = f + dt * np.minimum(best, 0)
    return out.astype(np.float32) if cast else out


def march(z, case, plant=None):
    ph = case["physics"]
    limits = (ph["p_min"], ph["p_max"], ph["e_min"], ph["e_max"], ph["p_acceleration"],
              ph["e_acceleration"], ph["p_curvature"], ph["e_curvature"])
    shape = tuple(int(s) for s in z["shape"])
    axes = (
        np.linspace(ph["p_min"], ph["p_max"], shape[0]),
        np.linspace(ph["e_min"], ph["e_max"], shape[1]),
        np.linspace(-np.pi, np.pi, shape[2], endpoint=False),
        np.linspace(-case["extent"], case["extent"], shape[3]),
        np.linspace(-case["extent"], case["extent"], shape[4]),
    )
    fields = z["fields"].reshape((-1,) + shape)
    f = fields[0]
    limit = float(z["cfl_dt_limit"])
    t, out = 0.0, []
    goals = np.minimum(np.arange(1, len(fields)) * case["output_dt"], case["horizon"])
    for goal in goals:
        while t < goal - 1e-12:
            dt = min(limit, goal - t)
            if case["order"] == 1:
                f = enumerated_step(f, axes, limits, dt, 1, plant=plant)
            else:
                s = enumerated_step(f, axes, limits, dt, 2, plant=plant)
                f = (0.5 * (f.astype(np.float64)
                            + enumerated_step(s, axes, limits, dt, 2, cast=False, plant=plant))
                     ).astype(np.float32)
            t += dt
        out.append(f)
    return fields, np.stack(out)


@pytest.mark.parametrize("name", CASES[1:])
def test_oracle_matches_explicit_control_enumeration(oracle_lib, name):
    z, case = golden(name)
    cfg = cfg_from_case(oracle_lib, case)
    r = oracle_lib.hji5d_solve(cfg, case["horizon"], case["output_dt"])
    shape = tuple(int(s) for s in z["shape"])
    frames = r["fields"].reshape((-1,) + shape)
    _, marched = march(z, case)
    np.testing.assert_allclose(frames[1:], marched, atol=2e-6, rtol=2e-6)


@pytest.mark.parametrize("plant", ["no_dissipation", "swap_roles", "sin_sign"])
def test_enumeration_check_catches_planted_mistakes(oracle_lib, plant):
    z, case = golden("unit_9x8x7x7_order2")
    cfg = cfg_from_case(oracle_lib, case)
    r = oracle_lib.hji5d_solve(cfg, case["horizon"], case["output_dt"])
    shape = tuple(int(s) for s in z["shape"])
    frames = r["fields"].reshape((-1,) + shape)
    _, marched = march(z, case, plant=plant)
    assert np.abs(frames[1:] - marched).max() > 1e-4


# ---------------------------------------------------------------- check 3

def test_fields_monotone_and_phi_at_time_zero(oracle_lib):
    z, case = golden("c0_a3_17x12x7x7")
    cfg = cfg_from_case(oracle_lib, case)
    r = oracle_lib.hji5d_solve(cfg, case["horizon"], case["output_dt"])
    assert np.diff(r["fields"], axis=0).max() <= 0
    n = case["positions"]
    xs = np.linspace(-case["extent"], case["extent"], n)
    phi = (np.hypot(xs[None, :], xs[:, None]) - case["physics"]["radius"]).astype(np.float32)
    shape = tuple(int(s) for s in z["shape"])
    f0 = r["fields"][0].reshape(shape)
    np.testing.assert_allclose(f0, np.broadcast_to(phi, shape), atol=1e-6)


def test_mirror_symmetry(oracle_lib):
    """V(x, -y, -psi) == V(x, y, psi): the heading grid -pi + k*2pi/n maps to
    itself under psi -> -psi (k -> n - k), y -> -y reverses the y axis."""
    z, case = golden("c0_a3_17x12x7x7")
    cfg = cfg_from_case(oracle_lib, case, store_f64=True)
    r = oracle_lib.hji5d_solve(cfg, case["horizon"], case["output_dt"], history=False)
    shape = tuple(int(s) for s in z["shape"])
    f = r["final"].reshape(shape)
    nh = shape[2]
    mirrored = f[:, :, (-np.arange(nh)) % nh][:, :, :, ::-1, :]
    # The ENO2 selection uses strict "<" (native_relative.cpp:65-66), so an
    # exact magnitude tie |sl| == |sc| picks a different branch in the mirror
    # image; phi = hypot - r is full of such ties.  The asymmetry this leaves
    # is O(h * second difference): about 2e-4 at most, 2e-7 on average.
    d = np.abs(f - mirrored)
    assert d.max() < 1e-3 and d.mean() < 1e-6


def test_cfl_limit_closed_form(oracle_lib):
    z, case = golden("c0_a3_17x12x7x7")
    cfg = cfg_from_case(oracle_lib, case)
    ph = case["physics"]
    dx = 2 * case["extent"] / (case["positions"] - 1)
    dphi = 2 * np.pi / case["headings"]
    dvp = (ph["p_max"] - ph["p_min"]) / (case["p_speeds"] - 1)
    dve = (ph["e_max"] - ph["e_min"]) / (case["e_speeds"] - 1)
    VP, VE, W = 2.0, 1.0, case["extent"]
    kp, ke = ph["p_curvature"], ph["e_curvature"]
    rate = ((VP + VE + VP * kp * W) + (VE + VP * kp * W)) / dx + (VP * kp + VE * ke) / dphi \
        + ph["p_acceleration"] / dvp + ph["e_acceleration"] / dve
    assert oracle_lib.hji5d_cfl_limit(cfg) == pytest.approx(0.8 / rate, rel=1e-14)
    assert oracle_lib.hji5d_cfl_limit(cfg) == pytest.approx(0.012187794918605, rel=1e-12)


def test_float64_storage_close_to_float32_storage(oracle_lib):
    """The FP64-storage oracle used to state GPU FP32 errors agrees with the
    float-storage reference up to the scheme's own rounding sensitivity:
    rounding decides ENO selections at exact ties, which moves a few nodes by
    O(h * second difference) (about 4e-4 at most on C0, 3e-7 on average)."""
    z, case = golden("c0_a3_17x12x7x7")
    r64 = oracle_lib.hji5d_solve(cfg_from_case(oracle_lib, case, store_f64=True),
                                 case["horizon"], case["output_dt"], history=False)
    d = np.abs(r64["final"] - z["final"].astype(np.float64))
    assert d.max() < 1e-3 and d.mean() < 1e-6


def _exit_cfg(oracle, exit_radius, nx=33, extent=4.0, threads=4):
    return oracle.HJI5DCfg(nx, 12, 7, 7, 2, threads, 0, extent, 0.8, 0.3,
                           -1.0, 2.0, -0.5, 1.0, 3.0, 3.0, 1.0, 2.0, exit_radius)


def test_exit_radius_mode_known_answers_and_locality(oracle_lib):
    """The oracle's local-exit mode (gpu_relative.py:126-131, float64):
    nodes beyond the decimal literal of r^2 hold the literal of r at every
    step (count from numpy, float64); the test compares with the literal,
    not r * r (2.3: r * r = 5.289999999999999, literal 5.29); nodes whose
    +-2 stencil touches no exit node evolve exactly as without the exit
    clamp."""
    r = 2.3
    lit_r2, lit_r = float(f"{r * r:.9e}"), float(f"{r:.9e}")
    assert lit_r2 != r * r
    cfg = _exit_cfg(oracle_lib, r)
    base = _exit_cfg(oracle_lib, 0.0)
    out = oracle_lib.hji5d_solve(cfg, 0.1, 0.05)
    ref = oracle_lib.hji5d_solve(base, 0.1, 0.05)
    n = 33
    x = -4.0 + np.arange(n) * (8.0 / (n - 1))
    xx, yy = np.meshgrid(x, x)
    outside = xx * xx + yy * yy > lit_r2
    init_outside = xx * xx + yy * yy > r * r
    f = out["final"].reshape(7, 7, 12, n, n)
    assert (f[..., outside] == np.float32(lit_r)).all()
    frame0 = out["fields"][0].reshape(7, 7, 12, n, n)
    assert (frame0[..., init_outside] == np.float32(r)).all()
    # one RK2 step (two stages reach 4 nodes in x and y) with and without
    # the clamp: equal wherever no exit node lies within 4 nodes
    f1, a1 = oracle_lib.hji5d_init(cfg)
    f2, a2 = oracle_lib.hji5d_init(base)
    dt = oracle_lib.hji5d_cfl_limit(cfg)
    oracle_lib.hji5d_time_steps(cfg, f1, a1, dt, 1)
    oracle_lib.hji5d_time_steps(base, f2, a2, dt, 1)
    from scipy.ndimage import binary_dilation

    far = ~binary_dilation(outside, np.ones((9, 9), bool))
    assert far.sum() > 30
    f1, f2 = f1.reshape(7, 7, 12, n, n), f2.reshape(7, 7, 12, n, n)
    np.testing.assert_array_equal(f1[..., far].view(np.uint32), f2[..., far].view(np.uint32))
    assert (f1[..., outside] == np.float32(lit_r)).all()
    assert (f <= frame0 + 1e-6).all()
