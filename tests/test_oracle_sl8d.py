"""Pins the 8D contact-game oracle (oracle/sl8d.c, oracle/geometry.c).

Backup: the factorised FP64 sweep against a brute force over joint (P, E)
commands with the full 16 x 16 product weights (the check of
tests/test_compact_dp.py:31-50, restated here) on seeded random tables;
planted mistakes (players' roles swapped, no tube projection) must fail it.
Geometry: bit-for-bit against the reference's own OpenMP library
(oracle/_ref/libnative_geometry.so, compiled from native_geometry.cpp) on
seeded states of a random-rectangle map, plus the reference's known answers
(tests/test_contact_game.py:21-121).
"""

import ctypes
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from pe_amd.local_game.geometry import (
    CircleArena,
    Footprint,
    random_rectangle_arena,
    rectangle_separation,
)
from pe_amd.reversible.model import VehicleLimits

ROOT = Path(__file__).resolve().parents[1]
REF_GEOM = ROOT / "oracle" / "_ref" / "libnative_geometry.so"


def fake_grid(n, na, rng):
    """Random stencil tables: ids uniform, weights normalised, command 0
    always allowed (tests/test_compact_dp.py:12-28)."""
    states = rng.uniform(-1, 1, (n, 4))
    ids = rng.integers(n, size=(n, na, 16), dtype=np.int32)
    w = rng.random((n, na, 16)).astype(np.float32)
    w /= w.sum(-1, keepdims=True)
    allowed = rng.random((n, na)) > 0.3
    allowed[:, 0] = True
    return SimpleNamespace(states=states, successors=ids, weights=w, allowed=allowed)


def inside_mask(p, e, radius):
    d = np.linalg.norm(p.states[:, None, :2] - e.states[None, :, :2], axis=-1)
    return d <= radius


def brute(f, p, e, inside, plant=None):
    """Joint enumeration: min over P commands of max over E commands of the
    16 x 16 product interpolation."""
    out = f.copy()
    for i in range(len(p.states)):
        for j in range(len(e.states)):
            if not inside[i, j]:
                continue
            values = []
            for a in np.flatnonzero(p.allowed[i]):
                resp = [
                    np.sum(f[np.ix_(p.successors[i, a], e.successors[j, b])]
                           * p.weights[i, a, :, None].astype(np.float64)
                           * e.weights[j, b, None, :].astype(np.float64))
                    for b in np.flatnonzero(e.allowed[j])
                ]
                values.append(min(resp) if plant == "swap_roles" else max(resp))
            best = max(values) if plant == "swap_roles" else min(values)
            out[i, j] = best if plant == "no_tube" else min(out[i, j], best)
    return out


@pytest.mark.parametrize("seed", [114, 57, 3])
def test_sweep_equals_joint_enumeration(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    p, e = fake_grid(5, 3, rng), fake_grid(7, 4, rng)
    inside = inside_mask(p, e, 1.2)
    V = rng.uniform(-0.5, 1.5, (5, 7))
    expected = V.copy()
    for _ in range(3):
        expected = brute(expected, p, e, inside)
        V, change = oracle_lib.sl_sweep(V, p, e, inside)
        np.testing.assert_allclose(V, expected, atol=1e-12)
    assert change >= 0


@pytest.mark.parametrize("plant", ["swap_roles", "no_tube"])
def test_enumeration_catches_planted_mistakes(oracle_lib, plant):
    rng = np.random.default_rng(114)
    p, e = fake_grid(5, 3, rng), fake_grid(7, 4, rng)
    inside = inside_mask(p, e, 1.2)
    V = rng.uniform(-0.5, 1.5, (5, 7))
    out, _ = oracle_lib.sl_sweep(V, p, e, inside)
    assert np.abs(out - brute(V, p, e, inside, plant=plant)).max() > 1e-3


def test_sweep_invariants(oracle_lib):
    rng = np.random.default_rng(9)
    p, e = fake_grid(40, 9, rng), fake_grid(50, 9, rng)
    inside = inside_mask(p, e, 1.0)
    V = rng.uniform(-0.3, 1.0, (40, 50))
    V[~inside] = 1.0
    prev = V
    for _ in range(4):
        nxt, change = oracle_lib.sl_sweep(prev, p, e, inside)
        assert (nxt <= prev).all()                     # tube projection
        np.testing.assert_array_equal(nxt[~inside], 1.0)  # exit values fixed
        assert change == pytest.approx(np.abs(nxt - prev).max())
        assert (nxt[prev <= 0] <= 0).all()             # captured stays captured
        prev = nxt


def test_initial_field_is_contact_payoff_inside(oracle_lib):
    rng = np.random.default_rng(5)
    p, e = fake_grid(6, 2, rng), fake_grid(8, 2, rng)
    inside = inside_mask(p, e, 1.2)
    pb, eb = Footprint(0.5, 0.3), Footprint(0.5, 0.3)
    V = oracle_lib.sl_init(p.states, e.states, inside, 1.2, pb, eb)
    expect = rectangle_separation(p.states[:, None], e.states[None], pb, eb)
    np.testing.assert_allclose(V, np.where(inside, expect, 1.2), atol=1e-14)


# ----------------------------------------------------------------- geometry

def _ref_lib():
    if not REF_GEOM.exists():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = ctypes.CDLL(str(REF_GEOM))
    P, Lg, I, D = ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_double
    L.batch_clearance.argtypes = [P, Lg, P, I, D, P, P, I]
    L.batch_backups.argtypes = [P, Lg, P, I, D, P, P, I]
    L.batch_actions.argtypes = [P, Lg, P, I, D, P, I, D, P, P, P, I]
    L.configuration_connections.argtypes = [P, P, Lg, P, I, D, P, P, I]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def seeded_states(n, seed):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(np.stack(
        [rng.uniform(-3, 3, n), rng.uniform(-3, 3, n), rng.uniform(-np.pi, np.pi, n),
         rng.uniform(-1, 2, n)], 1))


CONTROLS = np.array([[a, k] for a in (-3, 0, 3) for k in (-1, 0, 1)], float)


def test_geometry_matches_reference_library_bit_for_bit(oracle_lib):
    L = _ref_lib()
    arena = random_rectangle_arena(3)
    obs = np.ascontiguousarray(arena.circles)
    par = oracle_lib.model_params(VehicleLimits(-1, 2, 3, 1), Footprint(0.5, 0.3), 0.05)
    s = seeded_states(3000, 0)
    ref = np.empty(len(s))
    L.batch_clearance(_p(s), len(s), _p(obs), len(obs), 3.0, _p(par), _p(ref), 4)
    np.testing.assert_array_equal(oracle_lib.geom_clearance(s, obs, 3.0, par), ref)
    refb = np.empty(len(s), np.uint8)
    L.batch_backups(_p(s), len(s), _p(obs), len(obs), 3.0, _p(par), _p(refb), 4)
    np.testing.assert_array_equal(oracle_lib.geom_backups(s, obs, 3.0, par), refb.astype(bool))
    sub = np.ascontiguousarray(s[:800])
    ok_ref = np.empty((800, 9), np.uint8)
    nxt_ref = np.empty((800, 9, 4))
    L.batch_actions(_p(sub), 800, _p(CONTROLS), 9, 0.05, _p(obs), len(obs), 3.0, _p(par),
                    _p(ok_ref), _p(nxt_ref), 4)
    ok, nxt = oracle_lib.geom_actions(sub, CONTROLS, 0.05, obs, 3.0, par)
    np.testing.assert_array_equal(ok, ok_ref.astype(bool))
    np.testing.assert_array_equal(nxt, nxt_ref)
    t = np.ascontiguousarray(s[::-1] + np.random.default_rng(1).normal(0, 0.2, s.shape))
    con_ref = np.empty(len(s), np.uint8)
    L.configuration_connections(_p(s), _p(t), len(s), _p(obs), len(obs), 3.0, _p(par),
                                _p(con_ref), 4)
    np.testing.assert_array_equal(oracle_lib.geom_connections(s, t, obs, 3.0, par),
                                  con_ref.astype(bool))


def test_geometry_known_answers(oracle_lib):
    """tests/test_contact_game.py:64-70,109-121 with the A3 pursuer/evader."""
    P = VehicleLimits(-1.0, 2.0, 3.0, 1.0)
    E = VehicleLimits(-0.5, 1.0, 3.0, 2.0)
    par = oracle_lib.model_params(P, Footprint(0.8, 0.5), 0.05)
    empty = np.zeros((0, 3))
    got = oracle_lib.geom_backups([[2.3, 0, 0, 2], [2.3, 0, 0, 0], [1.8, 0, 0, 2]], empty, 3.0, par)
    np.testing.assert_array_equal(got, [False, True, True])
    ok, _ = oracle_lib.geom_actions([[0, 0, 0, 0.02]], [[-3, 0], [0, 0], [3, 0], [-0.4, 0]],
                                    0.05, empty, 0.40055, par)
    np.testing.assert_array_equal(ok[0], [False, False, False, True])
    epar = oracle_lib.model_params(E, Footprint(0.4, 0.25), 0.05)
    assert not oracle_lib.geom_backups([[2.7892, 0, 0, 0.25]], empty, 3.0, epar)[0]


def test_held_arc_closed_form(oracle_lib):
    """Exact arc of reversible/model.py:86-153 (saturated speed, sinc chord)."""
    lim = VehicleLimits(-1.0, 2.0, 3.0, 1.0)
    par = oracle_lib.model_params(lim, Footprint(0.5, 0.3), 0.05)
    s = seeded_states(500, 7)
    _, nxt = oracle_lib.geom_actions(s, CONTROLS, 0.05, np.zeros((0, 3)), 50.0, par)
    for u, (a, k) in enumerate(CONTROLS):
        v0 = s[:, 3]
        target = np.where(a >= 0, lim.speed_max, lim.speed_min)
        hit = np.where(abs(a) > 1e-14, (target - v0) / (a if a else 1.0), np.inf)
        act = np.minimum(0.05, np.maximum(0.0, hit))
        v1 = np.clip(v0 + a * 0.05, lim.speed_min, lim.speed_max)
        d = v0 * act + 0.5 * a * act**2 + v1 * (0.05 - act)
        turn = k * d
        chord = d * np.sinc(turn / (2 * np.pi))
        mid = s[:, 2] + turn / 2
        expect = np.stack([s[:, 0] + chord * np.cos(mid), s[:, 1] + chord * np.sin(mid),
                           (s[:, 2] + turn + np.pi) % (2 * np.pi) - np.pi, v1], -1)
        err = nxt[:, u] - expect
        err[:, 2] = (err[:, 2] + np.pi) % (2 * np.pi) - np.pi
        assert abs(err).max() < 1e-12


def test_rectangle_contact_zeros():
    """Contact payoff zeros of tests/test_contact_game.py:21-42."""
    p = np.array([0.0, 0, 0, 0])
    e = np.array([0.6, 0, 0, 0])
    assert rectangle_separation(p, e) == pytest.approx(0, abs=1e-15)
    e[0], e[1] = 0, 0.375
    assert rectangle_separation(p, e) == pytest.approx(0, abs=1e-15)
    e[1] = 0.5
    assert rectangle_separation(p, e) > 0
    e[2] = np.pi / 2
    e[1] = 0.45
    assert rectangle_separation(p, e) == pytest.approx(0, abs=1e-15)
    arena = CircleArena([[0.4, 0.7, 0.2]], 3)
    assert arena.capture_value(np.r_[p, 0.6, 0, 0, 0]) == pytest.approx(0, abs=1e-15)
