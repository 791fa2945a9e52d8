"""End-to-end contact solve through the reference API on the GPU, and the
reader of its saved history.

* `solve(config, arena, out, ...)` writes the reference's folder
  (meta.json, p_grid.npz, e_grid.npz, domain.npz, queries.npz, labels.npz,
  fields.f16; the checkpoint removed on completion) and a second call on a
  complete folder returns its meta without work (reference.py:176-184).
* ContactReference (reference.py:372-454) answers from the device-resident
  float16 history.  Pinned by what the stencil fixes: at a grid-node pair
  state the stencil is a single vertex of weight one, so every budget's value
  is the stored float16 value itself (clipped by the payoff); frame 0 is the
  payoff.  At the sampled continuous queries the reader agrees with the
  solver's own FP32 labels (written by pe_sl_query from the unrounded field)
  to float16 rounding: |dV| <= 2e-3 for |V| <= 4.
* time_value is the first budget where the field reaches zero.
"""

import json

import numpy as np
import pytest

from pe_amd.local_game.geometry import Footprint, random_rectangle_arena
from pe_amd.local_game.reference import ContactReference, contact_config, solve
from pe_amd.reversible.map_game_field import first_crossing

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solved(tmp_path_factory):
    out = tmp_path_factory.mktemp("contact") / "run"
    config = contact_config()
    body = Footprint(0.5, 0.3)
    arena = random_rectangle_arena(5, pursuer=body, evader=body, count=8)
    print("[gpu] contact solve (7 positions, 8 headings, 3 speeds)", flush=True)
    meta = solve(config, arena, out, headings=8, positions=7, half_width=3.0, p_speeds=3,
                 e_speeds=3, horizon=0.5, output_dt=0.25, retain_history=True, frames=4,
                 per_frame=16)
    return out, meta


def test_solve_writes_reference_folder_and_resumes(gpu, solved):
    out, meta = solved
    assert meta["status"] == "complete"
    for name in ("meta.json", "p_grid.npz", "e_grid.npz", "domain.npz", "queries.npz",
                 "labels.npz", "fields.f16"):
        assert (out / name).exists(), name
    assert not list(out.glob("checkpoint_*"))
    assert meta["times"] == pytest.approx([0.0, 0.25, 0.5])
    stored = json.loads((out / "meta.json").read_text())
    assert stored["schema"] == "local_contact_compact8d_v1"
    size = (out / "fields.f16").stat().st_size
    assert size == 3 * stored["pair_states"] * 2
    again = solve(contact_config(), random_rectangle_arena(5, count=8), out, headings=8,
                  positions=7, half_width=3.0, p_speeds=3, e_speeds=3, horizon=0.5,
                  output_dt=0.25, retain_history=True, frames=4, per_frame=16)
    assert again["status"] == "complete" and again["signature"] == stored["signature"]


def test_reader_at_grid_nodes_returns_stored_values(gpu, solved):
    out, _ = solved
    ref = ContactReference(out)
    d = ref.domain
    rng = np.random.default_rng(0)
    ip = rng.integers(len(ref.p.states), size=200)
    rows = []
    for i in ip:
        pos = d.p_position[i]
        ie = d.e_indices[pos, rng.integers(d.counts[pos])]
        rows.append((i, ie))
    ip, ie = np.array(rows).T
    s = np.concatenate([ref.p.states[ip], ref.e.states[ie]], axis=1)
    values, valid, _ = ref.fields(s)
    assert valid.all()
    packed = d.offsets[ip] + d.lookup[d.p_position[ip], ie]
    stored = np.asarray(ref.values[:, packed], np.float32)
    phi = ref.arena.capture_value(s)
    np.testing.assert_allclose(values[1:], np.minimum(stored[1:], phi[None]), atol=1e-6)
    np.testing.assert_allclose(values[0], phi, atol=1e-6)
    np.testing.assert_allclose(ref.field(s, 0.25), values[1], atol=1e-6)
    np.testing.assert_allclose(ref.field(s, 0.375), 0.5 * (values[1] + values[2]), atol=1e-6)
    np.testing.assert_array_equal(np.isnan(ref.time_value(s)),
                                  np.isnan(first_crossing(ref.times, values, valid)))
    ref.close()


def test_reader_matches_solver_labels(gpu, solved):
    out, _ = solved
    ref = ContactReference(out)
    q = np.load(out / "queries.npz")["states"]
    labels = np.load(out / "labels.npz")
    values, valid, _ = ref.fields(q)
    phi = ref.arena.capture_value(q)
    keep = valid
    want = np.minimum(labels["values"][1:], phi[None])
    d = np.abs(values[1:, keep] - want[:, keep])
    print(f"reader vs FP32 labels at {keep.sum()} queries: max|dV|={d.max():.2e}")
    assert d.max() <= 2e-3
    ref.close()


def test_reader_continuous_queries_match_product_formula(gpu, oracle_lib, solved):
    """ContactReference.fields (reference.py:405-431) at the solve's
    continuous query states, against a restatement in numpy: both cars'
    stencils from the oracle, the 16 x 16 product weights, the stored
    float16 value of each kept pair and the exit value for the others,
    summed in float64; then the payoff clip, the exit value outside the
    radius and the payoff at budget 0.  The reader sums in float32 on the
    GPU: agreement to 1e-6 at every budget."""
    out, _ = solved
    ref = ContactReference(out)
    q = np.load(out / "queries.npz")["states"].reshape(-1, 8)
    got, ok, _ = ref.fields(q)

    def stencil(g, s):
        geo = g.geometry
        ids, w, valid, _, _, _ = oracle_lib.car_stencil(g.axes, g.dense_to_compact, s,
                                                        geo.obstacles, geo.arena.half_width,
                                                        geo.params)
        return np.maximum(ids, 0), w.astype(np.float64), valid

    pi, pw, pv = stencil(ref.p, q[:, :4])
    ei, ew, ev = stencil(ref.e, q[:, 4:])
    d = ref.domain
    rank = d.lookup[d.p_position[pi][:, :, None], ei[:, None, :]]
    packed = d.offsets[pi][:, :, None] + np.maximum(rank, 0)
    history = np.asarray(ref.values, np.float64)
    terms = np.where(rank[None] >= 0, history[:, packed], d.exit_radius)
    want = np.einsum("na,nb,tnab->tn", pw, ew, terms)
    phi = ref.arena.capture_value(q)
    want = np.minimum(want, phi[None])
    inside = np.linalg.norm(q[:, :2] - q[:, 4:6], axis=1) <= d.exit_radius
    want = np.where(inside[None], want, d.exit_radius)
    valid = pv & ev
    want[0, valid & inside] = phi[valid & inside]
    np.testing.assert_array_equal(ok, valid & inside)
    assert np.isnan(got[:, ~valid]).all()
    diff = np.abs(got[:, valid] - want[:, valid])
    print(f"reader vs product formula at {valid.sum()} continuous queries, "
          f"{len(ref.times)} budgets: max|dV|={diff.max():.2e}")
    assert valid.sum() > 0 and diff.max() <= 1e-6
    ref.close()


def test_resume_after_interruption_equals_uninterrupted_solve(gpu, tmp_path, monkeypatch):
    """(reference.py:289-348) A solve interrupted right after the checkpoint
    of budget 0.5 s (host-side exception in the record step), then re-run
    into the same folder, resumes from that checkpoint and ends with the
    same labels, history and field as a solve that was never interrupted."""
    import pe_amd.local_game.reference as reference

    config = contact_config()
    body = Footprint(0.5, 0.3)
    arena = random_rectangle_arena(6, pursuer=body, evader=body, count=8)
    kw = dict(headings=8, positions=7, half_width=3.0, p_speeds=3, e_speeds=3, horizon=1.0,
              output_dt=0.25, retain_history=True, frames=4, per_frame=16,
              residual_tol=1e-3, max_sweeps=400)
    print("[gpu] contact solve, uninterrupted", flush=True)
    whole = solve(config, arena, tmp_path / "whole", **kw)
    original = reference._ContactSolve.record

    def record_then_fail(self, step):
        original(self, step)
        if step == 10:
            raise KeyboardInterrupt("interrupted after the 0.5 s checkpoint")

    monkeypatch.setattr(reference._ContactSolve, "record", record_then_fail)
    print("[gpu] contact solve, interrupted at 0.5 s", flush=True)
    with pytest.raises(KeyboardInterrupt):
        solve(config, arena, tmp_path / "cut", **kw)
    cut_meta = json.loads((tmp_path / "cut" / "meta.json").read_text())
    assert cut_meta["status"] == "running" and cut_meta["completed_step"] == 10
    assert (tmp_path / "cut" / "checkpoint_0010.f32").exists()
    monkeypatch.setattr(reference._ContactSolve, "record", original)
    print("[gpu] contact solve, resumed", flush=True)
    resumed = solve(config, arena, tmp_path / "cut", **kw)
    assert resumed["status"] == "complete" and resumed["times"] == whole["times"]
    for name in ("labels.npz",):
        a, b = np.load(tmp_path / "whole" / name), np.load(tmp_path / "cut" / name)
        np.testing.assert_array_equal(a["values"], b["values"])
        np.testing.assert_array_equal(a["times"], b["times"])
    for name in ("fields.f16", "converged.f32"):
        np.testing.assert_array_equal(np.fromfile(tmp_path / "whole" / name, np.uint8),
                                      np.fromfile(tmp_path / "cut" / name, np.uint8))
    assert not list((tmp_path / "cut").glob("checkpoint_*"))
