"""Table folders for the 5D reader tests (the reference's synthetic_table,
tests/test_reversible_free_dp.py:132-156: a field linear in x, y, vP, vE and
time, with a bogus arrival file that queries must ignore)."""

import json

import numpy as np

from pe_amd.reversible.free_dp import RelativeGrid, physics_signature
from pe_amd.reversible.model import ReversibleConfig


def synthetic_table(tmp_path):
    from pe_amd.reversible.relative_table import RelativeTable

    c = ReversibleConfig()
    g = RelativeGrid(9, 8, 7, 7, 2)
    axes = g.axes(c)
    vp, ve, psi, y, x = np.meshgrid(*axes, indexing="ij")
    times = np.array([0, 0.5, 1, 1.5, 2])
    field = np.array([x + 0.2 * y + 0.3 * vp - 0.4 * ve - t for t in times], np.float32)
    folder = tmp_path / "synthetic"
    folder.mkdir()
    field.tofile(folder / "fields.f32")
    np.full(g.shape, 123, dtype=np.float32).tofile(folder / "arrival.f32")
    meta = dict(schema="reversible_relative5d_v1", status="complete", history_saved=True,
                tube_projection="min over every internal update", times=times.tolist(),
                axes=[a.tolist() for a in axes], physics=physics_signature(c),
                config=c.to_dict(), grid=g.to_dict())
    (folder / "meta.json").write_text(json.dumps(meta))
    return RelativeTable(folder, c), c
