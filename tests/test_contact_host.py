"""Host pieces of the contact solve that need no GPU: the sign-preserving
float16 history (reference.py:19-25) and the solve's rule checks."""

import numpy as np
import pytest

from pe_amd.local_game.geometry import Footprint, random_rectangle_arena
from pe_amd.local_game.reference import contact_config, solve, write_half_preserving_sign


def test_half_history_keeps_every_sign(tmp_path):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=1000) * 10.0 ** rng.integers(-12, 4, 1000),
                        [0.0, -0.0, 1e-30, -1e-30, 6e-8, -6e-8, 3e-8, 60000.0, -60000.0]])
    x = x.astype(np.float32)
    with open(tmp_path / "h.f16", "wb") as f:
        write_half_preserving_sign(x, f)
    h = np.fromfile(tmp_path / "h.f16", np.float16)
    assert len(h) == len(x)
    nonzero = x != 0
    assert (np.sign(h[nonzero].astype(np.float32)) == np.sign(x[nonzero])).all()
    assert (h[~nonzero] == 0).all()
    tiny = np.float16(2.0 ** -24)
    flushed = nonzero & (x.astype(np.float16) == 0)
    assert flushed.sum() > 50
    np.testing.assert_array_equal(np.abs(h[flushed]), tiny)
    same = ~flushed
    np.testing.assert_array_equal(h[same].view(np.uint16),
                                  x[same].astype(np.float16).view(np.uint16))


def test_half_history_streams_in_chunks(tmp_path):
    x = np.linspace(-1e-9, 1e-9, 5_000_001, dtype=np.float32)
    with open(tmp_path / "h.f16", "wb") as f:
        write_half_preserving_sign(x, f)
    h = np.fromfile(tmp_path / "h.f16", np.float16)
    assert len(h) == len(x) and (np.sign(h) == np.sign(x)).all()


def test_solve_rule_checks_precede_any_work(tmp_path):
    body = Footprint(0.5, 0.3)
    arena = random_rectangle_arena(1, pursuer=body, evader=body, count=4)
    from pe_amd.reversible.model import ReversibleConfig

    with pytest.raises(ValueError, match="finite-body"):
        solve(ReversibleConfig(), arena, tmp_path / "a")
    with pytest.raises(ValueError, match="integer control periods"):
        solve(contact_config(), arena, tmp_path / "b", horizon=0.33)
    assert not (tmp_path / "a").exists() and not (tmp_path / "b").exists()
