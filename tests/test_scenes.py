"""The synthetic scene generators (host logic, no GPU)."""
from __future__ import annotations

import numpy as np

from fluid_simulation_amd import scenes


def test_dam_break_batch_shapes_and_placement():
    scs = scenes.dam_break_batch(64, 24, 130, seed=7)
    assert len(scs) == 64
    for sc in scs:
        assert 24 <= sc.nx <= 130 and 24 <= sc.ny <= 130 and sc.nx != sc.ny
        assert sc.length_x == sc.nx and sc.length_y == sc.ny
        p = sc.particles
        assert p.dtype == np.float32 and p.shape[1] == 4 and len(p) % 4 == 0
        # inside the interior, never in the border cells
        assert p[:, 0].min() >= 1.0 and p[:, 0].max() < sc.nx - 1
        assert p[:, 1].min() >= 1.0 and p[:, 1].max() < sc.ny - 1
        assert not p[:, 2:].any()
        if sc.solid_mask is not None:
            m = sc.solid_mask.astype(bool)
            assert m.shape == (sc.ny, sc.nx)
            assert 0 < m.sum() <= 0.15 * sc.nx * sc.ny
            assert not m[0].any() and not m[-1].any() and not m[:, 0].any() and not m[:, -1].any()
            cells = set(zip(p[:, 1].astype(int), p[:, 0].astype(int)))
            assert not any(m[j, i] for j, i in cells)
    # deterministic in the seed, different across seeds
    again = scenes.dam_break_batch(64, 24, 130, seed=7)
    assert all(np.array_equal(a.particles, b.particles) for a, b in zip(scs, again))
    other = scenes.dam_break_batch(64, 24, 130, seed=8)
    assert any(a.nx != b.nx for a, b in zip(scs, other))
