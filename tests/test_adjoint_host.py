"""Host logic of the reverse mode (no GPU): the checkpointed reverse-pass schedule and
the control-grid model."""
from __future__ import annotations

import numpy as np
import pytest

from fluid_simulation_amd import adjoint


class MockDomain:
    """Stands in for fsg.Domain: the state is the step index it holds."""

    def __init__(self):
        self.t = 0
        self.slots = {}
        self.adjoint_at = []
        self.forward_steps = 0

    def step(self, scheme, dt):
        self.t += 1
        self.forward_steps += 1

    def adjoint_step(self, scheme, dt):
        self.adjoint_at.append(self.t)
        self.t += 1  # the taped forward step advances the state

    def checkpoint_save(self, slot):
        self.slots[slot] = self.t

    def checkpoint_load(self, slot):
        self.t = self.slots[slot]

    def adjoint_get(self):
        return np.zeros((1, 4), np.float32), np.full((2, 2), 1.0), np.full((2, 2), 2.0)

    def adjoint_get_front(self):
        return np.full((2, 2), 10.0), np.full((2, 2), 20.0)


@pytest.mark.parametrize("steps,every", [(1, 1), (5, 2), (8, 8), (64, 8), (7, 3), (6, 10)])
def test_reverse_pass_visits_every_step_once_from_the_last(steps, every):
    dom = MockDomain()
    seen = {}

    def seed(d):
        seen["at"] = d.t
        return 1.25

    loss, _, gu, gv = adjoint.rollout_adjoint(dom, 3, 0.01, steps, every, seed)
    assert loss == 1.25
    # the velocity arrays set_state writes to both buffers: both adjoints summed
    assert np.all(gu == 11.0) and np.all(gv == 22.0)
    assert seen["at"] == steps                     # seeded at the final state
    assert dom.adjoint_at == list(reversed(range(steps)))
    assert dom.t == steps                          # left at the final state
    # forward steps: the run itself plus each segment's recompute (its length - 1)
    seg = [min(every, steps - s0) for s0 in range(0, steps, every)]
    assert dom.forward_steps == steps + sum(n - 1 for n in seg)


def test_control_model_is_bilinear_and_linear():
    torch = pytest.importorskip("torch")
    n = 20
    rng = np.random.default_rng(0)
    pos = rng.uniform(1.0, n - 1.0, size=(50, 2))
    m = adjoint.ControlModel(n, n, float(n), float(n), pos, k=6)
    # a constant control field gives that constant everywhere
    c = torch.full((2, 6, 6), 0.75, dtype=torch.float64)
    vel, u, v = m(c, dtype=torch.float64)
    assert torch.allclose(vel, torch.full_like(vel, 0.75))
    assert torch.allclose(u, torch.full_like(u, 0.75)) and torch.allclose(v, torch.full_like(v, 0.75))
    # linear in the controls
    c1 = torch.tensor(rng.standard_normal((2, 6, 6)))
    c2 = torch.tensor(rng.standard_normal((2, 6, 6)))
    a = m(2.0 * c1 - 3.0 * c2, dtype=torch.float64)
    b1, b2 = m(c1, dtype=torch.float64), m(c2, dtype=torch.float64)
    for x, y1, y2 in zip(a, b1, b2):
        assert torch.allclose(x, 2.0 * y1 - 3.0 * y2)
    # a field linear in x is reproduced exactly: node value a_x = x at node positions
    xs = np.linspace(0.0, n, 6)
    lin = torch.tensor(np.stack([np.tile(xs, (6, 1)), np.zeros((6, 6))]))
    vel, u, _ = m(lin, dtype=torch.float64)
    assert np.allclose(vel[:, 0].numpy(), pos[:, 0])
    assert np.allclose(u.numpy(), np.tile(np.arange(n, dtype=np.float64), (n, 1)))
