"""Reverse-mode differentiation of semi-Lagrangian, PIC and FLIP rollouts on the GPU.

The adjoint kernels (``fsg_adjoint_step``) carry the adjoint state of one step
(per particle d/d(x, y, vx, vy) and the front and back velocity buffers' d/d(u, v))
from the step's end to its start, with the step's discrete decisions held at the
forward run's.  A PIC / FLIP step reads only the back buffer of its start state, a
semi-Lagrangian step only the front.  ``rollout_adjoint`` drives a whole
run: checkpoints every ``every`` steps in the forward pass; in the reverse pass each
segment is recomputed from its checkpoint, its per-step states kept, and the steps'
adjoints applied from the last.

``Rollout`` wraps a rollout + particle-density loss as a ``torch.autograd.Function``
(the simulator itself runs in the HIP library; torch only carries the tensors), and
``ControlModel`` is the differentiable configuration: initial velocities from a
coarse control grid, bilinearly upsampled.
"""
from __future__ import annotations

import math

import numpy as np

from . import fsg


def forward_checkpointed(dom: "fsg.Domain", scheme: int, dt: float, steps: int, every: int):
    """Run ``steps`` steps, saving the state at every ``every``-th step start in slots
    0 .. ceil(steps/every) - 1 and the final state in slot ceil(steps/every) + every
    (the slots in between hold a segment's per-step states in the reverse pass)."""
    every = max(1, int(every))
    nseg = math.ceil(steps / every) if steps > 0 else 0
    for s in range(steps):
        if s % every == 0:
            dom.checkpoint_save(s // every)
        dom.step(scheme, dt)
    dom.checkpoint_save(nseg + every)


def reverse_pass(dom: "fsg.Domain", scheme: int, dt: float, steps: int, every: int, seed_fn):
    """After ``forward_checkpointed``: seed the adjoint at the final state with
    ``seed_fn(dom)`` (it sets the adjoint state, e.g. through
    ``adjoint_seed_density``, and returns the loss), then recompute each segment from
    its checkpoint, keep its per-step states and apply the steps' adjoints from the
    last.  Returns (loss, pbar (n, 4), gu, gv) at the initial state, gu / gv the
    adjoint of the velocity arrays ``set_state`` writes to both buffers (the back and
    the front buffer's adjoints summed); the domain is left at the final state."""
    every = max(1, int(every))
    nseg = math.ceil(steps / every) if steps > 0 else 0
    final_slot = nseg + every
    dom.checkpoint_load(final_slot)
    loss = seed_fn(dom)
    for seg in reversed(range(nseg)):
        s0 = seg * every
        s1 = min(steps, s0 + every)
        dom.checkpoint_load(seg)
        for s in range(s0, s1):
            dom.checkpoint_save(nseg + (s - s0))
            if s + 1 < s1:
                dom.step(scheme, dt)
        for s in reversed(range(s0, s1)):
            dom.checkpoint_load(nseg + (s - s0))
            dom.adjoint_step(scheme, dt)
    pbar, bu, bv = dom.adjoint_get()
    fu, fv = dom.adjoint_get_front()
    dom.checkpoint_load(final_slot)
    return loss, pbar, bu + fu, bv + fv


def rollout_adjoint(dom: "fsg.Domain", scheme: int, dt: float, steps: int, every: int,
                    seed_fn):
    """``forward_checkpointed`` then ``reverse_pass`` from the domain's state."""
    forward_checkpointed(dom, scheme, dt, steps, every)
    return reverse_pass(dom, scheme, dt, steps, every, seed_fn)


def rollout_decisions(dom: "fsg.Domain", scheme: int, dt: float, steps: int):
    """Step the domain ``steps`` times from its state and return each step's discrete
    decisions: (labels, face decisions, reject flags) per step.  Needs
    ``record_decisions = 1``; with ``deterministic = 1`` they are the decisions every
    rerun from the same state (a checkpointed recompute) takes, i.e. the branch of
    the piecewise-smooth rollout the adjoint differentiates."""
    recs = []
    for _ in range(steps):
        dom.step(scheme, dt)
        recs.append((dom.labels(), dom.step_decisions(), dom.reject_flags()))
    return recs


def density_loss_seed(target: np.ndarray):
    """seed_fn for the loss sum((rho - target)^2) (FP64 sum) of the particle density."""
    target = np.asarray(target, np.float64)

    def seed(dom):
        rho = dom.density().astype(np.float64)
        diff = rho - target
        dom.adjoint_seed_density((2.0 * diff).astype(np.float32))
        return float((diff * diff).sum())

    return seed


def state_of(dom: "fsg.Domain"):
    """(particles, back u, back v) of the domain."""
    bu, bv = dom.velocity(fsg.VEL_BACK)
    return dom.particles(), bu, bv


def set_state(dom: "fsg.Domain", particles, back_u, back_v):
    """The carried state a step reads: particles and one velocity field, written to
    both buffers (a PIC / FLIP step reads the back buffer, a semi-Lagrangian step the
    front)."""
    dom.set_particles(np.asarray(particles, np.float32))
    dom.set_velocity(back_u, back_v, fsg.VEL_BACK)
    dom.set_velocity(back_u, back_v, fsg.VEL_FRONT)


try:  # torch is plumbing here: the wrapper is defined only where torch imports
    import torch

    class Rollout(torch.autograd.Function):
        """loss = sum((density(final particles) - target)^2) after ``steps`` steps from
        (positions, velocities, grid u, grid v) (the grid written to both buffers);
        gradients by the adjoint kernels."""

        @staticmethod
        def forward(ctx, pos, vel, back_u, back_v, dom, scheme, dt, steps, every, target):
            parts = torch.cat([pos, vel], dim=1).detach().cpu().numpy().astype(np.float32)
            set_state(dom, parts, back_u.detach().cpu().numpy(), back_v.detach().cpu().numpy())
            ctx.dom, ctx.scheme, ctx.dt, ctx.steps, ctx.every = dom, scheme, dt, steps, every
            ctx.target = np.asarray(target, np.float64)
            ctx.device, ctx.dtype = pos.device, pos.dtype
            forward_checkpointed(dom, scheme, dt, steps, every)
            rho = dom.density().astype(np.float64)
            loss = float(((rho - ctx.target) ** 2).sum())
            return torch.tensor(loss, dtype=torch.float64, device=pos.device)

        @staticmethod
        def backward(ctx, gloss):
            g = float(gloss)
            tgt = ctx.target

            def seed(d):
                rho = d.density().astype(np.float64)
                d.adjoint_seed_density((2.0 * g * (rho - tgt)).astype(np.float32))
                return 0.0

            _, pbar, bu, bv = reverse_pass(ctx.dom, ctx.scheme, ctx.dt, ctx.steps, ctx.every,
                                           seed)
            to = dict(device=ctx.device, dtype=ctx.dtype)
            return (torch.as_tensor(pbar[:, :2], **to), torch.as_tensor(pbar[:, 2:], **to),
                    torch.as_tensor(bu, **to), torch.as_tensor(bv, **to),
                    None, None, None, None, None, None)

    class _Gather4(torch.autograd.Function):
        """out[n] = sum_k w[n, k] c[idx[n, k]]; the backward is one weighted bincount."""

        @staticmethod
        def forward(ctx, cflat, idx, w):
            ctx.save_for_backward(idx, w)
            ctx.size = cflat.numel()
            return (w * cflat[idx]).sum(1)

        @staticmethod
        def backward(ctx, g):
            idx, w = ctx.saved_tensors
            gc = torch.bincount(idx.reshape(-1), weights=(w * g[:, None]).reshape(-1),
                                minlength=ctx.size)
            return gc.to(g.dtype), None, None

    class ControlModel:
        """Initial velocities from a (2, K, K) control grid c: the bilinear field
        U(x, y) through control nodes at (a L_x / (K-1), b L_y / (K-1)).  The back
        (and front) buffer samples it at the face centres (u at (i dx, (j + 1/2) dy),
        v at ((i + 1/2) dx, j dy)) and every particle at its position.  Each sample is
        a gather of its 4 control nodes with fixed weights (FP64), so the model and its
        backward (a scatter-add) are linear in c."""

        def __init__(self, nx, ny, length_x, length_y, positions, k=16):
            self.nx, self.ny, self.k = nx, ny, k
            dx, dy = length_x / nx, length_y / ny
            self.lx, self.ly = length_x, length_y
            i = np.arange(nx, dtype=np.float64)
            j = np.arange(ny, dtype=np.float64)
            ux, uy = np.meshgrid(i * dx, (j + 0.5) * dy)
            vx, vy = np.meshgrid((i + 0.5) * dx, j * dy)
            self._wu = self._weights(ux.ravel(), uy.ravel())
            self._wv = self._weights(vx.ravel(), vy.ravel())
            pos = np.asarray(positions, np.float64)
            self._wp = self._weights(pos[:, 0], pos[:, 1])

        def _weights(self, x, y):
            """(node indices (n, 4), weights (n, 4)) of the bilinear samples at (x, y)."""
            k = self.k
            sx = np.clip(x / self.lx * (k - 1), 0.0, k - 1 - 1e-12)
            sy = np.clip(y / self.ly * (k - 1), 0.0, k - 1 - 1e-12)
            i0, j0 = np.floor(sx).astype(np.int64), np.floor(sy).astype(np.int64)
            fx, fy = sx - i0, sy - j0
            idx = np.stack([i0 + j0 * k, i0 + 1 + j0 * k, i0 + (j0 + 1) * k,
                            i0 + 1 + (j0 + 1) * k], 1)
            w = np.stack([(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy], 1)
            return torch.from_numpy(idx), torch.from_numpy(w)

        @staticmethod
        def _sample(wt, cflat):
            return _Gather4.apply(cflat, wt[0], wt[1])

        def __call__(self, c, dtype=torch.float32):
            """(particle velocities (P, 2), back u, back v) for controls c (2, K, K)."""
            cu = c[0].reshape(-1).to(torch.float64)
            cv = c[1].reshape(-1).to(torch.float64)
            u = self._sample(self._wu, cu).reshape(self.ny, self.nx)
            v = self._sample(self._wv, cv).reshape(self.ny, self.nx)
            vel = torch.stack([self._sample(self._wp, cu), self._sample(self._wp, cv)], 1)
            return vel.to(dtype), u.to(dtype), v.to(dtype)

except ImportError:  # pragma: no cover - torch is part of this image
    torch = None
