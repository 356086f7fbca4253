"""FP64 oracle rollouts that replay a GPU run's discrete decisions.

TEST INFRASTRUCTURE ONLY (see oracle.h): the gradient checks of the reverse mode use
it as the checker; the product never imports it.

A PIC / FLIP step (src/FluidSolver.cpp:136-251) is a piecewise-smooth map of the
particle state and the back buffer: its discrete decisions (the labels of
classifyCells, src/FluidDomain.cpp:150-180; the P2G weight test,
src/FluidSolver.cpp:898-917; the two enforceDirichlet passes, :297-321; the obstacle
rejection of advectAndEnsureOutsideObstacles, src/MarkerParticleSet.cpp:48-62)
select the smooth piece.  A semi-Lagrangian step (:99-134) is one of the particles
and the front buffer, its decisions the labels and the two Dirichlet passes.  The adjoint kernels differentiate the piece the GPU run
took.  An FP64 rollout that takes the same decisions stays on that piece, so its
central differences are the derivative the adjoint must match, even where an FP64
rollout of its own would take a different branch within a few steps.
"""
from __future__ import annotations

import numpy as np

from .binding import OracleSim


def particle_density(parts, nx, ny, dx, dy):
    """Bilinear particle-splat density in FP64 (Grid::addToValueInterpolated,
    include/Grid.h:152-184, weight 1 per particle at its position): truncating cell
    index, fraction taken before the clamp, the upper corner clamped after +1."""
    p = np.asarray(parts, np.float64)
    xs, ys = p[:, 0] / dx, p[:, 1] / dy
    i, j = np.trunc(xs), np.trunc(ys)
    fi, fj = xs - i, ys - j
    i, j = i.astype(np.int64), j.astype(np.int64)
    i0, j0 = np.clip(i, 0, nx - 1), np.clip(j, 0, ny - 1)
    i1, j1 = np.clip(i + 1, 0, nx - 1), np.clip(j + 1, 0, ny - 1)
    rho = np.zeros(nx * ny, np.float64)
    for ii, jj, w in ((i0, j0, (1 - fi) * (1 - fj)), (i1, j0, fi * (1 - fj)),
                      (i0, j1, (1 - fi) * fj), (i1, j1, fi * fj)):
        np.add.at(rho, ii + jj * nx, w)
    return rho.reshape(ny, nx)


class ReplayRollout:
    """Rollouts of ``len(decisions)`` steps in FP64 from given initial states, each
    step replaying one record (labels (ny, nx) uint8, face decisions (ny, nx) uint8 as
    fsg_get_step_decisions returns them, reject flags (P,) uint8).

    ``params`` go to the oracle (closed_walls, project_mean, ...); the pressure solve
    runs to |r| / |b| <= ``cg_tol``."""

    def __init__(self, nx, ny, length_x, length_y, density, pic_ratio, decisions,
                 cg_tol=1e-13, **params):
        self.geom = (nx, ny, length_x, length_y, density, pic_ratio)
        self.decisions = list(decisions)
        self.params = dict(params, cg_tol=cg_tol, cg_max_iter=1000000)

    def run(self, particles, back_u, back_v, scheme, dt):
        """The final oracle state after the replayed steps from (particles (P, 4),
        back u, back v); the front buffer starts equal to the back one (a PIC / FLIP
        step reads the back, a semi-Lagrangian step the front)."""
        o = OracleSim(*self.geom, precision="f64")
        o.set_params(**self.params)
        o.set_particles(np.asarray(particles, np.float64))
        o.set_vel(back_u, back_v, "front")
        o.set_vel(back_u, back_v, "back")
        for lab, face, rej in self.decisions:
            o.set_replay(label=lab, face=face, reject=rej)
            o.step(scheme, dt)
        o.set_replay()
        return o

    def density_loss(self, particles, back_u, back_v, scheme, dt, target):
        """sum((density(final particles) - target)^2) in FP64, and the final sim."""
        o = self.run(particles, back_u, back_v, scheme, dt)
        nx, ny = self.geom[0], self.geom[1]
        rho = particle_density(o.particles, nx, ny, o.s.dx, o.s.dy)
        return float(((rho - np.asarray(target, np.float64)) ** 2).sum()), o

    def density_losses(self, states, scheme, dt, target, workers=8):
        """density_loss of each (particles, back u, back v) in ``states``, on a pool of
        threads (the oracle keeps no global state and ctypes releases the GIL)."""
        from concurrent.futures import ThreadPoolExecutor

        def one(st):
            return self.density_loss(*st, scheme, dt, target)[0]

        with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
            return list(ex.map(one, states))
