"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np

from fluid_simulation_amd import scenes


def rel_err(a, b) -> float:
    """max|a - b| / max(max|b|, 1e-30) (the per-field relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    return float(np.abs(a - b).max() / max(float(np.abs(b).max()), 1e-30))


def oracle_after(n=128, steps=0, precision="f32", tol=1e-6, max_iter=10000, scheme=3):
    """Dam break advanced `steps` PICFLIP steps by the oracle."""
    from oracle import OracleSim
    sc = scenes.dam_break(n)
    sim = OracleSim.from_scene(sc, precision)
    sim.set_params(cg_tol=tol, cg_max_iter=max_iter)
    for _ in range(steps):
        sim.step(scheme, sc.dt)
    return sc, sim


def copy_state(src, dst):
    """Copy the full carried state of one oracle sim into another (e.g. f32 -> f64)."""
    dst.set_particles(src.particles)
    for which in ("front", "back"):
        u, v = src.vel(which)
        dst.set_vel(u, v, which)
    dst.u_prev[...] = src.u_prev
    dst.v_prev[...] = src.v_prev
    dst.u_diff[...] = src.u_diff
    dst.v_diff[...] = src.v_diff
    dst.label[...] = src.label
    if src.solid_mask is not None:
        dst.set_solid_mask(src.solid_mask)
    if src.phi is not None:
        dst.set_phi(src.phi)


def gpu_from_oracle(sim, scene, **params):
    """A GPU Domain holding exactly the oracle's carried state."""
    from fluid_simulation_amd import fsg
    d = fsg.Domain(sim.nx, sim.ny, scene.length_x, scene.length_y, scene.density,
                   scene.pic_ratio)
    kw = dict(cg_tol=float(sim.s.cg_tol), cg_max_iter=int(sim.s.cg_max_iter))
    kw.update(params)
    d.set_params(**kw)
    if len(sim.particles):
        d.set_particles(sim.particles.astype(np.float32))
    uf, vf = sim.vel("front")
    ub, vb = sim.vel("back")
    d.set_velocity(uf, vf, fsg.VEL_FRONT)
    d.set_velocity(ub, vb, fsg.VEL_BACK)
    d.set_velocity(sim.u_prev, sim.v_prev, fsg.VEL_PREV)
    d.set_velocity(sim.u_diff, sim.v_diff, fsg.VEL_DIFF)
    d.set_labels(sim.label)
    if sim.solid_mask is not None:
        d.set_solid_mask(sim.solid_mask)
    if sim.phi is not None:
        d.set_level_set(sim.phi)
    return d
