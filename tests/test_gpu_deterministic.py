"""The deterministic form (fsg_params.deterministic: fixed-point P2G and density splats)
and the step decision record (fsg_params.record_decisions).

- A deterministic run is reproducible bit for bit, also with the particles' storage
  order shuffled (integer sums do not depend on the order of the atomics).
- Its P2G grid is within 1e-6 of the FP64 oracle's (tighter than the float atomics:
  each sum is the exact sum of the float terms, rounded once).
- The decision record of a developed dam-break step equals the oracle's record of
  the same step (src/FluidSolver.cpp:898-917 weight test, :297-321 Dirichlet), and
  the labels and reject flags equal the oracle's.
"""
from __future__ import annotations

import numpy as np
import pytest

from fluid_simulation_amd import fsg, scenes

from helpers import copy_state, gpu_from_oracle, oracle_after, rel_err

pytestmark = pytest.mark.gpu


def _run(sc, parts, steps, **params):
    d = fsg.Domain.from_scene(sc, solver=fsg.SOLVER_MG, cg_tol=1e-8, cg_max_iter=200,
                              deterministic=1, **params)
    d.set_particles(parts)
    for _ in range(steps):
        d.step(fsg.PICFLIP, sc.dt)
    out = (d.particles(), *d.velocity(fsg.VEL_FRONT), d.density())
    d.close()
    return out


def test_deterministic_run_is_reproducible_in_any_particle_order():
    sc = scenes.dam_break(256)
    print(f"dam break 256, {len(sc.particles)} particles, 30 PICFLIP steps", flush=True)
    a = _run(sc, sc.particles, 30)
    b = _run(sc, sc.particles, 30)
    perm = np.random.default_rng(5).permutation(len(sc.particles))
    c = _run(sc, sc.particles[perm], 30)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    # the shuffled run: the same particles (in the shuffled order), the same grids
    assert np.array_equal(a[0][perm].view(np.uint32), c[0].view(np.uint32))
    for x, y in zip(a[1:], c[1:]):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_deterministic_p2g_against_fp64():
    from oracle import OracleSim
    sc, o32 = oracle_after(128, 40)
    o = OracleSim(sc.nx, sc.ny, sc.length_x, sc.length_y, sc.density, sc.pic_ratio, "f64")
    copy_state(o32, o)
    g = gpu_from_oracle(o32, sc, deterministic=1)
    for s in (o, g):
        s.run_phase("CLASSIFY")
        s.run_phase("P2G")
    u, v = g.velocity(fsg.VEL_FRONT)
    uo, vo = o.vel("front")
    eu, ev = rel_err(u, uo), rel_err(v, vo)
    print(f"deterministic p2g vs fp64: u {eu:.2e} v {ev:.2e}", flush=True)
    g.close()
    assert eu <= 1e-6 and ev <= 1e-6


@pytest.mark.parametrize("closed", [0, 1])
def test_decision_record_matches_the_oracle(closed):
    sc, o = oracle_after(96, 60)
    o.set_params(closed_walls=closed, cg_tol=1e-8)
    g = gpu_from_oracle(o, sc, solver=fsg.SOLVER_MG, deterministic=1, record_decisions=1,
                        closed_walls=closed, cg_tol=1e-8, cg_max_iter=200)
    o.record_faces(True)
    o.step(3, sc.dt)
    g.step(fsg.PICFLIP, sc.dt)
    dec = g.step_decisions()
    assert np.array_equal(g.labels(), o.label)
    assert np.array_equal(g.reject_flags(), o.reject_flag)
    diff = np.nonzero(dec != o.face_record)
    print(f"closed {closed}: {int((dec >> 2 & 3).astype(bool).sum())} first-pass zeroed cells, "
          f"{int((dec >> 4 & 3).astype(bool).sum())} second-pass, "
          f"{len(diff[0])} cells differ", flush=True)
    assert np.array_equal(dec, o.face_record)
    g.close()
