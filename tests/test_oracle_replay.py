"""Decision replay in the CPU oracle (the checker of the reverse mode's gradients).

A step that replays the decisions it recorded itself must be the same step, bit for
bit, over many steps of a dam break (free surface, wall contact, particles pushed back
from the walls); a replayed record with one decision changed must change the result.
The face record's bits are checked against the reference's rules on the state the
step saw (src/FluidSolver.cpp:898-917 weight test, :297-321 Dirichlet)."""
import numpy as np
import pytest

from fluid_simulation_amd import scenes
from oracle import OracleSim
from oracle.binding import SCHEME_FLIP, SCHEME_PIC, SCHEME_PICFLIP, SOLID
from oracle.replay import particle_density


@pytest.fixture(autouse=True)
def _lib(oracle_lib):
    return oracle_lib


def _record_run(sc, scheme, steps, precision="f32"):
    o = OracleSim.from_scene(sc, precision)
    o.set_params(cg_tol=1e-6, cg_max_iter=5000)
    o.record_faces(True)
    recs, states = [], []
    for _ in range(steps):
        o.step(scheme, sc.dt)
        recs.append((o.label.copy(), o.face_record.copy(), o.reject_flag.copy()))
        states.append((o.particles.copy(), *(a.copy() for a in o.vel("front"))))
    return recs, states


@pytest.mark.parametrize("scheme", [SCHEME_PIC, SCHEME_FLIP, SCHEME_PICFLIP])
def test_replaying_its_own_decisions_is_the_same_step(scheme):
    sc = scenes.dam_break(32)
    recs, states = _record_run(sc, scheme, 40)
    # the run saw every kind of decision
    assert any(r[2].any() for r in recs) or scheme != SCHEME_PICFLIP
    assert any(((r[1] >> 2) & 3).any() for r in recs)
    o = OracleSim.from_scene(sc, "f32")
    o.set_params(cg_tol=1e-6, cg_max_iter=5000)
    for (lab, face, rej), (p, u, v) in zip(recs, states):
        o.set_replay(label=lab, face=face, reject=rej if scheme == SCHEME_PICFLIP else None)
        o.step(scheme, sc.dt)
        assert np.array_equal(o.particles, p)
        uu, vv = o.vel("front")
        assert np.array_equal(uu, u) and np.array_equal(vv, v)


def test_a_changed_decision_changes_the_step():
    sc = scenes.dam_break(32)
    recs, states = _record_run(sc, SCHEME_PICFLIP, 6)
    lab, face, rej = recs[5]
    # change one first-pass Dirichlet decision at a zeroed wall face of a LIQUID cell
    liq = lab == 0
    liq_u = liq | np.roll(liq, 1, axis=1)
    jj, ii = np.nonzero((face & 4).astype(bool) & liq_u)
    assert len(jj)
    face2 = face.copy()
    face2[jj[0], ii[0]] &= ~np.uint8(4)
    o = OracleSim.from_scene(sc, "f32")
    o.set_params(cg_tol=1e-6, cg_max_iter=5000)
    for k, (l_, f_, r_) in enumerate(recs):
        o.set_replay(label=l_, face=face2 if k == 5 else f_, reject=r_)
        o.step(SCHEME_PICFLIP, sc.dt)
    assert not np.array_equal(o.particles, states[5][0])


def test_face_record_bits_follow_the_reference_rules():
    """P2G bit = weight test w > 1e-6 of the step's own accumulators; a first-pass
    Dirichlet bit only where the face touches a SOLID cell; every face next to SOLID
    with closed walls."""
    sc = scenes.dam_break(24)
    o = OracleSim.from_scene(sc, "f64")
    o.set_params(cg_tol=1e-10, cg_max_iter=5000)
    o.record_faces(True)
    for _ in range(10):
        o.step(SCHEME_PICFLIP, sc.dt)
    wu, wv = o.acc[1], o.acc[3]
    assert np.array_equal((o.face_record & 1) != 0, wu > 1e-6)
    assert np.array_equal((o.face_record & 2) != 0, wv > 1e-6)
    lab = o.label
    solid = lab == SOLID
    touch_u = solid | np.roll(solid, 1, axis=1)
    touch_u[:, 0] = solid[:, 0]
    touch_v = solid | np.roll(solid, 1, axis=0)
    touch_v[0, :] = solid[0, :]
    for bit, touch in ((4, touch_u), (8, touch_v), (16, touch_u), (32, touch_v)):
        assert not np.any(((o.face_record & bit) != 0) & ~touch)
    o.set_params(closed_walls=1)
    o.step(SCHEME_PICFLIP, sc.dt)
    lab = o.label
    solid = lab == SOLID
    touch_u = solid | np.roll(solid, 1, axis=1)
    touch_u[:, 0] = solid[:, 0]
    assert np.array_equal((o.face_record & 4) != 0, touch_u)


def test_particle_density_partition_of_unity():
    """Each particle splats total weight 1 when its four corners are distinct cells;
    a particle on a cell centre splats 1/4 to each corner of its cell square."""
    rng = np.random.default_rng(3)
    parts = np.zeros((500, 4))
    parts[:, :2] = rng.uniform(1.0, 15.0, (500, 2))
    rho = particle_density(parts, 16, 16, 1.0, 1.0)
    assert abs(rho.sum() - 500.0) < 1e-9
    one = particle_density(np.array([[3.5, 5.5, 0, 0]]), 16, 16, 1.0, 1.0)
    assert np.allclose(one[5:7, 3:5], 0.25) and abs(one.sum() - 1.0) < 1e-15
