"""The multigrid-preconditioned CG (FSG_SOLVER_MG) against the FP64 oracle on
irregular domains: obstacles, free surfaces, odd and elongated grid shapes, the
all-Neumann closed box, and determinism.  The solve replaces Eigen's
ConjugateGradient (src/FluidSolver.cpp:362-425); the velocity update after it is
src/FluidSolver.cpp:427-482."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from fluid_simulation_amd import fsg

from helpers import gpu_from_oracle, oracle_after, rel_err

pytestmark = pytest.mark.gpu

LIQ, AIR, SOL = 0, 1, 2


def irregular_labels(nx, ny, seed, liquid_frac=0.7, n_obst=3):
    """SOLID border, a few SOLID rectangles, LIQUID / AIR noise elsewhere; every
    LIQUID component touches AIR (so the system is nonsingular) and no LIQUID
    cell has four SOLID neighbours (no empty row)."""
    rng = np.random.default_rng(seed)
    lab = np.where(rng.random((ny, nx)) < liquid_frac, LIQ, AIR).astype(np.uint8)
    for _ in range(n_obst):
        w, h = rng.integers(1, max(2, nx // 4)), rng.integers(1, max(2, ny // 4))
        x0, y0 = rng.integers(0, nx - w + 1), rng.integers(0, ny - h + 1)
        lab[y0:y0 + h, x0:x0 + w] = SOL
    lab[0, :] = lab[-1, :] = SOL
    lab[:, 0] = lab[:, -1] = SOL
    comp, n = ndimage.label(lab == LIQ)
    air_nb = ndimage.binary_dilation(lab == AIR)
    for k in range(1, n + 1):
        if not (air_nb & (comp == k)).any():
            lab[comp == k] = AIR
    return lab


def _solve_pair(lab, seed, tol=1e-10, project_mean=0, length=None, smoother=1, **gpu_params):
    from oracle import OracleSim
    ny, nx = lab.shape
    lx, ly = (float(nx), float(ny)) if length is None else length
    rng = np.random.default_rng(seed + 100)
    u = rng.standard_normal((ny, nx)).astype(np.float32)
    v = rng.standard_normal((ny, nx)).astype(np.float32)
    d = OracleSim(nx, ny, lx, ly, 0.01, 0.05, "f64")
    d.set_params(cg_tol=1e-13, cg_max_iter=200000, project_mean=project_mean)
    d.label[...] = lab
    d.set_vel(u, v, "front")
    d.set_vel(np.zeros_like(u), np.zeros_like(v), "back")
    d.run_phase("PRESSURE", 0.01)
    g = fsg.Domain(nx, ny, lx, ly, 0.01, 0.05)
    g.set_params(solver=fsg.SOLVER_MG, cg_tol=tol, cg_max_iter=500, project_mean=project_mean,
                 mg_smoother=smoother, **gpu_params)
    g.set_labels(lab)
    g.set_velocity(u, v, fsg.VEL_FRONT)
    g.set_velocity(np.zeros_like(u), np.zeros_like(v), fsg.VEL_BACK)
    print(f"MG pressure {nx}x{ny} seed {seed}", flush=True)
    g.run_phase("PRESSURE", 0.01)
    return d, g


@pytest.mark.parametrize("smoother", [1, 0], ids=["chebyshev", "jacobi"])
@pytest.mark.parametrize("nx,ny,seed", [(64, 64, 0), (77, 133, 1), (200, 50, 2), (7, 5, 3),
                                        (513, 257, 4), (1030, 20, 5)])
def test_mg_irregular_against_fp64(nx, ny, seed, smoother):
    lab = irregular_labels(nx, ny, seed)
    d, g = _solve_pair(lab, seed, smoother=smoother)
    st = g.stats()
    liq = lab == LIQ
    ep = rel_err(g.pressure()[liq], d.pressure[liq])
    u, v = g.velocity(fsg.VEL_FRONT)
    ud, vd = d.vel("front")
    eu, ev = rel_err(u, ud), rel_err(v, vd)
    print(f"F {st.n_fluid} iters {st.cg_iterations} |r|/|b| {st.cg_error:.2e}; "
          f"vs fp64: p {ep:.2e} u {eu:.2e} v {ev:.2e}")
    assert st.n_fluid == int(liq.sum())
    assert st.cg_error <= 1e-10
    assert st.cg_iterations <= 60
    assert ep <= 1e-6 and eu <= 1e-6 and ev <= 1e-6


@pytest.mark.parametrize("sweeps,fuse", [(8, 1), (16, 0), (0, 0)])
def test_mg_coarse_and_fusion_options_against_fp64(sweeps, fuse):
    """The truncated coarse solve (mg_coarse_sweeps: Jacobi sweeps on the first
    one-workgroup level instead of its V-cycle) and the unfused update kernel
    (mg_fuse_update = 0) change the preconditioner and the launch sequence, not the
    system: the same FP64 residual bound and the same solution."""
    lab = irregular_labels(513, 257, 4)
    d, g = _solve_pair(lab, 4, mg_coarse_sweeps=sweeps, mg_fuse_update=fuse)
    st = g.stats()
    liq = lab == LIQ
    ep = rel_err(g.pressure()[liq], d.pressure[liq])
    print(f"coarse sweeps {sweeps} fuse {fuse}: {st.cg_iterations} iterations, "
          f"|r|/|b| {st.cg_error:.2e}, p {ep:.2e}")
    assert st.cg_error <= 1e-10
    assert ep <= 1e-6


@pytest.mark.parametrize("threads", [512, 1024])
@pytest.mark.parametrize("fuse", [1, 0])
def test_mg_level0_tile_threads_against_fp64(threads, fuse):
    """Level-0 tiles of 512 or 1024 threads (mg_tile_threads) in the fused and the
    unfused form: the same system solved to the same FP64 residual bound."""
    lab = irregular_labels(513, 257, 4)
    d, g = _solve_pair(lab, 4, mg_tile_threads=threads, mg_fuse_update=fuse)
    st = g.stats()
    liq = lab == LIQ
    ep = rel_err(g.pressure()[liq], d.pressure[liq])
    print(f"tile threads {threads} fuse {fuse}: {st.cg_iterations} iterations, "
          f"|r|/|b| {st.cg_error:.2e}, p {ep:.2e}")
    assert st.cg_error <= 1e-10
    assert ep <= 1e-6


def test_mg_closed_box_project_mean():
    """All-LIQUID interior in a SOLID box: K is singular (constants); with the mean
    of b removed the pressure is defined up to a constant and the velocity update
    is unique."""
    n = 96
    lab = np.full((n, n), LIQ, np.uint8)
    lab[0, :] = lab[-1, :] = SOL
    lab[:, 0] = lab[:, -1] = SOL
    d, g = _solve_pair(lab, 7, tol=1e-9, project_mean=1)
    liq = lab == LIQ
    pg, pd = g.pressure()[liq].astype(np.float64), d.pressure[liq]
    ep = rel_err(pg - pg.mean(), pd - pd.mean())
    u, v = g.velocity(fsg.VEL_FRONT)
    ud, vd = d.vel("front")
    st = g.stats()
    print(f"closed box: iters {st.cg_iterations} err {st.cg_error:.2e} p {ep:.2e} "
          f"u {rel_err(u, ud):.2e} v {rel_err(v, vd):.2e}")
    assert st.cg_error <= 1e-9
    assert ep <= 1e-5
    assert max(rel_err(u, ud), rel_err(v, vd)) <= 1e-5


def test_mg_project_mean_with_air_is_nonsingular():
    """A SOLID box of LIQUID with a few AIR cells inside: with the mean of b removed
    (project_mean) the system is still non-singular (the AIR neighbours are Dirichlet
    links), so its solution is unique and no constant may be added: the pressure
    itself, not only its deviation from the mean, matches the FP64 oracle's
    Jacobi-PCG from x0 = 0, and so does the velocity update."""
    n = 96
    lab = np.full((n, n), LIQ, np.uint8)
    lab[0, :] = lab[-1, :] = SOL
    lab[:, 0] = lab[:, -1] = SOL
    lab[40, 50] = lab[41, 50] = lab[70, 20] = AIR
    d, g = _solve_pair(lab, 9, tol=1e-9, project_mean=1)
    liq = lab == LIQ
    pg, pd = g.pressure()[liq].astype(np.float64), d.pressure[liq]
    ep = rel_err(pg, pd)
    u, v = g.velocity(fsg.VEL_FRONT)
    ud, vd = d.vel("front")
    print(f"box with AIR: p {ep:.2e} u {rel_err(u, ud):.2e} v {rel_err(v, vd):.2e}")
    assert ep <= 1e-5
    assert max(rel_err(u, ud), rel_err(v, vd)) <= 1e-5


def test_mg_deterministic_and_no_fluid():
    lab = irregular_labels(150, 90, 11)
    _, g1 = _solve_pair(lab, 11)
    _, g2 = _solve_pair(lab, 11)
    assert np.array_equal(g1.pressure().view(np.uint32), g2.pressure().view(np.uint32))
    # no LIQUID cell: the solve is skipped exactly like the plain form
    lab0 = np.where(lab == LIQ, AIR, lab).astype(np.uint8)
    d, g = _solve_pair(lab0, 12)
    assert g.stats().n_fluid == 0
    u, v = g.velocity(fsg.VEL_FRONT)
    ud, vd = d.vel("front")
    assert np.array_equal(u, ud.astype(np.float32)) and np.array_equal(v, vd.astype(np.float32))


@pytest.mark.parametrize("graph_iters,fuse", [(2, 1), (0, 1), (0, 0), (3, 0)])
def test_graph_replay_matches_eager_launches(graph_iters, fuse):
    """The multigrid iterations replayed from captured HIP graphs (the default) give
    the pressure of the same solve launched kernel by kernel, bit for bit, over
    several solves of different right-hand sides on one handle (every replay
    compared, with eager work and synchronisations in between): graphs of graph_iters
    iterations, or (0) one graph per solve length followed by one per further
    iteration, for the fused and the separate update."""
    sc, sim = oracle_after(96, steps=3)
    rng = np.random.default_rng(21)
    doms = [gpu_from_oracle(sim, sc, cg_tol=1e-9, cg_max_iter=300, solver=fsg.SOLVER_MG,
                            graph=g, graph_iters=graph_iters, mg_fuse_update=fuse)
            for g in (1, 0)]
    u0, v0 = sim.vel("front")
    for k in range(4):
        du = (rng.standard_normal(u0.shape) * 0.1 * k).astype(np.float32)
        dv = (rng.standard_normal(v0.shape) * 0.1 * k).astype(np.float32)
        out = []
        for d in doms:
            d.set_velocity(u0 + du, v0 + dv, fsg.VEL_FRONT)
            d.run_phase("PRESSURE", sc.dt)
            out.append((d.pressure(), d.stats().cg_iterations))
        (pg, ig), (pe, ie) = out
        print(f"solve {k}: {ig} iterations (graph) / {ie} (eager)")
        assert ig == ie
        assert np.array_equal(pg.view(np.uint32), pe.view(np.uint32))


@pytest.mark.parametrize("sweeps,sweeps0", [(2, 2), (3, 3), (4, 2), (5, 3), (6, 2), (6, 3)])
@pytest.mark.parametrize("nx,ny,seed", [(513, 257, 21), (1030, 20, 22), (300, 300, 23)])
def test_mg_level_sweeps_against_fp64(nx, ny, seed, sweeps, sweeps0):
    """2 .. 6 smoothing sweeps on the tile levels below level 0 (mg_level_sweeps) and 2
    or 3 on level 0 (mg_level0_sweeps; more than two: the degree-S Chebyshev weights):
    a different symmetric preconditioner, the same system solved to the same FP64
    residual bound."""
    lab = irregular_labels(nx, ny, seed)
    d, g = _solve_pair(lab, seed, mg_level_sweeps=sweeps, mg_level0_sweeps=sweeps0)
    st = g.stats()
    liq = lab == LIQ
    ep = rel_err(g.pressure()[liq], d.pressure[liq])
    print(f"level sweeps {sweeps}, level 0 {sweeps0}: {st.cg_iterations} iterations, "
          f"|r|/|b| {st.cg_error:.2e}, p vs fp64 {ep:.2e}")
    assert st.cg_error <= 1e-10
    assert ep <= 1e-6


@pytest.mark.parametrize("nx,ny,seed", [(520, 390, 31), (1030, 257, 32)])
def test_spread_tiles_against_fp64(nx, ny, seed):
    """The level-0 tile kernels with the live tiles dealt out evenly over the XCDs
    (mg_spread_tiles 1, a block -> tile map formed per solve) and with one band of
    consecutive tiles per XCD (0): the same system solved to the same FP64 residual
    bound, on a grid whose unknowns fill only some tiles (a LIQUID block in one corner
    and a few pools elsewhere; the tile count is not a multiple of 8)."""
    rng = np.random.default_rng(seed)
    lab = np.full((ny, nx), AIR, np.uint8)
    lab[0, :] = lab[-1, :] = lab[:, 0] = lab[:, -1] = SOL
    lab[1:ny * 3 // 4, 1:nx * 2 // 5] = LIQ
    for _ in range(4):
        i, j = int(rng.integers(nx // 2, nx - 40)), int(rng.integers(1, ny - 40))
        lab[j:j + 30, i:i + 30] = LIQ
    out = []
    for spread in (1, 0):
        d, g = _solve_pair(lab, seed, mg_spread_tiles=spread)
        st = g.stats()
        liq = lab == LIQ
        ep = rel_err(g.pressure()[liq], d.pressure[liq])
        print(f"spread {spread}: {st.cg_iterations} iterations, {st.mg_live_tiles} live tiles, "
              f"|r|/|b| {st.cg_error:.2e}, p vs fp64 {ep:.2e}")
        assert st.cg_error <= 1e-10
        assert ep <= 1e-6
        out.append(g.pressure())
    assert rel_err(out[0], out[1]) <= 1e-7
