"""Benchmark: FLIP steps/s on the 1024x1024 dam break (one MI355X).

One step = FluidSolver::stepPICFLIP (src/FluidSolver.cpp:211-251) on the GPU:
classify, P2G, gravity, Dirichlet, 2-layer extension, pressure projection to a
relative residual of 1e-6, Dirichlet, PIC/FLIP G2P (0.95 FLIP), advection with
obstacle rejection.  The chosen pressure solver is the multigrid-preconditioned
CG with FP64 solution and residual; the plain form is FP32 Jacobi-PCG.  Workload: 1024x1024 MAC grid, unit cells,
fluid block 409 x 614 cells at (1,1), 4 jittered particles per fluid cell
(1,004,504 particles), dt = density = 0.01.  Inputs are resident in HBM before
the timed region; W warmup steps are the first W steps of the trajectory and the
K timed steps continue it.

The printed JSON line reports the chosen solver form ("value"), the plain FP32
3-kernel form timed over the same trajectory segment, the dominant kernel's
achieved bandwidth against the 8 TB/s HBM peak (its algorithmic bytes per launch
over its average duration, timed with HIP events on the solver's stream inside the
timed region; the dominant kernel is picked from a per-kernel profile of a few
untimed steps before it), the pressure solver's aggregate bandwidth, the
single-step error of the chosen form against the FP64 oracle, and the FP32 CPU
oracle on one core.

Two more workloads are reported under "config" (not in "value"):
- "batch": 256 non-square dam breaks of sides 64..512 with 1-4 obstacles each, stepped
  together as one packed batch, beside the same scenes stepped one Domain at a time
  and the CPU oracle on the first 16 scenes;
- "grid_only": semi-Lagrangian advection + projection on the fully fluid 2048^2 grid;
- "adjoint": the reverse-mode gradient of a 64-step 512^2 rollout (closed all-fluid
  box, density loss) through the adjoint kernels, timed, beside FP32 differences.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FLIP steps/s @1024² (chosen vs plain GPU, ≤1e-5 rel err) + CG HBM GB/s"
HBM_PEAK_GBS = 8000.0


def mg_levels(nx, ny, coarse_cells=1024):
    """Level sizes of the multigrid hierarchy (halved, rounded up, to <= 4 cells a
    side) and the first level solved in one workgroup (<= 1024 cells)."""
    lv = [(nx, ny)]
    while lv[-1][0] > 4 or lv[-1][1] > 4:
        lv.append(((lv[-1][0] + 1) // 2, (lv[-1][1] + 1) // 2))
    lc = next(l for l, (a, b) in enumerate(lv) if a * b <= coarse_cells)
    return [a * b for a, b in lv], lc


def kernel_bytes(name, nx, ny, P, levels=None):
    """Algorithmic HBM bytes of one launch of a kernel over dense cell grids, each
    array read or written once: nx * ny cells, P particles.  `levels` = (cells per
    level, first one-workgroup level) overrides the single-grid hierarchy (a batch's
    packed levels, whose last one is cycled scene by scene).  Level-0 tile kernels of
    one domain skip the tiles without an unknown: solve_bytes counts those."""
    C = nx * ny
    cells, lc = levels or mg_levels(nx, ny)
    table = {
        # particles 16 B read + label 1 B; float4 accumulators read and written once
        "particle_scatter": 17.0 * P + 32.0 * C,
        # particles 16 B read + written; 4 grids sampled (16 B per cell); labels 1 B
        "particle_gather": 32.0 * P + 17.0 * C,
        # info 1, p 4, x_old 8, b 8 read; x_new 8, r 4 written
        "cg_mg_xr": 33.0 * C,
        # info 1, z 4, p_old 4 read; p_new 4 written
        "cg_mg_dir": 13.0 * C,
        # right-hand side and info of the one-workgroup levels read, solution written
        "cg_mg_coarse": 8.0 * cells[lc] + float(sum(cells[lc:])),
    }
    if name in table:
        return table[name]
    for pre, fine_b, coarse_b in (("cg_mg_down", 9.0, 5.0), ("cg_mg_up", 13.0, 5.0)):
        if name.startswith(pre):
            tail = name[len(pre):].rstrip("+")
            l = int(tail)
            b = fine_b * cells[l] + coarse_b * cells[l + 1]
            if pre == "cg_mg_down" and levels is None and l == lc - 1 and lc >= 2:
                # the last tile level's down kernel also runs the one-workgroup stage
                b += table["cg_mg_coarse"]
            return b
    return None


# Level-0 tiles of one domain: 32 x 32 cells (and their 16 x 16 level-1 cells).  The
# level-0 kernels process only the tiles holding an unknown ("live"); the others
# write their zero outputs in the first iterations of a solve and return.
TILE = 32 * 32


def solve_bytes(name, nx, ny, solves):
    """Algorithmic bytes of every launch of solver kernel `name` over a list of
    solves (iterations n, live level-0 tiles L) of one domain (two or more tile
    levels: the fused form).  A solve of n iterations runs level 0's down kernel n + 1
    times (iteration 0 without the update, 1 .. n with the previous iteration's x /
    residual update; the stop test of iteration n fires in level 1's down kernel) and
    every other kernel n times; launches enqueued after the stop test move nothing."""
    cells, lc = mg_levels(nx, ny)
    tiles = ((nx + 31) // 32) * ((ny + 31) // 32)
    tot = 0.0
    for n, live in solves:
        dead = tiles - live
        if name == "cg_mg_down0" and lc >= 2:
            # live tile: r, info read and xs written on 1024 cells (9 B), r and info of
            # the 256 level-1 cells (5 B); with the update also info, p, x_old, b read
            # and x_new, r written (33 B per cell).  Dead tiles: zero xs and level-1 r
            # at iteration 0, zero x_new (8 B) and r (4 B) at 1, x_new at 2.
            plain = 9.0 * TILE + 5.0 * 256
            tot += live * plain + dead * (4.0 * TILE + 4.0 * 256)
            tot += n * live * (plain + 33.0 * TILE)
            tot += dead * ((12.0 * TILE if n >= 1 else 0.0) + (8.0 * TILE if n >= 2 else 0.0))
        elif name == "cg_mg_up0" and lc >= 1:
            # live tile: r, xs, info read, z written (13 B), level-1 x and info (5 B);
            # dead tiles write z = 0 at iteration 0
            tot += n * live * (13.0 * TILE + 5.0 * 256) + (dead * 4.0 * TILE if n >= 1 else 0.0)
        elif name == "cg_mg_dir" and lc >= 1:
            # live tile: info, z, p_old read, p_new written (13 B); dead tiles write
            # p = 0 at iterations 0 and 1
            tot += n * live * 13.0 * TILE + dead * 4.0 * TILE * min(n, 2)
        else:
            b = kernel_bytes(name, nx, ny, 0)
            if b is None:
                return None
            tot += n * b
    return tot


# HBM traffic per launch on the 1024^2 dam break (bytes), from two rocprofv3 --pmc
# passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, averaged over the dispatches
# of the last 8 steps, 250-257) of
#   bench.py --graph 0 --steps 5 --warmup 2 --start 250 --run-steps 0 --plain-steps 0
#            --refined-steps 0 --profile-steps 1 --no-fp64-check --no-cpu
#            --batch-scenes 0 --grid-steps 0 --adj-steps 0 --autotune 0
# scaled by a 1 GiB device copy (bench.py --copy-probe 1024) read by the same counters:
# FETCH_SIZE reads 0.5 GiB for its 1 GiB read (factor 2.0), WRITE_SIZE 1 GiB (1.0).
# traffic = (2.0 * FETCH_SIZE + 1.0 * WRITE_SIZE) * 1024.
MEASURED_TRAFFIC_1024 = {
    "cg_mg_down0": (2.0 * 3201.4 + 6303.8) * 1024,
    "cg_mg_up0": (2.0 * 1828.3 + 1309.8) * 1024,
    "cg_mg_dir": (2.0 * 1229.6 + 1637.5) * 1024,
    "cg_mg_down4": (2.0 * 66.2 + 20.8) * 1024,
    "particle_scatter": (2.0 * 8065.4 + 18469.3) * 1024,
    "particle_gather": (2.0 * 10294.5 + 16676.3) * 1024,
    "face_update": (2.0 * 12821.0 + 28672.2) * 1024,
}


# timed steps per autotuner candidate on a single domain
TUNE_STEPS = 8

# a kernel moving less than this per launch is bound by launch and barrier latency
ROOFLINE_MIN_BYTES = 1e6

# timers that wrap more than one kernel (or one kernel at several levels)
MULTI_KERNEL_TIMERS = {"cg_mg_setup", "cg_mg_finish", "cg_mg_down7+", "cg_mg_up7+", "bucket_order"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# multigrid iterations replayed from captured HIP graphs (--graph); the multigrid
# smoother (--mg-smoother: 0 damped Jacobi, 1 two-step Chebyshev)
GRAPH = 1
SMOOTHER = 1
# the multigrid settings fsg_autotune chooses
TUNED = ("mg_level_sweeps", "mg_level0_sweeps", "mg_coarse_sweeps", "mg_fuse_update", "graph_iters", "splat_order",
         "mg_tile_threads")


def make_domain(fsg, scene, solver, tol, max_iter, warm=0):
    return fsg.Domain.from_scene(scene, cg_tol=tol, cg_max_iter=max_iter, solver=solver,
                                 refine_inner_tol=1e-3, refine_max_outer=20, check_every=64,
                                 warm_start=warm, graph=GRAPH, mg_smoother=SMOOTHER)


def run_steps(fsg, d, n, dt, scheme=None):
    scheme = fsg.PICFLIP if scheme is None else scheme
    for _ in range(n):
        d.step(scheme, dt)


def solver_gbs(prof, nx, ny, P, levels=None, solves=None, live_frac=1.0):
    """The pressure solver's algorithmic bytes over its kernel time (GB/s): live-tile
    counts (`solves`: (iterations, live tiles) per solve of the profiled steps) for one
    domain; for a batch (`levels`) the level-0 and vector kernels' dense counts scaled
    by the fraction of level-0 tiles that hold an unknown (`live_frac`: the others are
    skipped), the smaller levels' dense."""
    tot_b = tot_ms = 0.0
    for k, (ms, n) in prof.items():
        if not k.startswith("cg_") or k in MULTI_KERNEL_TIMERS:
            continue
        if solves is not None:
            b = solve_bytes(k, nx, ny, solves)
            if b is not None:
                tot_b += b
                tot_ms += ms * n * len(solves)
        else:
            b = kernel_bytes(k, nx, ny, P, levels)
            if b is not None:
                if k in ("cg_mg_down0", "cg_mg_up0", "cg_mg_xr", "cg_mg_dir", "cg_mg_rz"):
                    b *= live_frac
                tot_b += b * n
                tot_ms += ms * n
    return tot_b / (tot_ms * 1e-3) / 1e9 if tot_ms > 0 else None


def batch_levels(sizes, W, H, coarse_cells=4096):
    """A batch's packed multigrid levels: halved up to the first level `la` at which
    every scene's rectangle has at most `coarse_cells` cells (>= 1)."""
    la = 1
    while any(((a + (1 << la) - 1) >> la) * ((b + (1 << la) - 1) >> la) > coarse_cells
              for a, b in sizes):
        la += 1
    lv = [(W, H)]
    for _ in range(la):
        lv.append(((lv[-1][0] + 1) // 2, (lv[-1][1] + 1) // 2))
    return [a * b for a, b in lv], la


def kernel_profile(fsg, d, steps, dt, scheme=None, solves=None):
    """Per-kernel HIP-event times over `steps` steps (continuing the trajectory):
    {name: (ms per launch, launches per step)}; `solves` receives each step's
    (iterations, live level-0 tiles)."""
    d.synchronize()
    d.reset_timing()
    d.set_timing(1)
    scheme = fsg.PICFLIP if scheme is None else scheme
    for _ in range(steps):
        d.step(scheme, dt)
        if solves is not None:
            st = d.stats()
            solves.append((st.cg_iterations, st.mg_live_tiles))
    d.synchronize()
    tm = d.timing()
    d.set_timing(0)
    return {k: (ms / max(n, 1), n / steps) for k, (ms, n) in tm.items()}


def timed_segment(fsg, d, steps, dt, kernel=None, scheme=None):
    """K steps bracketed by synchronisations; with `kernel`, that kernel (and only
    it) is timed with HIP events on the solver's stream inside the region, and each
    step's solver iterations and live level-0 tiles are recorded (host-side
    counters, no sync)."""
    d.synchronize()
    d.reset_stats()
    d.reset_timing()
    if kernel:
        d.set_timing_prefix(kernel)
        d.set_timing(2)
    clock = {"cg_mg_coarse": "coarse", "cg_mg_down0": "down0"}.get(kernel)
    if clock:  # the kernel's own device clock as well (no launch gaps)
        d.set_kernel_clock(True, clock)
    scheme = fsg.PICFLIP if scheme is None else scheme
    solves = []
    d.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        d.step(scheme, dt)
        if kernel:
            st = d.stats()
            solves.append((st.cg_iterations, st.mg_live_tiles))
    d.synchronize()
    t1 = time.perf_counter()
    st = d.stats()
    tm = d.timing() if kernel else {}
    if clock:
        us, n = d.kernel_clock()
        d.set_kernel_clock(False)
        if n > 0:
            tm["__clock__"] = (us * 1e-3, n)
    if kernel:
        tm["__solves__"] = solves
    d.set_timing(0)
    return t1 - t0, st, tm


def full_run(fsg, scene, solver, tol, steps, autotune=False, tune_every=100):
    """The configuration's whole run (`steps` steps from the initial state) on a
    domain of its own, autotuned (untimed) at its initial state and every `tune_every`
    steps if asked (the flow changes what is fastest: the particles leave their
    initial lattice order within ~40 steps, which decides the splat order): wall time
    of the steps, and every solve's iterations and FP64 relative residual (host-side
    counters read after each step, no extra synchronisation)."""
    d = make_domain(fsg, scene, solver, tol, 200 if solver == fsg.SOLVER_MG else 20000)
    its, errs, tunes = [], [], {}
    seconds = tune_s = 0.0
    seg = tune_every if autotune else max(steps, 1)
    for s0 in range(0, steps, seg):
        if autotune:
            t0 = time.perf_counter()
            tunes[str(s0)] = d.autotune(fsg.PICFLIP, scene.dt, steps=TUNE_STEPS)
            tune_s += time.perf_counter() - t0
        d.synchronize()
        t0 = time.perf_counter()
        for _ in range(min(seg, steps - s0)):
            d.step(fsg.PICFLIP, scene.dt)
            st = d.stats()
            its.append(st.cg_iterations)
            errs.append(st.cg_error)
        d.synchronize()
        seconds += time.perf_counter() - t0
    d.close()
    tune = ({k: {q: t[q] for q in TUNED} for k, t in tunes.items()} | {"seconds": tune_s}
            if autotune else None)
    return {"steps": steps, "seconds": seconds, "steps_per_s": steps / seconds,
            "cg_iterations_min": int(min(its)), "cg_iterations_mean": float(np.mean(its)),
            "cg_iterations_max": int(max(its)), "cg_rel_residual_max": float(max(errs)),
            "all_solves_within_tol": bool(max(errs) <= tol),
            "cg_iterations_by_step": {str(k): int(its[k]) for k in range(0, steps, 50)},
            "autotune": tune}


def fp64_single_step_error(fsg, scene, solver, tol, state=None):
    """Error of one chosen-form step against the FP64 oracle from the same input:
    the scene's initial state, or a state saved from the GPU trajectory."""
    from oracle import OracleSim
    from oracle.binding import SCHEME_PICFLIP
    d = make_domain(fsg, scene, solver, tol, 20000)
    o = OracleSim.from_scene(scene, "f64")
    if state is not None:
        d.load_state(state)
        o.set_particles(state["particles"])
        for name in ("front", "back"):
            o.set_vel(*state[name], name)
        o.u_prev[...], o.v_prev[...] = state["prev"]
        o.u_diff[...], o.v_diff[...] = state["diff"]
        o.label[...] = state["labels"]
    d.step(fsg.PICFLIP, scene.dt)
    o.set_params(cg_tol=1e-13, cg_max_iter=200000)
    o.step(SCHEME_PICFLIP, scene.dt)
    pg, po = d.particles(), o.particles
    ug, vg = d.velocity(fsg.VEL_FRONT)
    uo, vo = o.vel("front")

    def rel(a, b):
        return float(np.abs(a - b).max() / max(float(np.abs(b).max()), 1e-30))

    same_labels = bool(np.array_equal(d.labels(), o.label))
    errs = {"x": rel(pg[:, 0], po[:, 0]), "y": rel(pg[:, 1], po[:, 1]),
            "vx": rel(pg[:, 2], po[:, 2]), "vy": rel(pg[:, 3], po[:, 3]),
            "u": rel(ug, uo), "v": rel(vg, vo)}
    d.close()
    return errs, same_labels


def cpu_baseline(scene, steps):
    from oracle import OracleSim
    from oracle.binding import SCHEME_PICFLIP
    o = OracleSim.from_scene(scene, "f32")
    o.set_params(cg_tol=1e-6, cg_max_iter=10000)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(SCHEME_PICFLIP, scene.dt)
    dt = time.perf_counter() - t0
    return steps / dt, o.s.cg_iterations


SOLVERS = {"plain": 0, "persistent": 1, "refined": 2, "mg": 3}


def batch_leg(fsg, scenes, args):
    """The packed batch of scenes: ``--batch-scenes`` non-square dam breaks of sides
    ``--batch-lo..--batch-hi`` with 1-4 obstacles each, stepped together (one launch
    per phase, one multigrid-preconditioned solve bounding every scene's relative
    residual).  For comparison, the same scenes stepped one Domain at a time over a
    sample of ``--batch-seq-scenes`` scenes."""
    scs = scenes.dam_break_batch(args.batch_scenes, args.batch_lo, args.batch_hi)
    cells = sum(sc.nx * sc.ny for sc in scs)
    parts = sum(len(sc.particles) for sc in scs)
    b = fsg.Batch.from_scenes(scs, cg_tol=args.tol, cg_max_iter=200, graph=GRAPH,
                              mg_smoother=SMOOTHER)
    log(f"batch: {len(scs)} scenes, {cells} cells, {parts} particles, packed {b.nx}x{b.ny}")
    run_steps(fsg, b, args.batch_warmup, scs[0].dt)
    tune = None
    if args.autotune:
        t0 = time.perf_counter()
        tune = b.autotune(fsg.PICFLIP, scs[0].dt, steps=1)
        tune["seconds"] = time.perf_counter() - t0
        log(f"batch autotune: {tune}")
    prof = kernel_profile(fsg, b, 2, scs[0].dt)
    per_step = {k: ms * n for k, (ms, n) in prof.items()}
    st_p = b.stats()
    unknowns, live = st_p.n_fluid, st_p.mg_live_tiles
    elapsed, st, _ = timed_segment(fsg, b, args.batch_steps, scs[0].dt)
    W, H = b.nx, b.ny
    b.close()
    sps = args.batch_steps / elapsed
    live_frac = live / (((W + 31) // 32) * ((H + 31) // 32))
    gbs = solver_gbs(prof, W, H, parts, batch_levels([(sc.nx, sc.ny) for sc in scs], W, H),
                     live_frac=live_frac)
    log(f"batch: {sps:.3f} steps/s = {sps * len(scs):.1f} scene-steps/s "
        f"({st.cg_iterations} MG iterations in the last solve)")
    seq = None
    k = min(args.batch_seq_scenes, len(scs))
    if k > 0:
        doms = [make_domain(fsg, sc, fsg.SOLVER_MG, args.tol, 200) for sc in scs[:k]]
        for d in doms:
            run_steps(fsg, d, args.batch_warmup, scs[0].dt)
        for d in doms:
            d.synchronize()
        nsteps = 5
        t0 = time.perf_counter()
        for _ in range(nsteps):
            for d in doms:
                d.step(fsg.PICFLIP, scs[0].dt)
        for d in doms:
            d.synchronize()
        t1 = time.perf_counter()
        cells_k = sum(sc.nx * sc.ny for sc in scs[:k])
        # cell-steps/s of the one-at-a-time form, converted to the batch's scene-steps
        seq = {"scenes": k, "steps": nsteps,
               "scene_steps_per_s": k * nsteps / (t1 - t0),
               "cell_steps_per_s": cells_k * nsteps / (t1 - t0)}
        for d in doms:
            d.close()
        log(f"one scene at a time ({k} scenes): {seq['scene_steps_per_s']:.1f} scene-steps/s")
    cpu = None
    k = min(args.batch_cpu_scenes, len(scs))
    if k > 0 and not args.no_cpu:
        from oracle import OracleSim
        from oracle.binding import SCHEME_PICFLIP
        t_cpu = 0.0
        for sc in scs[:k]:
            o = OracleSim.from_scene(sc, "f32")
            o.set_params(cg_tol=1e-6, cg_max_iter=10000)
            t0 = time.perf_counter()
            o.step(SCHEME_PICFLIP, sc.dt)
            t_cpu += time.perf_counter() - t0
        cells_k = sum(sc.nx * sc.ny for sc in scs[:k])
        cpu = {"value": k / t_cpu, "unit": "scene-steps/s", "cores": 1, "kind": "port",
               "cell_steps_per_s": cells_k / t_cpu,
               "batch_steps_per_s_extrapolated": cells_k / t_cpu / cells,
               "sample": f"first step of the first {k} scenes, FP32 oracle, Jacobi-PCG rel "
                         f"tol 1e-6, 1 thread; the batch rate is extrapolated by cells"}
        log(f"batch cpu oracle: {k / t_cpu:.3f} scene-steps/s over {k} scenes ({t_cpu:.1f} s)")
    # the per-scene CG vector kernels: algorithmic bytes per unknown (each array read or
    # written once; the cells of their tiles that are not unknowns move bytes that are
    # not counted) per launch, over their HIP-event time (which also counts launch
    # gaps, so this is a lower bound on the bandwidth)
    vec_bytes = {"cg_mg_xr": 33.0 * unknowns, "cg_mg_dir": 13.0 * unknowns,
                 "cg_mg_rz": 9.0 * unknowns}
    vec_gbs = {k: vec_bytes[k] / (prof[k][0] * 1e-3) / 1e9 for k in vec_bytes if k in prof}
    return {
        "workload": f"dam_break_batch: {len(scs)} non-square scenes, sides {args.batch_lo}.."
                    f"{args.batch_hi}, 1-4 obstacles each, {cells} cells, {parts} particles, "
                    f"stepPICFLIP dt 0.01, 0.95 FLIP, per-scene rel tol {args.tol:g}",
        "steps": args.batch_steps,
        "autotune": tune,
        "steps_per_s": sps,
        "ms_per_step": 1e3 / sps,
        "scene_steps_per_s": sps * len(scs),
        "cell_steps_per_s": sps * cells,
        "cg_iterations_last_solve": st.cg_iterations,
        "cg_hbm_gbs": gbs,
        "cg_unknowns": unknowns,
        "level0_live_tile_fraction": live_frac,
        "cg_vector_kernel_gbs_per_unknown_bytes": vec_gbs,
        "one_scene_at_a_time": seq,
        "cpu_baseline": cpu,
        "speedup_vs_one_at_a_time": (sps * cells / seq["cell_steps_per_s"]) if seq else None,
        "kernel_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
    }


def grid_leg(fsg, scenes, args):
    """Grid-only semi-Lagrangian advection + projection on the fully fluid
    ``--grid-n``^2 grid (divergence-free 64-mode Fourier velocity, RK2 gather
    backtrace, dt = 0.5 / max|u|, mean-removed right-hand side, closed walls: the
    reference's one-sided wall rule makes this closed box inject fluid through its
    walls and overflow within a few steps)."""
    sc = scenes.grid_only(args.grid_n)
    d = fsg.Domain.from_scene(sc, cg_tol=args.tol, cg_max_iter=200, solver=fsg.SOLVER_MG, graph=GRAPH,
                              integrator=2, project_mean=1, closed_walls=1, mg_smoother=SMOOTHER)
    run_steps(fsg, d, 2, sc.dt, fsg.GRID_ADVECT)
    tune = d.autotune(fsg.GRID_ADVECT, sc.dt, steps=2) if args.autotune else None
    if tune:
        log(f"grid-only autotune: {tune}")
    prof = kernel_profile(fsg, d, 2, sc.dt, fsg.GRID_ADVECT)
    per_step = {k: ms * n for k, (ms, n) in prof.items()}
    elapsed, st, _ = timed_segment(fsg, d, args.grid_steps, sc.dt, scheme=fsg.GRID_ADVECT)
    d.close()
    sps = args.grid_steps / elapsed
    log(f"grid-only {args.grid_n}: {sps:.2f} steps/s ({st.cg_iterations} MG iterations)")
    return {
        "workload": f"grid_only_{args.grid_n}: gather-RK2 semi-Lagrangian advection + "
                    f"projection, fully fluid {args.grid_n}^2, closed walls, "
                    f"dt {sc.dt:.4g}, rel tol {args.tol:g}",
        "steps": args.grid_steps,
        "autotune": tune,
        "steps_per_s": sps,
        "ms_per_step": 1e3 / sps,
        "cg_iterations_last_solve": st.cg_iterations,
        "cg_hbm_gbs": solver_gbs(prof, args.grid_n, args.grid_n, 0),
        "kernel_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
    }


def replay_gradient_check(fsg, scenes, n, steps, every, n_dirs, eps_list, scale, tol, workers):
    """The differentiable configuration at n^2 (closed all-fluid box, (2, 16, 16)
    controls x scale, ``steps`` PICFLIP steps, checkpoints every ``every``): the
    adjoint gradient (deterministic rollout) along ``n_dirs`` seeded unit directions
    against central FP64 differences (step ``eps``) of the oracle replaying the GPU
    run's decisions (labels, P2G weight tests, Dirichlet passes, rejections), which
    follows the branch of the piecewise-smooth rollout the GPU run took.  Runs after
    the timed legs; the oracle is the checker here, not the thing measured."""
    import torch
    from fluid_simulation_amd import adjoint
    from oracle.replay import ReplayRollout
    t0 = time.perf_counter()
    dt = 0.01
    jj, ii = np.meshgrid(np.arange(1, n - 1), np.arange(1, n - 1), indexing="ij")
    parts = scenes.jittered_particles(ii.ravel(), jj.ravel(), n, 1.0, 1.0, 0)
    model = adjoint.ControlModel(n, n, float(n), float(n), parts[:, :2])
    d = fsg.Domain(n, n, float(n), float(n), dt, 0.05)
    d.set_params(solver=fsg.SOLVER_MG, cg_tol=tol, cg_max_iter=500, project_mean=1,
                 closed_walls=1, deterministic=1, record_decisions=1, mg_smoother=SMOOTHER)

    def start(c):
        vel, u, v = model(c)
        adjoint.set_state(d, np.concatenate([parts[:, :2], vel.detach().numpy()], 1),
                          u.detach().numpy(), v.detach().numpy())

    start(torch.randn(2, 16, 16, generator=torch.Generator().manual_seed(1)) * scale)
    for _ in range(steps):
        d.step(fsg.PICFLIP, dt)
    target = d.density().astype(np.float64)
    c0 = torch.randn(2, 16, 16, generator=torch.Generator().manual_seed(0)) * scale
    cg = c0.clone().requires_grad_(True)
    vel, u, v = model(cg)
    loss = adjoint.Rollout.apply(torch.tensor(parts[:, :2]), vel, u, v, d, fsg.PICFLIP, dt,
                                 steps, every, target)
    loss.backward()
    grad = cg.grad.numpy().astype(np.float64)
    start(c0)
    recs = adjoint.rollout_decisions(d, fsg.PICFLIP, dt, steps)
    d.close()
    rr = ReplayRollout(n, n, float(n), float(n), dt, 0.05, recs, cg_tol=1e-13, closed_walls=1,
                       project_mean=1)

    def state(c):
        pv, uu, vv = model(torch.as_tensor(c), dtype=torch.float64)
        return (np.concatenate([parts[:, :2].astype(np.float64), pv.numpy()], 1), uu.numpy(),
                vv.numpy())

    base = c0.numpy().astype(np.float64)
    rng = np.random.default_rng(16)
    dirs = []
    for _ in range(n_dirs):
        dv = rng.standard_normal((2, 16, 16))
        dirs.append(dv / np.linalg.norm(dv))
    l0 = rr.density_losses([state(base)], 3, dt, target, workers=1)[0]
    gnorm = float(np.linalg.norm(grad))
    out = {
        "workload": f"closed all-fluid {n}^2 box, {len(parts)} particles, {steps} PICFLIP "
                    f"steps, checkpoints every {every}, deterministic rollout",
        "directions": n_dirs,
        "loss_gpu_fp32": float(loss.detach()), "loss_fp64_replay": l0,
        "loss_rel_diff": abs(l0 - float(loss.detach())) / abs(l0),
        "grad_norm": gnorm, "oracle_threads": workers, "by_eps": {},
    }
    for eps in eps_list:
        states = []
        for dv in dirs:
            states += [state(base + eps * dv), state(base - eps * dv)]
        ls = rr.density_losses(states, 3, dt, target, workers=workers)
        rows = []
        for k, dv in enumerate(dirs):
            fd = (ls[2 * k] - ls[2 * k + 1]) / (2 * eps)
            ad = float((grad * dv).sum())
            rows.append((ad, fd, abs(ad - fd) / max(abs(fd), abs(ad), 1e-30),
                         abs(ad - fd) / gnorm))
        r = {"rel_median": float(np.median([q[2] for q in rows])),
             "rel_max": float(max(q[2] for q in rows)),
             "rel_to_grad_norm_max": float(max(q[3] for q in rows)),
             "directions": [{"adjoint": q[0], "fd": q[1]} for q in rows]}
        out["by_eps"][f"{eps:g}"] = r
        log(f"adjoint vs FP64 replayed differences ({n}^2, {steps} steps, {n_dirs} directions, "
            f"eps {eps:g}): median rel {r['rel_median']:.2e}, max {r['rel_max']:.2e}, max rel "
            f"to |grad| {r['rel_to_grad_norm_max']:.2e}")
    out["seconds"] = time.perf_counter() - t0
    log(f"replayed check: loss FP32 vs FP64 replay {out['loss_rel_diff']:.1e} "
        f"({out['seconds']:.1f} s)")
    return out


def adjoint_leg(fsg, scenes, args):
    """The differentiable form: a closed all-fluid ``--adj-n``^2 box (4 particles per
    interior cell), initial velocities from a (2, 16, 16) control grid bilinearly
    upsampled, ``--adj-steps`` PICFLIP steps with checkpoints every ``--adj-every``,
    loss = sum((particle density - target)^2) with the target from control seed 1 and
    the evaluation at seed 0, deterministic (fixed-point) splats.  Times the
    checkpointed forward and the reverse pass, and compares the gradient with central
    FP32 differences of the GPU rollout (each ``--adj-eps``) along ``--adj-dirs``
    seeded unit directions, counting the label changes between each pair of runs (a
    discrete decision the differences cross), beside the loss's run-to-run spread at
    fixed controls (zero for the deterministic rollout).  The well-posed check, against
    FP64 differences replaying the run's decisions, runs at ``--adj-check-n``^2."""
    import torch
    from fluid_simulation_amd import adjoint
    n, steps, every, dt = args.adj_n, args.adj_steps, args.adj_every, 0.01
    jj, ii = np.meshgrid(np.arange(1, n - 1), np.arange(1, n - 1), indexing="ij")
    parts = scenes.jittered_particles(ii.ravel(), jj.ravel(), n, 1.0, 1.0, 0)
    model = adjoint.ControlModel(n, n, float(n), float(n), parts[:, :2])
    d = fsg.Domain(n, n, float(n), float(n), dt, 0.05)
    d.set_params(solver=fsg.SOLVER_MG, cg_tol=args.tol, cg_max_iter=200, project_mean=1,
                 closed_walls=1, deterministic=1, mg_smoother=SMOOTHER)
    pos = torch.tensor(parts[:, :2].astype(np.float32))

    def run(c, record=False):
        vel, u, v = model(c)
        adjoint.set_state(d, np.concatenate([parts[:, :2], vel.detach().numpy()], 1),
                          u.detach().numpy(), v.detach().numpy())
        labels = []
        for _ in range(steps):
            d.step(fsg.PICFLIP, dt)
            if record:
                labels.append(d.labels())
        return d.density().astype(np.float64), labels

    scale = args.adj_scale
    target, _ = run(torch.randn(2, 16, 16, generator=torch.Generator().manual_seed(1)) * scale)
    c = (torch.randn(2, 16, 16, generator=torch.Generator().manual_seed(0)) * scale)
    # untimed: the control model's first forward / backward (torch's one-time set-up)
    cw = torch.zeros(2, 16, 16, requires_grad=True)
    torch.autograd.backward(list(model(cw)), [torch.ones(len(parts), 2), torch.ones(n, n),
                                              torch.ones(n, n)])
    c.requires_grad_(True)
    d.synchronize()
    tm0 = time.perf_counter()
    vel, u, v = model(c)
    t0 = time.perf_counter()
    loss = adjoint.Rollout.apply(pos, vel, u, v, d, fsg.PICFLIP, dt, steps, every, target)
    d.synchronize()
    t1 = time.perf_counter()
    # the reverse pass on the GPU (Rollout.backward: checkpointed recompute, adjoint
    # steps, the adjoint state copied out), then the control model's backward (torch,
    # host CPU), timed apart
    g_in = torch.autograd.grad(loss, [vel, u, v])
    d.synchronize()
    t2 = time.perf_counter()
    torch.autograd.backward([vel, u, v], list(g_in))
    t3 = time.perf_counter()
    grad = c.grad.detach().numpy().astype(np.float64)
    log(f"adjoint {n}^2, {len(parts)} particles, {steps} steps: control model "
        f"{1e3 * (t0 - tm0):.1f} ms, forward {1e3 * (t1 - t0):.1f} ms, "
        f"reverse {1e3 * (t2 - t1):.1f} ms, control model backward {1e3 * (t3 - t2):.1f} ms, "
        f"loss {loss.item():.6e}, |grad| {np.linalg.norm(grad):.3e}")
    # run-to-run spread of the loss at the same controls (zero: fixed-point splats)
    with torch.no_grad():
        l_a = float((((run(c)[0]) - target) ** 2).sum())
        l_b = float((((run(c)[0]) - target) ** 2).sum())
    noise = abs(l_a - l_b)
    fd_out = {}
    for eps in (args.adj_eps if args.adj_dirs > 0 else []):
        rng = np.random.default_rng(16)
        rows = []
        for k in range(args.adj_dirs):
            dirn = rng.standard_normal((2, 16, 16))
            dirn /= np.linalg.norm(dirn)
            with torch.no_grad():
                rp, lp = run(c + eps * torch.tensor(dirn, dtype=torch.float32), record=True)
                rm, lm = run(c - eps * torch.tensor(dirn, dtype=torch.float32), record=True)
            flips = int(sum((a != b).sum() for a, b in zip(lp, lm)))
            fd = (((rp - target) ** 2).sum() - ((rm - target) ** 2).sum()) / (2 * eps)
            adj = float((grad * dirn).sum())
            rel = abs(adj - fd) / max(abs(fd), abs(adj), 1e-30)
            rows.append((adj, fd, rel, flips))
        rels = [r[2] for r in rows]
        clean = [r[2] for r in rows if r[3] == 0]
        log(f"adjoint vs FP32 differences, eps {eps:g}, {len(rows)} directions: median rel "
            f"{np.median(rels):.2e}, max {max(rels):.2e}; {len(clean)} without label changes; "
            f"loss noise {noise:.2e} vs |dL| ~ {2 * eps * np.median([abs(r[0]) for r in rows]):.2e}")
        fd_out[f"{eps:g}"] = {
            "directions": [{"adjoint": a, "fd": f, "rel": r, "label_changes": fl}
                           for a, f, r, fl in rows],
            "rel_median": float(np.median(rels)),
            "rel_max": float(max(rels)),
            "directions_without_label_changes": len(clean),
            "rel_max_without_label_changes": float(max(clean)) if clean else None,
        }
    d.close()
    check = None
    if args.adj_check_n > 0 and args.adj_check_dirs > 0:
        check = replay_gradient_check(fsg, scenes, args.adj_check_n, steps, every,
                                      args.adj_check_dirs, (1e-5, 1e-3), scale, 1e-10,
                                      args.adj_check_threads)
    return {
        "fp64_replay_check": check,
        "workload": f"closed all-fluid {n}^2 box, {len(parts)} particles, controls (2, 16, 16) "
                    f"x {scale:g} bilinearly upsampled, {steps} PICFLIP steps dt {dt}, "
                    f"checkpoints every {every}, loss sum((density - target)^2)",
        "control_model_ms": 1e3 * (t0 - tm0),
        "forward_ms": 1e3 * (t1 - t0),
        "forward_ms_per_step": 1e3 * (t1 - t0) / steps,
        "reverse_ms": 1e3 * (t2 - t1),
        "control_model_backward_ms": 1e3 * (t3 - t2),
        "reverse_ms_per_step": 1e3 * (t2 - t1) / steps,
        "loss": loss.item(),
        "grad_norm": float(np.linalg.norm(grad)),
        "loss_run_to_run_spread": noise,
        "fd": fd_out,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--solver", default="mg", choices=list(SOLVERS))
    ap.add_argument("--tol", type=float, default=1e-6,
                    help="pressure solve: relative residual (FP64 residual for mg / refined)")
    ap.add_argument("--start", type=int, default=250,
                    help="first timed step: the trajectory is advanced untimed to it (the "
                         "configuration's 500-step run is in its developed flow there)")
    ap.add_argument("--run-steps", type=int, default=500,
                    help="the configuration's whole run, timed on its own domain")
    ap.add_argument("--refined-steps", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--plain-tol", type=float, default=1e-6)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--roofline-kernel", default=None,
                    help="timer name of the roofline kernel (default: the one with the most "
                         "time per step in the profile pass)")
    ap.add_argument("--graph", type=int, default=1,
                    help="multigrid iterations replayed from captured HIP graphs (1) or "
                         "launched kernel by kernel (0)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: fsg_autotune chooses the multigrid settings (untimed) before the "
                         "timed region of each workload")
    ap.add_argument("--mg-smoother", type=int, default=1,
                    help="multigrid smoothing: 0 damped Jacobi (0.8), 1 two-step Chebyshev")
    ap.add_argument("--warm-start", type=int, default=0,
                    help="chosen form: 1 starts each solve from the previous pressure")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-lo", type=int, default=64)
    ap.add_argument("--batch-hi", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--batch-warmup", type=int, default=3)
    ap.add_argument("--batch-seq-scenes", type=int, default=32)
    ap.add_argument("--batch-cpu-scenes", type=int, default=16)
    ap.add_argument("--grid-n", type=int, default=2048)
    ap.add_argument("--grid-steps", type=int, default=20)
    ap.add_argument("--adj-n", type=int, default=512)
    ap.add_argument("--adj-steps", type=int, default=64)
    ap.add_argument("--adj-every", type=int, default=8)
    ap.add_argument("--adj-dirs", type=int, default=16)
    ap.add_argument("--adj-scale", type=float, default=4.0)
    ap.add_argument("--adj-check-n", type=int, default=128,
                    help="grid of the FP64 replayed gradient check (0: off)")
    ap.add_argument("--adj-check-dirs", type=int, default=8)
    ap.add_argument("--adj-check-threads", type=int, default=8)
    ap.add_argument("--adj-eps", type=lambda t: [float(x) for x in t.split(",")],
                    default=[1e-3, 3e-2], help="comma-separated FD steps in control space")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 5 times (calibrates the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    global GRAPH, SMOOTHER
    GRAPH = args.graph
    SMOOTHER = args.mg_smoother
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from fluid_simulation_amd import fsg, scenes
    fsg.load()
    scene = scenes.dam_break(args.n)
    dt = scene.dt
    P = len(scene.particles)
    solver = SOLVERS[args.solver]
    max_iter = 200 if solver == fsg.SOLVER_MG else 20000
    log(f"workload dam_break_{args.n}: {P} particles; solver {args.solver} tol {args.tol:g}")

    # ---- chosen form: W warmup steps, the trajectory advanced untimed to step
    # --start (the developed flow), the autotuner (untimed; it restores the state), a
    # per-kernel profile, then the timed region
    d = make_domain(fsg, scene, solver, args.tol, max_iter, args.warm_start)
    run_steps(fsg, d, args.warmup, dt)
    advance = max(args.start - args.warmup - args.profile_steps, 0)
    run_steps(fsg, d, advance, dt)
    tune = None
    if args.autotune and solver == fsg.SOLVER_MG:
        t0 = time.perf_counter()
        # 8 timed steps per candidate: a weaker preconditioner (the truncated coarse solve)
        # can win over 3 steps where the flow happens to need no more iterations, and
        # lose them over the window that follows
        tune = d.autotune(fsg.PICFLIP, dt, steps=TUNE_STEPS)
        tune["seconds"] = time.perf_counter() - t0
        log(f"autotune at step {args.warmup + advance}: {tune}")
    prof_solves = []
    prof = (kernel_profile(fsg, d, args.profile_steps, dt, solves=prof_solves)
            if args.profile_steps > 0 else {})
    t_start = args.warmup + advance + args.profile_steps
    state0 = d.save_state() if (args.plain_steps > 0 or args.refined_steps > 0) else None
    per_step = {k: ms * n for k, (ms, n) in prof.items()}
    # the roofline kernel: the one kernel with the most time per step among those that
    # move at least ROOFLINE_MIN_BYTES of algorithmic bytes per launch (timers that
    # cover several kernels are not candidates); below that a kernel is bound by launch
    # and barrier latency, not by HBM: those are reported beside it ("latency_bound")
    single = {k: v for k, v in per_step.items() if k not in MULTI_KERNEL_TIMERS}
    one_domain_mg = solver == fsg.SOLVER_MG and mg_levels(args.n, args.n)[1] >= 2

    def launch_bytes(k):
        if one_domain_mg and k.startswith("cg_") and prof_solves:
            b = solve_bytes(k, args.n, args.n, prof_solves)
            n = prof[k][1] * len(prof_solves)
            return b / n if b is not None and n else None
        return kernel_bytes(k, args.n, args.n, P)

    hbm = {k: v for k, v in single.items() if (launch_bytes(k) or 0.0) >= ROOFLINE_MIN_BYTES}
    latency = sorted(((k, v) for k, v in single.items() if k not in hbm), key=lambda kv: -kv[1])
    pool = hbm or single
    dominant = args.roofline_kernel or (max(pool, key=pool.get) if pool else None)
    log("kernel profile (ms per step): " +
        ", ".join(f"{k} {v:.3f}" for k, v in sorted(per_step.items(), key=lambda kv: -kv[1])[:12]))
    # the timed region: the multigrid iterations replayed from their graphs, nothing
    # instrumented
    elapsed, st, _ = timed_segment(fsg, d, args.steps, dt)
    state = d.save_state() if not args.no_fp64_check else None
    sps = args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"chosen: {sps:.3f} steps/s ({ms_per_step:.3f} ms/step) over steps {t_start}.."
        f"{t_start + args.steps - 1}, dominant kernel {dominant}")
    # the roofline kernel timed live over the next K steps of the same trajectory: HIP
    # events around each of its launches on the solver's stream (per-kernel events
    # launch every kernel, not the captured graphs: this segment's own rate is
    # reported beside, not as the value)
    elapsed_k, _, tm = timed_segment(fsg, d, args.steps, dt, kernel=dominant)
    d.close()
    log(f"instrumented segment (eager launches, {dominant} timed): "
        f"{args.steps / elapsed_k:.3f} steps/s")

    # ---- the configuration's whole run (500 steps from the initial state), on a
    # domain of its own, tuned at its own start: average rate, iterations and the
    # largest FP64 residual
    full = (full_run(fsg, scene, solver, args.tol, args.run_steps,
                     bool(args.autotune) and solver == fsg.SOLVER_MG)
            if args.run_steps > 0 else None)
    if full:
        log(f"whole run: {full['steps']} steps at {full['steps_per_s']:.1f} steps/s, "
            f"{full['cg_iterations_min']}-{full['cg_iterations_max']} iterations "
            f"(mean {full['cg_iterations_mean']:.1f}), max rel residual "
            f"{full['cg_rel_residual_max']:.2e}")

    # roofline of the dominant kernel, timed live over the instrumented segment: HIP
    # events around each of its launches on the solver's stream (they also count the
    # dispatch before the kernel starts), and the kernel's own device clock beside
    solves = tm.get("__solves__", [])
    roof = None
    if dominant and dominant in tm:
        ev_ms_total, n_launch = tm[dominant]
        ev_ms = ev_ms_total / max(n_launch, 1)
        if one_domain_mg and dominant.startswith("cg_"):
            total_b = solve_bytes(dominant, args.n, args.n, solves)
        else:
            kb = kernel_bytes(dominant, args.n, args.n, P)
            total_b = kb * n_launch if kb is not None else None
        k_bytes = total_b / n_launch if total_b is not None else None
        ach = k_bytes / (ev_ms * 1e-3) / 1e9 if k_bytes else None
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS if ach else None,
                "traffic": MEASURED_TRAFFIC_1024.get(dominant) if args.n == 1024 else None,
                "kernel": dominant, "timing": "HIP events",
                "kernel_us_per_launch": 1e3 * ev_ms, "launches": n_launch,
                "algorithmic_bytes_per_launch": k_bytes}
        if one_domain_mg and solves:
            roof["live_level0_tiles_mean"] = float(np.mean([lv for _, lv in solves]))
            roof["level0_tiles"] = ((args.n + 31) // 32) ** 2
            roof["cg_iterations_per_solve"] = [n for n, _ in solves]
        if "__clock__" in tm and k_bytes:
            ck_ms = tm["__clock__"][0] / tm["__clock__"][1]
            roof["kernel_us_per_launch_device_clock"] = 1e3 * ck_ms
            roof["achieved_device_clock"] = k_bytes / (ck_ms * 1e-3) / 1e9
            roof["frac_device_clock"] = roof["achieved_device_clock"] / HBM_PEAK_GBS
    # every solver kernel of the profile pass (HIP events per launch, live-tile bytes):
    # the latency-bound small levels beside the roofline kernel
    solver_kernels = {}
    if one_domain_mg and prof_solves:
        for k, (ms, n) in sorted(prof.items()):
            if not k.startswith("cg_") or k in MULTI_KERNEL_TIMERS:
                continue
            b = solve_bytes(k, args.n, args.n, prof_solves)
            launches = n * len(prof_solves)
            if b is None or launches == 0:
                continue
            gbs = b / launches / (ms * 1e-3) / 1e9
            solver_kernels[k] = {"us_per_launch_events": 1e3 * ms, "launches_per_step": n,
                                 "bytes_per_launch": b / launches, "gbs": gbs,
                                 "frac": gbs / HBM_PEAK_GBS}
    # "CG HBM GB/s": algorithmic bytes of every pressure-solver kernel over their time
    cg_gbs = solver_gbs(prof, args.n, args.n, P,
                        solves=prof_solves if one_domain_mg and prof_solves else None)

    # ---- the plain FP32 Jacobi-PCG form and the refined precision form (FP32 PCG
    # inside FP64 iterative refinement), each from the state at the timed window's
    # start (the same trajectory segment)
    def other_form(form, steps, tol):
        dp = make_domain(fsg, scene, form, tol, 20000)
        dp.load_state(state0)
        e_p, st_p, _ = timed_segment(fsg, dp, steps, dt)
        dp.close()
        return steps / e_p, st_p

    plain_sps = refined = None
    if args.plain_steps > 0 and solver != fsg.SOLVER_PLAIN:
        pk = min(args.plain_steps, args.steps)
        plain_sps, st_p = other_form(fsg.SOLVER_PLAIN, pk, args.plain_tol)
        log(f"plain: {plain_sps:.3f} steps/s over {pk} steps ({st_p.cg_iterations} iterations "
            f"in the last solve)")
    elif solver == fsg.SOLVER_PLAIN:
        plain_sps = sps
    if args.refined_steps > 0 and solver != fsg.SOLVER_REFINED:
        rk = min(args.refined_steps, args.steps)
        r_sps, st_r = other_form(fsg.SOLVER_REFINED, rk, args.tol)
        refined = {"steps_per_s": r_sps, "steps": rk, "tol": args.tol,
                   "inner_tol": 1e-3, "cg_iterations_last_solve": st_r.cg_iterations,
                   "refine_sweeps_last_solve": st_r.refine_sweeps,
                   "rel_residual_last_solve": st_r.cg_error,
                   "speedup_vs_plain": r_sps / plain_sps if plain_sps else None}
        log(f"refined: {r_sps:.3f} steps/s over {rk} steps ({st_r.cg_iterations} inner "
            f"iterations, {st_r.refine_sweeps} FP64 sweeps in the last solve)")

    # ---- single-step error against the FP64 oracle, from the state the timed run
    # reached (at step 0 the projected field is hydrostatic, ~0, and relative
    # errors of a near-zero field say little)
    errs, same_labels = (None, None)
    if not args.no_fp64_check:
        t0 = time.perf_counter()
        errs, same_labels = fp64_single_step_error(fsg, scene, solver, args.tol, state)
        log(f"fp64 check at step {t_start + args.steps} "
            f"({time.perf_counter() - t0:.1f} s): {errs} labels equal {same_labels}")
        if refined is not None:
            e_r, same_r = fp64_single_step_error(fsg, scene, fsg.SOLVER_REFINED, args.tol, state)
            refined["rel_err_vs_fp64_single_step"] = e_r
            refined["max_rel_err_vs_fp64"] = max(e_r.values())
            refined["labels_equal_fp64"] = same_r
            log(f"refined fp64 check: max {max(e_r.values()):.2e}, labels equal {same_r}")

    # ---- packed batch of scenes; grid-only advection + projection
    batch = batch_leg(fsg, scenes, args) if args.batch_scenes > 0 else None
    grid = grid_leg(fsg, scenes, args) if args.grid_n > 0 and args.grid_steps > 0 else None
    adj = adjoint_leg(fsg, scenes, args) if args.adj_n > 0 and args.adj_steps > 0 else None

    # ---- CPU baseline: the FP32 oracle, one thread, bounded sample
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        t0 = time.perf_counter()
        cpu_sps, cpu_its = cpu_baseline(scene, args.cpu_steps)
        log(f"cpu oracle: {cpu_sps:.4f} steps/s ({time.perf_counter() - t0:.1f} s, "
            f"{cpu_its} CG iterations in the last step)")
        cpu = {"value": cpu_sps, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"first {args.cpu_steps} steps of the same {args.n}^2 dam break, FP32 "
                         f"oracle (C restatement, gcc -O2), Jacobi-PCG rel tol 1e-6, 1 thread"}

    out = {
        "metric": METRIC,
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"dam_break_{args.n}: stepPICFLIP on a {args.n}x{args.n} MAC grid, "
                        f"{P} particles, dt 0.01, density 0.01, 0.95 FLIP",
            "solver": args.solver,
            "graph": args.graph,
            "mg_smoother": ["damped Jacobi 0.8", "two-step Chebyshev [0.4, 2]"][args.mg_smoother]
                           + " on the levels that take two sweeps (the one-workgroup levels;"
                           " level 0 takes mg_level0_sweeps, the tile levels below it"
                           " mg_level_sweeps, degree-S Chebyshev where S > 2)",
            "mg_level_sweeps": (tune or {}).get("mg_level_sweeps", 5),
            "mg_level0_sweeps": (tune or {}).get("mg_level0_sweeps", 3),
            "tol": args.tol,
            "timed_steps": [t_start, t_start + args.steps - 1],
            "autotune": tune,
            "instrumented_segment_steps_per_s": args.steps / elapsed_k,
            "untimed_steps_before": t_start,
            "whole_run": full,
            "plain_tol": args.plain_tol,
            "plain_steps_per_s": plain_sps,
            "speedup_vs_plain": (sps / plain_sps) if plain_sps else None,
            "refined": refined,
            "cg_hbm_gbs": cg_gbs,
            "solver_kernels": solver_kernels,
            "latency_bound_kernels": {k: {"ms_per_step": round(v, 4),
                                          "bytes_per_launch": launch_bytes(k)}
                                      for k, v in latency[:6]},
            "cg_iterations_last_solve": st.cg_iterations,
            "rel_err_vs_fp64_single_step": errs,
            "max_rel_err_vs_fp64": max(errs.values()) if errs else None,
            "labels_equal_fp64": same_labels,
            "kernel_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
            "batch": batch,
            "grid_only": grid,
            "adjoint": adj,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
