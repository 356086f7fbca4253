/*
 * fsg.h -- C ABI of the MI355X (gfx950) FLIP / semi-Lagrangian / pressure
 * projection hot path (libfsg.so).
 *
 * The reference (kbladin/Fluid_Simulation) has no FFI: its boundary is the C++
 * class API in include/{FluidDomain,MacGrid,MarkerParticleSet,FluidSolver}.h.
 * Each entry point below names the reference member it replaces.  The C++ drop-in
 * classes (fluid_simulation_amd/cpp/include) are written on top of this ABI.
 *
 * Conventions
 *  - Every function returns FSG_OK (0) or a negative FSG_E* status; the message of
 *    the last failure on a handle is fsg_last_error(h).
 *  - Grids are row-major float32 / uint8 arrays of nx*ny values, index i + j*nx
 *    (include/Grid.h:27-31).  u(i,j) is the left face of cell (i,j), v(i,j) the
 *    bottom face (include/MacGrid.h:66-79).
 *  - Particles are AoS float32 (x, y, vx, vy) (include/MarkerParticleSet.h:50-56).
 *  - Host arrays passed in are copied; the handle owns all device memory.
 *  - One handle owns one HIP stream on the device current at fsg_create.  A handle
 *    is not thread-safe; distinct handles are independent.
 *  - Calls are stream-ordered: fsg_step / fsg_run_phase return once the work is
 *    enqueued (except where a form needs a host decision, see fsg_params.solver);
 *    the fsg_get_* calls synchronise the stream.
 */
#ifndef FSG_H
#define FSG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSG_OK 0
#define FSG_E_INVALID (-1)   /* bad argument */
#define FSG_E_HIP (-2)       /* HIP runtime failure */
#define FSG_E_VALIDATE (-3)  /* FluidSolver::validate failure (src/FluidSolver.cpp:101-107) */
#define FSG_E_NOMEM (-4)

/* Cell labels: enum CellType (include/MacGrid.h:14-17). */
#define FSG_LIQUID 0
#define FSG_AIR 1
#define FSG_SOLID 2

/* Step schemes: FluidSolver::step* (include/FluidSolver.h:51-54) + grid-only. */
#define FSG_SCHEME_SEMI_LAGRANGIAN 0 /* stepSemiLagrangian, src/FluidSolver.cpp:99-134 */
#define FSG_SCHEME_PIC 1             /* stepPIC, src/FluidSolver.cpp:136-170 */
#define FSG_SCHEME_FLIP 2            /* stepFLIP, src/FluidSolver.cpp:173-209 */
#define FSG_SCHEME_PICFLIP 3         /* stepPICFLIP, src/FluidSolver.cpp:211-251 */
#define FSG_SCHEME_GRID_ADVECT 4     /* new: labels from the level set, gather
                                        advection, gravity, Dirichlet, projection */

/* Phases for fsg_run_phase (one reference member each). */
#define FSG_PH_CLASSIFY 1          /* FluidDomain::classifyCells(MarkerParticleSet&), src/FluidDomain.cpp:150-180 */
#define FSG_PH_P2G 2               /* transferVelocityToGridSpread, src/FluidSolver.cpp:873-919 */
#define FSG_PH_SAVE_PREV 3         /* MacGrid::updatePreviousVelocityBuffer, src/MacGrid.cpp:52-56 */
#define FSG_PH_GRAVITY 4           /* addExternalAcceleration, src/FluidSolver.cpp:276-295 */
#define FSG_PH_DIRICHLET 5         /* enforceDirichlet, src/FluidSolver.cpp:297-321 */
#define FSG_PH_EXTEND 6            /* extendVelocityIndividual, src/FluidSolver.cpp:485-622 */
#define FSG_PH_PRESSURE 7          /* pressureSolve, src/FluidSolver.cpp:323-483 */
#define FSG_PH_DIFF 8              /* MacGrid::updateVelocityDiffBuffer, src/MacGrid.cpp:58-70 */
#define FSG_PH_G2P_PIC 9           /* transferVelocityToParticlesPIC, src/FluidSolver.cpp:921-931 */
#define FSG_PH_G2P_FLIP 10         /* transferVelocityToParticlesFLIP, src/FluidSolver.cpp:933-943 */
#define FSG_PH_G2P_PICFLIP 11      /* transferVelocityToParticlesPICFLIP, src/FluidSolver.cpp:945-963 */
#define FSG_PH_ADVECT 12           /* MarkerParticleSet::advect, src/MarkerParticleSet.cpp:40-46 */
#define FSG_PH_ADVECT_REJECT 13    /* advectAndEnsureOutsideObstacles, src/MarkerParticleSet.cpp:48-62 */
#define FSG_PH_SL_SCATTER 14       /* advectVelocitySemiLagrangian, src/FluidSolver.cpp:709-772 */
#define FSG_PH_SL_GATHER 15        /* new: gather-form semi-Lagrangian advection + swap */
#define FSG_PH_ADVECT_GRID 16      /* advectParticlesWithGrid, src/FluidSolver.cpp:774-791 */
#define FSG_PH_CLASSIFY_LEVELSET 17 /* FluidDomain::classifyCells(LevelSet&), src/FluidDomain.cpp:129-148 */
#define FSG_PH_EXTEND_AVERAGE 18   /* extendVelocityAvarageing(grid, ext_layers), src/FluidSolver.cpp:625-707 */

/* Integrators for getAdvectedPosition (include/FluidSolver.h:143-145). */
#define FSG_RK3 0    /* RK3::step, include/OdeSolver.h:102-113 (the reference's choice) */
#define FSG_EULER 1  /* EulerExplicit::step, include/OdeSolver.h:78-86 */
#define FSG_RK2 2    /* new: explicit midpoint */
#define FSG_RK4 3    /* RK4::step, include/OdeSolver.h:117-142 (with the Euler-style scalar
                        shift of both coordinates its Vec2 form implies) */

/* Pressure solver forms (all replace Eigen::ConjugateGradient, FluidSolver.h:114). */
#define FSG_SOLVER_PLAIN 0      /* FP32 Jacobi-PCG, 3 kernels per iteration, host loop */
#define FSG_SOLVER_PERSISTENT 1 /* FP32 Jacobi-PCG in one persistent cooperative kernel */
#define FSG_SOLVER_REFINED 2    /* FP32 PCG inside FP64 iterative refinement, FP64
                                   residual and reductions, persistent kernel */
#define FSG_SOLVER_MG 3         /* CG preconditioned by a geometric-multigrid V-cycle
                                   (FP32), FP64 solution, residual recomputed in FP64
                                   every iteration, FP64 reductions; stops on the FP64
                                   residual |b - A x| <= cg_tol |b| */

/* Velocity buffers for fsg_{get,set}_velocity (include/MacGrid.h:162-176). */
#define FSG_VEL_FRONT 0
#define FSG_VEL_BACK 1
#define FSG_VEL_PREV 2
#define FSG_VEL_DIFF 3

typedef struct fsg_domain fsg_domain;

/* Tunables; fsg_params_default gives the reference's hard-coded values. */
typedef struct fsg_params {
  float gravity_x;        /* 0      (src/FluidSolver.cpp:232) */
  float gravity_y;        /* -9.82f (src/FluidSolver.cpp:232) */
  double cg_tol;          /* FLT_EPSILON: Eigen's default for float (no setTolerance call) */
  int32_t cg_max_iter;    /* 100    (src/FluidSolver.cpp:81) */
  int32_t ext_layers;     /* 2      (src/FluidSolver.cpp:234) */
  int32_t integrator;     /* FSG_RK3 */
  int32_t solver;         /* FSG_SOLVER_PLAIN */
  int32_t project_mean;   /* 0: remove the mean of b (fully Neumann systems) */
  int32_t check_every;    /* plain solver: iterations enqueued between host checks */
  double refine_inner_tol;/* refined solver: relative tolerance of each FP32 inner solve */
  int32_t refine_max_outer;/* refined solver: maximum FP64 refinement sweeps */
  int32_t record_ext_info; /* 1: fsg_step also records the extension sources */
  int32_t pcg_blocks;     /* persistent solvers: blocks (0: automatic) */
  int32_t pcg_kpt;        /* persistent solvers: unknowns per thread, 1/2/4 (refined form:
                             1/2), 0: automatic */
  int32_t warm_start;     /* persistent, refined and multigrid solvers: 1 starts CG
                             from the previous solve's pressure (same stopping test
                             |r| < tol |b|); 0 (default) starts from x0 = 0 like
                             Eigen's solve() */
  float particle_pressure_k; /* 0 (src/FluidSolver.cpp:443): k of the particle-count
                                pressure term k * (n > 9 ? n - 8 : 0) */
  int32_t closed_walls;   /* 0 (the reference): enforceDirichlet zeroes only flow into
                             SOLID (src/FluidSolver.cpp:306-319); 1: every face next to
                             a SOLID cell is zeroed, both directions (no-penetration).
                             With the mean of b removed (closed all-fluid domains) the
                             one-sided rule lets the update, which takes SOLID
                             neighbours as p = 0, inject fluid through the walls, and
                             the flow grows without bound; 1 keeps such runs bounded. */
  int32_t graph;          /* multigrid solver: 1 (default) replays its iterations from
                             HIP graphs captured once per handle (two iterations per
                             graph); 0 launches every kernel.  Per-kernel timing and
                             the kernel clocks launch every kernel. */
  int32_t deterministic;  /* 0 (default): the P2G splat and the particle density add
                             their float terms with float atomics, in no fixed order
                             (like any parallel sum; the reference sums in particle
                             order); 1: the terms are added as 64-bit fixed-point
                             integers (2^-40 units, |sum| < 2^23 per face), so every
                             step is bit-reproducible: a rerun from the same state
                             (a checkpointed recompute) takes the same trajectory */
  int32_t record_decisions; /* 1: each PIC / FLIP / PICFLIP / semi-Lagrangian step
                               records its discrete decisions per cell for
                               fsg_get_step_decisions */
  int32_t mg_smoother;    /* multigrid solver's smoothing (two sweeps before and after
                             the coarse correction on every level): 0 damped Jacobi,
                             weight 0.8; 1 (default) the two-step Chebyshev smoother,
                             weights the reciprocal roots of the degree-2 Chebyshev
                             polynomial on [0.4, 2] (the high end of the spectrum of
                             D^-1 A): 11.9 instead of 13.4 iterations per solve over the
                             1024^2 dam break's 500 steps */
  int32_t mg_coarse_sweeps; /* multigrid solver's coarse solve (the levels of <= 1024
                               cells, one workgroup): 0 (default) a V-cycle down to 4 x 4
                               cells; k in 1..64: k damped-Jacobi sweeps on the first of
                               those levels only (cheaper; as good where AIR cells pin
                               the coarse problem, much weaker for all-Neumann boxes) */
  int32_t graph_iters;    /* multigrid solver with graph = 1: iterations per captured
                             graph, 1..16; 0 (default): one graph per solve, as long as
                             the longer of the previous two solves, then one graph per
                             further iteration (fewer iterations enqueued after the
                             stop test, whose kernels return at once) */
  int32_t mg_fuse_update; /* multigrid solver, one domain of two or more tile levels:
                             1 (default) the CG update of x and the residual runs inside
                             the next iteration's level-0 down kernel; 0 in a kernel of
                             its own */
  int32_t splat_order;    /* P2G / classify splat: 1 (default) reads a copy of the
                             particles grouped by 16 x 16-cell bucket (a counting sort
                             each step, storage order kept within a bucket; the
                             particles' own order is untouched); 0 reads them in
                             storage order (cheaper while particles that are neighbours
                             in storage are neighbours in space) */
  int32_t mg_tile_threads;/* multigrid solver: threads per 32 x 32 level-0 tile of the
                             down / up kernels, 256, 512 (default) or 1024 (more waves
                             per tile hide more memory latency where few tiles hold
                             unknowns; fewer leave room for more tiles per CU) */
  int32_t mg_level_sweeps;/* multigrid solver: smoothing sweeps before and after the
                             coarse correction on the tile levels below level 0 (level 0
                             and the one-workgroup levels keep two), 2 .. 6: 2 uses the
                             mg_smoother weights; S > 2 the degree-S Chebyshev weights
                             on [lo_S, 2] (lo = 0.2, 0.2, 0.03, 0.05 for S = 3 .. 6).
                             Those levels are bound by launch and barrier latency, so
                             the extra sweeps cost little and save iterations: 5
                             (default) takes the developed 1024^2 dam break from 13 to 9
                             iterations per solve */
  int32_t mg_level0_sweeps;/* multigrid solver: the same on level 0, 2 (the mg_smoother
                             weights) or 3 (default: the degree-3 Chebyshev weights on
                             [0.2, 2]; one more LDS phase in each level-0 kernel, about
                             7 % fewer iterations on the dam break) */
  int32_t mg_spread_tiles;/* multigrid solver, one domain: 1 (default) the level-0 tile
                             kernels deal the tiles holding unknowns out evenly over the
                             XCDs (a block -> tile map formed per solve); 0 each XCD
                             takes one band of consecutive tiles, live or not */
} fsg_params;

/* Counters of the last pressure solve and running totals. */
typedef struct fsg_stats {
  int32_t n_fluid;        /* unknowns of the last solve */
  int32_t cg_iterations;  /* iterations of the last solve (Eigen's count) */
  double cg_error;        /* |r|/|b| reported by the last solve */
  int32_t refine_sweeps;  /* FP64 sweeps of the last refined solve */
  int32_t mg_live_tiles;  /* multigrid form: level-0 32 x 32 tiles holding an unknown in
                             the last solve (the tiles its level-0 kernels process) */
  int64_t total_unknown_iterations; /* sum over solves of n_fluid * SpMV count */
  int64_t total_solves;
  int64_t reserved[4];
} fsg_stats;

/* ---- lifecycle ------------------------------------------------------------ */
const char *fsg_version(void);
void fsg_params_default(fsg_params *p);
/* FluidDomain(size_x, size_y, length_x, length_y, density, pic_ratio)
   (include/FluidDomain.h:40-46) + MacGrid(...) (include/MacGrid.h:22). */
int fsg_create(fsg_domain **out, int32_t nx, int32_t ny, float length_x, float length_y,
               float density, float pic_ratio);
int fsg_destroy(fsg_domain *h);
const char *fsg_last_error(const fsg_domain *h);
int fsg_set_params(fsg_domain *h, const fsg_params *p);
int fsg_get_params(const fsg_domain *h, fsg_params *p);
/* FluidDomain::setPicRatio (src/FluidDomain.cpp:99-102, clamps to [0,1]). */
int fsg_set_pic_ratio(fsg_domain *h, float pic_ratio);
/* sizes, deltas, density, pic ratio: GridInterface getters (include/Grid.h:50-55),
   FluidDomain::density/picRatio (src/FluidDomain.cpp:210-218). */
int fsg_get_geometry(const fsg_domain *h, int32_t *nx, int32_t *ny, float *dx, float *dy,
                     float *density, float *pic_ratio);
int fsg_synchronize(fsg_domain *h);

/* ---- particles: MarkerParticleSet (include/MarkerParticleSet.h:53-78) ------- */
/* replaces the particle vector wholesale (FluidDomain::resetParticleSet + addParticle) */
int fsg_set_particles(fsg_domain *h, const float *xyuv, int64_t n);
/* MarkerParticleSet::addParticle (include/MarkerParticleSet.h:56), batched append */
int fsg_add_particles(fsg_domain *h, const float *xyuv, int64_t n);
/* MarkerParticleSet::size / iteration (include/MarkerParticleSet.h:65-74) */
int fsg_num_particles(const fsg_domain *h, int64_t *n);
int fsg_get_particles(fsg_domain *h, float *xyuv, int64_t capacity);

/* ---- grid state: MacGrid (include/MacGrid.h:22-159) ------------------------ */
int fsg_set_velocity(fsg_domain *h, int32_t which, const float *u, const float *v);
int fsg_get_velocity(fsg_domain *h, int32_t which, float *u, float *v);
/* MacGrid::swapVelocityBuffers (src/MacGrid.cpp:89-93) */
int fsg_swap_velocity(fsg_domain *h);
/* MacGrid::setCellType / cellType (include/MacGrid.h:92-97,134-137) */
int fsg_set_labels(fsg_domain *h, const uint8_t *labels);
int fsg_get_labels(fsg_domain *h, uint8_t *labels);
/* new: static interior solids applied wherever the border is (NULL clears) */
int fsg_set_solid_mask(fsg_domain *h, const uint8_t *mask);
/* LevelSet values used by FSG_PH_CLASSIFY_LEVELSET (include/LevelSet.h:8-18) */
int fsg_set_level_set(fsg_domain *h, const float *phi);

/* ---- stepping: FluidSolver (include/FluidSolver.h:45-54) ------------------- */
int fsg_step(fsg_domain *h, int32_t scheme, float dt);
int fsg_run_phase(fsg_domain *h, int32_t phase, float dt);
/* FluidSolver::validate (src/FluidSolver.cpp:89-97): the pool a solver was built
   from must match the domain; returns FSG_OK or FSG_E_VALIDATE with the
   reference's message. */
int fsg_validate_pool(fsg_domain *h, int32_t pool_nx, int32_t pool_ny, float pool_dx,
                      float pool_dy);

/* ---- inspection (parity outputs) ------------------------------------------ */
/* pressure x per cell (LIQUID cells; 0 elsewhere) and b per cell, last solve */
int fsg_get_pressure(fsg_domain *h, float *p);
int fsg_get_rhs(fsg_domain *h, float *b);
/* FluidSolverMemoryPool::fluid_indices (include/FluidSolver.h:29) */
int fsg_get_fluid_indices(fsg_domain *h, int32_t *idx);
/* FluidSolverMemoryPool valid masks (front pair, include/FluidSolver.h:33-36) and
   per-face extension sources: (layer << 4) | neighbour bits, neighbour bit order
   (i-1,j) (i,j-1) (i,j+1) (i+1,j), layer 0 = valid before extension, 15 = never */
int fsg_get_valid_masks(fsg_domain *h, uint8_t *mask_u, uint8_t *mask_v);
int fsg_get_ext_info(fsg_domain *h, uint8_t *info_u, uint8_t *info_v);
/* per particle: 1 where advectAndEnsureOutsideObstacles stepped back */
int fsg_get_reject_flags(fsg_domain *h, uint8_t *flags);
/* The discrete decisions of the last step (fsg_params.record_decisions = 1), one byte
 * per cell: bit 0 / 1 u / v normalised by the P2G weight test w > 1e-6
 * (src/FluidSolver.cpp:898-917), bit 2 / 3 u / v zeroed by the first enforceDirichlet
 * (:233), bit 4 / 5 by the second (:240).  With the labels and the reject flags they
 * fix the branch of the piecewise-smooth step a gradient is taken on. */
int fsg_get_step_decisions(fsg_domain *h, uint8_t *dec);
int fsg_get_stats(fsg_domain *h, fsg_stats *s);
int fsg_reset_stats(fsg_domain *h);

/* ---- emission and level set ------------------------------------------------ */
/* FluidSource::update (src/FluidDomain.cpp:29-52) on the device: appends
   nxp * nyp particles (xs[i], ys[j], vx, vy), j outer, i inner.  xs / ys are the
   emitter's MyFloat accumulations (x = x_min; x < x_max; x += delta_x / 2.5). */
int fsg_emit_lattice(fsg_domain *h, const float *xs, int32_t nxp, const float *ys, int32_t nyp,
                     float vx, float vy);
/* FluidDomain::advectLevelSet (src/FluidDomain.cpp:104-127) on the level set set by
   fsg_set_level_set, with the front velocity. */
int fsg_advect_level_set(fsg_domain *h, float dt);
int fsg_get_level_set(fsg_domain *h, float *phi);

/* ---- packed batch of scenes (no counterpart in the reference) ---------------
   Many independent scenes stepped together: each scene (nx[s] x ny[s] cells of
   spacing delta, its own 1-cell SOLID border) gets a rectangle of one packed grid,
   the rest of which is SOLID.  The handle is an fsg_domain over the packed grid:
   fsg_set_params, fsg_step (PIC / FLIP / PICFLIP, FSG_SOLVER_MG, no particle
   pressure term), fsg_get_stats and the timing calls apply to the whole batch, one
   kernel launch per phase for all scenes.  Particles are kept in scene-local
   coordinates.  The pressure solve runs one multigrid-preconditioned CG per scene in
   lock step (per-scene step lengths and stopping test), so every scene meets
   |r_s| < cg_tol |b_s|; fsg_get_stats reports the largest scene's |r_s| / |b_s|.
   A scene with no LIQUID cell still swaps with the others (the reference's
   no-fluid early return is per packed grid). */
int fsg_batch_create(fsg_domain **out, int32_t n_scenes, const int32_t *nx, const int32_t *ny,
                     float delta, float density, float pic_ratio);
int fsg_batch_num_scenes(const fsg_domain *h, int32_t *n);
/* rect4 = first cell (x, y) in the packed grid, nx, ny */
int fsg_batch_scene_rect(const fsg_domain *h, int32_t scene, int32_t *rect4);
/* all scenes at once: counts[s] particles of scene s, concatenated in scene order */
int fsg_batch_set_particles(fsg_domain *h, const int64_t *counts, const float *xyuv);
/* interior SOLID cells of one scene (ny x nx, row-major) */
int fsg_batch_set_solid_mask(fsg_domain *h, int32_t scene, const uint8_t *mask);
int fsg_batch_num_particles(const fsg_domain *h, int32_t scene, int64_t *n);
int fsg_batch_get_particles(fsg_domain *h, int32_t scene, float *xyuv, int64_t capacity);
int fsg_batch_get_velocity(fsg_domain *h, int32_t scene, int32_t which, float *u, float *v);
int fsg_batch_get_labels(fsg_domain *h, int32_t scene, uint8_t *labels);

/* ---- checkpoints and reverse mode (no counterpart in the reference) -----------
   The carried state of a step is the particles plus the front and back velocity
   buffers; a checkpoint slot (0 <= slot < 65536) keeps a device copy.  The reverse
   pass of a run: for each step from the last, load (or recompute) the state at the
   step's start and call fsg_adjoint_step, which steps forward once with a tape and
   carries the adjoint state (per particle d/d(x, y, vx, vy); the front and back
   buffers' d/d(u, v)) from the step's end to its start.  A PIC / FLIP / PICFLIP step
   reads only the back buffer of its start state (the front is overwritten first), a
   semi-Lagrangian step (FluidSolver::stepSemiLagrangian, src/FluidSolver.cpp:99-134)
   only the front (the back is zeroed first): the other buffer's adjoint is zero at
   the step's start.  Discrete decisions (labels, weight tests, Dirichlet, extension
   masks, obstacle rejection) are held at the forward run's; the adjoint of the
   pressure projection is a second multigrid-preconditioned solve on the same
   operator.  Needs FSG_SOLVER_MG with warm_start = 0 and at least one LIQUID cell. */
int fsg_checkpoint_save(fsg_domain *h, int32_t slot);
int fsg_checkpoint_load(fsg_domain *h, int32_t slot);
int fsg_checkpoint_clear(fsg_domain *h);
/* adjoint state: float[4*n] per particle (x, y, vx, vy) and the back buffer's
   (ny x nx) u and v; null arguments set zeros / are skipped.  fsg_adjoint_set also
   zeroes the front buffer's adjoint. */
int fsg_adjoint_set(fsg_domain *h, const float *pbar_xyuv, const float *bbar_u,
                    const float *bbar_v);
int fsg_adjoint_get(fsg_domain *h, float *pbar_xyuv, float *bbar_u, float *bbar_v);
/* the front buffer's adjoint (ny x nx u and v); null arguments set zeros / are skipped */
int fsg_adjoint_set_front(fsg_domain *h, const float *fbar_u, const float *fbar_v);
int fsg_adjoint_get_front(fsg_domain *h, float *fbar_u, float *fbar_v);
int fsg_adjoint_step(fsg_domain *h, int32_t scheme, float dt);
/* iterations of the last adjoint pressure solve */
int fsg_adjoint_last_iterations(const fsg_domain *h, int32_t *iters);
/* particle density: every particle splats weight 1 at (x/dx, y/dy) with
   Grid::addToValueInterpolated's weights (include/Grid.h:152-184), ny x nx */
int fsg_particle_density(fsg_domain *h, float *rho);
/* adjoint state := (d density / d positions)^T rho_bar, zero elsewhere (both
   buffers' adjoints included) */
int fsg_adjoint_seed_density(fsg_domain *h, const float *rho_bar);

/* ---- timing --------------------------------------------------------------- */
/* Per-kernel HIP-event timing on the handle's stream: 0 off (default), 1 every
   kernel, 2 only the kernels whose timing name starts with the prefix set by
   fsg_set_timing_prefix (default "cg_": the pressure-solver kernels). */
int fsg_set_timing(fsg_domain *h, int32_t enable);
/* A multigrid kernel's own clock (no counterpart in the reference): its duration
   per launch from the device wall clock, free of the launch gaps that HIP events
   around a launch include.  enable = 1: the one-workgroup kernel k_mg_coarse (start
   to end), which runs as its own kernel only where the grid has fewer than two tile
   levels (elsewhere the last tile level's down kernel runs that stage, and mode 1
   records no launch); enable = 2: level 0's down kernel (earliest block start to
   latest block end, the first 4096 launches; the span buffer is kept until
   fsg_destroy); 0 stops it.  A nonzero value clears the total, which then covers
   every launch, early exits included. */
int fsg_set_kernel_clock(fsg_domain *h, int32_t enable);
int fsg_get_kernel_clock(fsg_domain *h, double *us_total, int64_t *launches);
/* Name prefix selected by timing mode 2 (at most 63 characters). */
int fsg_set_timing_prefix(fsg_domain *h, const char *prefix);
/* Kernel names and their summed durations (ms) and launch counts since the last
   reset; returns the number of kernels (writes at most `cap` entries). */
int fsg_get_timing(fsg_domain *h, int32_t cap, const char **names, double *ms,
                   int64_t *launches);
int fsg_reset_timing(fsg_domain *h);

/* ---- autotuner (no counterpart in the reference) -----------------------------
 * Chooses the step's settings for the handle's workload by timing whole steps
 * (scheme, dt) from its current state: every combination of the smoothing sweeps
 * (mg_level_sweeps, mg_level0_sweeps) {(5, 3), (5, 2), (2, 2)}, mg_coarse_sweeps
 * {0, 8, 16}, mg_fuse_update {1, 0} (one domain), graph_iters {0, 2}, splat_order
 * {1, 0} (with particles) and mg_tile_threads {256, 512} (144 for one domain with
 * particles, 72 for a batch); each runs one untimed step, then `steps`
 * (1..64) timed steps, each from the saved state.  The fastest is set on the handle
 * unless the handle's own configuration is within 2 % of it, and the state
 * (particles, velocity grids, labels, reject flags, statistics) is restored, so the
 * next fsg_step continues the trajectory the call found.  Needs FSG_SOLVER_MG;
 * allocates a copy of the state for the duration of the call. */
typedef struct fsg_tune {
  int32_t mg_smoother, mg_coarse_sweeps, mg_fuse_update, graph_iters; /* chosen */
  int32_t splat_order;                                                /* chosen */
  int32_t candidates;       /* configurations timed */
  int32_t cg_iterations;    /* the chosen configuration's iterations (last timed step) */
  int32_t mg_tile_threads;  /* chosen */
  double ms_per_step;       /* the chosen configuration's mean step time */
  double ms_per_step_before;/* the handle's configuration before the call (if a candidate) */
  int32_t mg_level_sweeps;  /* chosen */
  int32_t mg_level0_sweeps; /* chosen */
  int64_t reserved[2];
} fsg_tune;
int fsg_autotune(fsg_domain *h, int32_t scheme, float dt, int32_t steps, fsg_tune *out);

/* ---- an entry point that allocates nothing ----------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The P2G transfer of fsg_run_phase(FSG_PH_P2G)
 * (transferVelocityToGridSpread, src/FluidSolver.cpp:873-919) on caller-owned
 * device buffers: n particles (device AoS x, y, vx, vy) splatted with the
 * deterministic fixed-point sums of fsg_params.deterministic = 1, then
 * back := sum / weight at the faces whose weight exceeds 1e-6; other faces keep their
 * values (the caller swaps the buffers, as MacGrid::swapVelocityBuffers does).  The
 * grid is nx x ny cells of length_x / nx by length_y / ny, like fsg_create's.  The
 * sums live in the library's static device scratch of FSG_NOALLOC_MAX_CELLS cells
 * (32 MiB per GPU, held until the process ends): bigger grids, negative sizes and
 * null buffers return FSG_E_INVALID before any device call.  Every call on one GPU
 * shares that scratch, so calls that overlap on one GPU race; calls on one stream do
 * not.  Work is enqueued on `stream` (a hipStream_t; NULL: the null stream) and the
 * call returns without synchronising. */
#define FSG_NOALLOC_MAX_CELLS (1 << 20)
int fsg_p2g_noalloc(const float *particles_xyuv, int64_t n, int32_t nx, int32_t ny,
                    float length_x, float length_y, float *back_u, float *back_v,
                    void *stream);

#ifdef __cplusplus
}
#endif
#endif
