// fsg_kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <math.h>

#include "fsg_device.h"

namespace fsg {

inline bool is_pow2(float x) {
  if (!(x > 0.f) || !isfinite(x)) return false;
  int e;
  const float m = frexpf(x, &e);
  return m == 0.5f;
}

// MacGrid geometry: spacing = length / size in float (src/MacGrid.cpp:8), lengths
// = size * spacing (GridInterface::lengthX, include/Grid.h:54-55).
inline Geom make_geom(int nx, int ny, float length_x, float length_y) {
  Geom g{};
  g.nx = nx;
  g.ny = ny;
  g.dx = length_x / (float)nx;
  g.dy = length_y / (float)ny;
  g.pow2_x = is_pow2(g.dx);
  g.pow2_y = is_pow2(g.dy);
  g.inv_dx = g.pow2_x ? 1.0f / g.dx : 0.f;
  g.inv_dy = g.pow2_y ? 1.0f / g.dy : 0.f;
  g.len_x = (float)nx * g.dx;
  g.len_y = (float)ny * g.dy;
  return g;
}

// Batched form: scenes packed into one grid.  Scene s occupies the rectangle whose
// first cell is `cell_off` with row stride `pitch` (the packed grid's nx); its
// geometry `g` is the scene's own (sizes, spacing, lengths), and its particles are
// pt[part_base, part_base + np) in scene-local coordinates.
struct BatchScene {
  Geom g;
  int64_t part_base;
  int64_t np;
  int cell_off;
  int pitch;
  int bucket_base;  // the splat's bucket order: the scene's first bucket over all scenes
  int bucket_bx;    // and its buckets per row (16 x 16 cells each)
};
// Particle tile t (256 particles) belongs to scene ptile[t].x, local tile ptile[t].y.
struct BatchTables {
  const BatchScene* scene;
  const int2* ptile;
  int n_ptile;
};

// k_face_update phase flags (applied in this order)
enum { FU_P2G = 1, FU_PREV = 2, FU_GRAVITY = 4, FU_DIRICHLET = 8, FU_EXTINIT = 16,
       FU_CLOSED = 32 };

void launch_cell_init(hipStream_t s, const Geom& g, uint8_t* label, const uint8_t* solid,
                      float4* acc, int do_labels, int do_acc);
// The cell set-up of a step: label (if set) := SOLID on the border and where solid is
// set, else AIR; acc (if set) := 0.
struct CellInit {
  uint8_t* label = nullptr;
  const uint8_t* solid = nullptr;
  float4* acc = nullptr;
};
// acc_fx set: the splat accumulates in 64-bit fixed point there (4 per cell, zero on
// entry; launch_fx_to_float turns the sums into acc and zeroes them again).
void launch_particle_scatter(hipStream_t s, const Geom& g, const float4* pt, int64_t np,
                             uint8_t* label, float4* acc, int do_classify, int do_p2g,
                             const BatchTables* batch = nullptr, long long* acc_fx = nullptr);
void launch_fx_to_float(hipStream_t s, int64_t n, long long* fx, float* out);
// The splat's bucket order for one domain (np < 2^31, bucket_count(g) <= kMaxBuckets):
// cnt (zero on entry, zero again on return) and off hold bucket_count(g) ints, rank np
// ints; sorted receives the np particles grouped by bucket.  ci: the step's cell set-up
// (launch_cell_init's work), done by the count kernel as well.
constexpr int kMaxBuckets = 16384;  // the count kernel's LDS histogram: 64 KiB
int bucket_count(const Geom& g);
void launch_bucket_order(hipStream_t s, const Geom& g, const float4* pt, int64_t np, int* cnt,
                         int* off, int* rank, float4* sorted, const CellInit& ci = CellInit{});
// A packed batch: every scene's particles grouped by the 16 x 16-cell buckets of its own
// grid (BatchScene::bucket_base / bucket_bx), within the scene's own particle range, so
// the batch's particle tiles address the sorted copy unchanged.  n_buckets over all
// scenes; max_scene_buckets sizes the count kernel's LDS histogram.
void launch_bucket_order_batch(hipStream_t s, const float4* pt, int64_t np, int* cnt, int* off,
                               int* rank, float4* sorted, const BatchTables& tb, int n_buckets,
                               int max_scene_buckets);
void launch_face_update(hipStream_t s, const Geom& g, int flags, float* u, float* v,
                        const float4* acc, float* prev_u, float* prev_v, const uint8_t* label,
                        float gx_dt, float gy_dt, float* ext_u, float* ext_v, uint8_t* mxf,
                        uint8_t* mxb, uint8_t* myf, uint8_t* myb, uint8_t* info_u,
                        uint8_t* info_v, uint8_t* dec = nullptr);
void launch_extend_sweep(hipStream_t s, const Geom& g, const uint8_t* label, float* ub,
                         float* vb, const uint8_t* mxf, uint8_t* mxb, const uint8_t* myf,
                         uint8_t* myb, uint8_t* info_u, uint8_t* info_v, int layer);
void launch_classify_levelset(hipStream_t s, const Geom& g, const float* phi,
                              const uint8_t* solid, uint8_t* label);
struct CgScalars;
// The multigrid solve's last step, done by the pressure update (single domain, no mean
// projection): x = x[iter & 1] at the unknowns (inf0 & kMgUnknown), 0 elsewhere, into
// x64 (the solve's x[0]) and xg, as the separate finish kernel would.
struct MgFinishFuse {
  const uint8_t* inf0 = nullptr;
  const double* x1 = nullptr;
  const CgScalars* sc = nullptr;
};
void launch_pressure_update(hipStream_t s, const Geom& g, const int* n_fluid,
                            const uint8_t* label, float* xg, double* x64,
                            const int* counts, float ppk, float* au, float* av,
                            float* bu, float* bv, float dt_over_density, int do_dirichlet,
                            int do_diff, const float* prev_u, const float* prev_v, float* diff_u,
                            float* diff_v, uint8_t* zmask = nullptr, uint8_t* dec = nullptr,
                            const MgFinishFuse& fin = MgFinishFuse{});
void launch_diff(hipStream_t s, const Geom& g, const float* u, const float* v, const float* pu,
                 const float* pv, float* du, float* dv);
void launch_particle_gather(hipStream_t s, const Geom& g, float4* pt, int64_t np, const float* u,
                            const float* v, const float* du, const float* dv, float pic_ratio,
                            int g2p, int move, float dt, const uint8_t* label, uint8_t* reject,
                            const BatchTables* batch = nullptr);
// fx != null: the deterministic form (2 C fixed-point sums, zero on entry and on exit;
// the back buffer is written whole, no zero fill needed)
void launch_sl_scatter(hipStream_t s, const Geom& g, const uint8_t* label, const float* uf,
                       const float* vf, float* ub, float* vb, int integrator, float dt,
                       long long* fx = nullptr);
void launch_sl_gather(hipStream_t s, const Geom& g, const uint8_t* label, const float* uf,
                      const float* vf, float* ub, float* vb, int integrator, float dt);
void launch_advect_grid(hipStream_t s, const Geom& g, float4* pt, int64_t np, const float* uf,
                        const float* vf, int integrator, float dt);
void launch_fill(hipStream_t s, float* a, int64_t n, float val);
void launch_count_particles(hipStream_t s, const Geom& g, const float4* pt, int64_t np,
                            int* counts);
// init != 0: masks + back := front; else one sweep (front mask mf, back mask mb)
void launch_extend_average(hipStream_t s, const Geom& g, const uint8_t* label, const float* uf,
                           const float* vf, float* ub, float* vb, uint8_t* mf, uint8_t* mb,
                           int init);
void launch_emit_lattice(hipStream_t s, float4* pt, int64_t base, const float* xs, int nxp,
                         const float* ys, int nyp, float vx, float vy);
void launch_advect_level_set(hipStream_t s, const Geom& g, const float* phi, float* out,
                             const float* u, const float* v, float dt);

// ---------------------------------------------------------------- pressure
constexpr int kScanChunk = 2048;  // cells per block of the liquid count / compaction

// Info bit of the multigrid form's level-0 cells: the cell is an unknown (a LIQUID cell
// with a non-SOLID neighbour).
constexpr uint8_t kMgUnknown = 0x80;

// Device-resident scalars of one solve.
struct CgScalars {
  double rhs_norm2;      // |b|^2 (after the optional mean removal)
  double threshold;      // max(tol^2 |b|^2, min normal)
  double rr;             // |r|^2 after the last update
  double err;            // sqrt(|r|^2 / |b|^2)
  float rz[2];           // r.z of the current / next direction (iteration parity)
  float beta;
  int iter;              // iterations (Eigen's count for the plain form)
  int spmv;              // SpMV applications
  int done;              // 1 once the solve has stopped
  int refine_sweeps;
  int live_tiles;        // multigrid form: level-0 32 x 32 tiles holding an unknown
  double b64_norm2;      // refined form: |b|^2 in FP64
  double res64_norm2;    // refined form: |b - A x|^2 in FP64
  unsigned long long total_unknown_iters;
  long long total_solves;
  double rho[2];         // multigrid form: r.z of the current / previous direction
};

// The scalars a solve's host loop reads back after each batch of iterations, written by
// one small kernel into pinned host memory (one launch instead of two device-to-host
// copies).  n_fluid is -1 when not asked for.
struct HostScalars {
  CgScalars sc;
  int n_fluid;
  int pad;
};
void launch_read_scalars(hipStream_t s, const CgScalars* sc, const int* n_fluid,
                         HostScalars* out);

struct CgParams {
  int max_iter;
  float tol;             // in float: Eigen's RealScalar tolerance
  float off;             // 1/dx^2 as float (src/FluidSolver.cpp:381-402)
  float diag_tab[5];     // -n/dx^2 as float for n non-SOLID neighbours (:409-410)
  float invd_tab[5];     // Eigen's DiagonalPreconditioner: 1/diag, or 1 where diag == 0
  int project_mean;
};

void launch_count_liquid(hipStream_t s, const Geom& g, const uint8_t* label, int* counts);
void launch_scan_counts(hipStream_t s, int nb, const int* counts, int* offsets, int* n_fluid);
void launch_compact(hipStream_t s, const Geom& g, const uint8_t* label, const int* offsets,
                    int* fluid_index, int* cell_of, const float* uf, const float* vf, float* b,
                    uint8_t* info);
void launch_updown(hipStream_t s, int nx, int F_bound, const int* n_fluid, const int* cell_of,
                   const uint8_t* info, const int* fluid_index, int2* updn);

// plain form (partials: >= 12288 doubles)
void launch_cg_init(hipStream_t s, int F, const CgParams& cp, const int* cell_of,
                    const uint8_t* info, const float* b, float* r, float* pg, float* xg,
                    double* partials, CgScalars* sc);
void launch_cg_iteration(hipStream_t s, int F, int nx, int it, const CgParams& cp,
                         const int* cell_of, const uint8_t* info, float* r, float* ap, float* pg,
                         float* xg, double* partials, CgScalars* sc);

// persistent forms (pipelined PCG, optionally inside FP64 refinement)
struct PersistentArgs {
  int nx;
  int F_max;             // capacity of the compact arrays
  int F_actual;          // unknowns of this solve (host copy, sizes the launch)
  const int* n_fluid;
  const int* cell_of;
  const uint8_t* info;
  const int2* updn;
  const float* b;
  float* xg;             // solution, grid layout
  float* exch;           // 3 x F_max floats: device-coherent z and p (x2) exchange
  double* x64;           // F_max doubles: FP64 solution exchange (refined form)
  double* partials;      // 2 x grid x 4 doubles
  unsigned* bar;         // [0] arrival counter, [1] abort flag
  CgScalars* sc;
  CgParams cp;
  int refined;
  int max_outer;
  float inner_tol;
  double outer_tol;
  int warm_start;        // 1: x0 = previous solution (xg), else x0 = 0
};
int pcg_max_kpt();
size_t pcg_lds_bytes(int F, int grid, int nx);
hipError_t launch_pcg(hipStream_t s, const PersistentArgs& a, int grid, int kpt);
hipError_t pcg_set_lds_limit(size_t max_bytes);

// multigrid-preconditioned CG with FP64 solution and residual (dense cell grids)
constexpr int kMgMaxLevels = 16;
constexpr int kMgMaxVecBlocks = 1024;  // blocks of the vector kernels (partials per kernel)
constexpr int kMgMaxSweeps = 6;     // smoothing sweeps of the tile levels below level 0, at most
constexpr int kMgCoarseCells = 1024;   // levels at or below this size run in one workgroup

// Smoothing: every level's pre- and post-smoothing is two relaxation sweeps
// x += w (r - K x) / n, the first with omega, the second with omega2 (damped Jacobi:
// both 0.8; the two-step Chebyshev smoother: the reciprocals of the roots of the
// degree-2 Chebyshev polynomial on [0.4, 2], the high end of the spectrum of D^-1 K).
struct MgCoarseArgs {
  int lc, nlev, total, sweeps;
  float omega, omega2;
  int trunc;  // 0: V-cycle over levels lc .. nlev-1; k > 0: k Jacobi sweeps on level lc only
  int nx[kMgMaxLevels], ny[kMgMaxLevels], off[kMgMaxLevels];
  const uint8_t* inf_all;  // info of levels lc .. nlev-1, contiguous (offsets off[l] - off[lc])
  const float* r_in;  // right-hand side of level lc
  float* x_out;       // cycle result of level lc
};

// Packed batch of scenes: scene s's own hierarchy below level `la` of the packed one.
constexpr int kSceneCoarseCells = 4096;  // cells of a scene's rectangle at level la, at most
constexpr int kSceneMaxLevels = 8;
struct SceneCoarse {
  int cx, cy;     // first cell of the scene's rectangle in the packed grid of level la
  int nl, total;  // its levels (the first is the rectangle); cells over all, a multiple of 4
  int nx[kSceneMaxLevels], ny[kSceneMaxLevels], off[kSceneMaxLevels];
};
struct SceneCoarseArgs {
  const SceneCoarse* tab = nullptr;
  int n_scenes = 0;
  int pitch = 0;              // width of the packed grid of level la
  int stride = 0;             // bytes of info per scene in inf_all (>= every total, mult. of 4)
  uint8_t* inf_all = nullptr; // scene s's info at s * stride
  const uint8_t* lab = nullptr;  // packed labels of level la
  const float* r_in = nullptr;   // packed right-hand side of level la
  float* x_out = nullptr;        // packed cycle result of level la
  float omega = 0.8f, omega2 = 0.8f;
  int trunc = 0;  // as MgCoarseArgs::trunc, on each scene's first level
};

// Packed batch: per-scene CG scalars (one CG per scene in lock step).
struct MgbArgs {
  const int4* tile = nullptr;  // (x0, y0, scene, 0): 64 x 16 tiles that lie in one scene
  const int2* tend = nullptr;  // per tile: its scene's column and row ends (exclusive)
  const int4* rect = nullptr;  // per scene: (x, y, nx, ny) in the packed grid
  int n_tiles = 0, n_scenes = 0, W = 0, C = 0;  // C: cells of the packed grid
  double* rz = nullptr;  // [2][S] r.z by iteration parity
  double* pq = nullptr;  // [S]
  double* rr = nullptr;  // [S]
  double* bb = nullptr;  // [S]
  const uint8_t* vflag = nullptr;  // per tile: 1 if it holds an unknown (set each solve)
};

struct MgHost {
  bool per_scene = false;  // a batch: stop test and CG scalars per scene (MgbArgs)
  unsigned long long* clk = nullptr;  // k_mg_coarse's own clock: ticks, launches
  bool clk_on = false;
  // level 0's down kernel clock: per recorded launch and block, (start, end) ticks
  unsigned long long* clk_span = nullptr;
  int clk_span_cap = 0;      // launches the buffer holds
  int clk_span_nb = 0;       // blocks per launch
  int clk_span_n = 0;        // launches recorded
  bool clk_span_on = false;
  uint8_t* tflag = nullptr;  // level-0 tiles (32 x 32) holding an unknown on level 0 or 1
  // one domain: the level-0 tile kernels' block -> tile map (k_mg_tile_map), formed per
  // solve from tflag: tile, bit 31 set where the tile is empty
  int* tmap = nullptr;
  int* tmap_buf = nullptr;  // its storage (tmap is null where the map is not used)
  // a batch: the same flags for its tile levels 1 .. lc-1 (most of a packed grid is air,
  // obstacles or gaps on every level)
  uint8_t* tflag_l[kMgMaxLevels] = {};
  // fused form (one domain, two or more tile levels): level 0's down kernel also does
  // the previous iteration's x and residual update, level 1's runs the stop test
  bool fused = false;
  int l0_threads = 256;  // threads per level-0 tile of the down / up kernels (256/512/1024)
  // sweeps before and after the coarse correction on the tile levels below level 0: 2
  // (the level's omega, omega2) or 3 .. kMgMaxSweeps (degree-S Chebyshev weights)
  int level_sweeps = 2;
  int level0_sweeps = 2;  // the same on level 0: 2 (omega, omega2) or 3 (degree-3 Chebyshev)
  // the last tile level's down kernel also runs the one-workgroup stage (its last
  // workgroup to finish), so no k_mg_coarse launch; ticket: its arrival counter
  bool coarse_tail = true;
  unsigned* ticket = nullptr;
  int nlev = 0, lc = 0;
  int nx[kMgMaxLevels] = {}, ny[kMgMaxLevels] = {};
  uint8_t* lab[kMgMaxLevels] = {};  // [0]: the domain's labels
  uint8_t* inf[kMgMaxLevels] = {};
  float* rl[kMgMaxLevels] = {};     // right-hand side per level ([0]: the CG residual r)
  float* xs[kMgMaxLevels] = {};     // after pre-smoothing
  float* xo[kMgMaxLevels] = {};     // after the cycle ([0]: z)
  double* b = nullptr;              // b' = -dx^2 b (FP64)
  double* x[2] = {nullptr, nullptr};  // FP64 solution (ping-pong)
  float* p[2] = {nullptr, nullptr};   // search direction (ping-pong)
  double *part_sum = nullptr, *part_cnt = nullptr, *part_bb = nullptr, *part_rr = nullptr,
         *part_pq = nullptr, *part_rz = nullptr;
  double* part_liq = nullptr;  // LIQUID cells per setup block (the solve without compaction)
  MgCoarseArgs coarse{};
  size_t coarse_lds = 0;
};

// Packed batch of scenes: each scene's right-hand side is scaled to unit norm before
// the joint solve and its solution scaled back, so that one stopping test bounds
// every scene's own relative residual.
struct MgBatchScale {
  int n_scenes;
  const int4* rect;  // per scene: first cell (x, y) in the packed grid, nx, ny
  const int* sid;    // per packed cell: scene, or -1 outside every scene
  double* norm;      // per scene: |b'_s|
  double* inv;       // per scene: 1 / |b'_s| (0 when b'_s = 0)
};

int mg_vec_blocks(int C);
bool mg_tail_coarse(const MgHost& m);  // the one-workgroup stage runs in the down kernel
int mg_part_rr_size(const MgHost& m);  // entries of MgHost::part_rr
int mg_up_blocks(int nx, int ny);
int mg_dir_blocks(const MgHost& m);  // blocks of k_mg_dir (partials of p.Kp)
hipError_t mg_set_lds_limit(size_t bytes);
// adj_rhs != null: the right-hand side is the adjoint solve's cell field (the
// pressure's adjoint) instead of the divergence of (uf, vf); see k_mg_adj_rhs0.
// rhs, n_fluid (optional, divergence right-hand side only): the reference's FP32 b at
// every LIQUID cell (0 elsewhere) and the number of LIQUID cells, the outputs the
// compaction of the other solver forms gives, so the multigrid form skips it.
void launch_mg_setup(hipStream_t s, const Geom& g, const MgHost& m, const uint8_t* label,
                     const float* uf, const float* vf, int project_mean, int warm,
                     CgScalars* sc, const MgBatchScale* batch = nullptr,
                     const double* adj_rhs = nullptr, float* rhs = nullptr,
                     int* n_fluid = nullptr);
// a batch's per-scene CG: init after the setup; the stop test at the start of each
// iteration; the vector kernels (which: 0 r.z, 1 direction, 2 x / r update)
void launch_mgb_init(hipStream_t s, const MgbArgs& a, const double* norm, const MgHost& m);
void launch_mgb_check(hipStream_t s, const MgbArgs& a, int it, int max_iter, double tol,
                      CgScalars* sc);
void launch_mgb_vectors(hipStream_t s, const MgbArgs& a, const MgHost& m, int it, CgScalars* sc,
                        int which);
// removes the plain mean of the FP64 solution over the unknowns (adjoint solves)
void launch_mg_remove_mean(hipStream_t s, const MgHost& m);
// one CG iteration: down(l) for l < lc, coarse, up(l) for l = lc-1 .. 0, dir, xr
// (launches through launch_mg_down record level 0's block spans while clk_span_on)
void launch_mg_down(hipStream_t s, MgHost& m, int l, int it, int max_iter, double tol,
                    CgScalars* sc);
int mg_down_blocks0(const MgHost& m);  // blocks of level 0's down kernel
void launch_mg_coarse(hipStream_t s, const MgHost& m, int it, int max_iter, double tol,
                      CgScalars* sc);
// ---- adjoint (reverse-mode) kernels of the PIC / FLIP step, kernels_adj.hip
void launch_adj_g2p(hipStream_t s, const Geom& g, const float4* x0, int64_t np, const float* hu,
                    const float* hv, const float* du, const float* dv, float r, float dt,
                    const uint8_t* reject, float4* pbar, float* su, float* sv);
void launch_adj_post(hipStream_t s, const Geom& g, const uint8_t* label, const uint8_t* zmask,
                     float r, int flip, const float* su, const float* sv, const float* fu,
                     const float* fv, float* eu, float* ev, float* g4u, float* g4v, float* g1u,
                     float* g1v);
void launch_adj_pgrad(hipStream_t s, const Geom& g, const uint8_t* label, const float* eu,
                      const float* ev, float dt_over_density, double* pbar);
void launch_adj_div(hipStream_t s, const Geom& g, const uint8_t* inf, const double* bbar,
                    const float* nu, const float* nv, float* eu, float* ev);
void launch_adj_mask_init(hipStream_t s, const Geom& g, const uint8_t* label, uint8_t* r0,
                          uint8_t* b0);
void launch_adj_mask_step(hipStream_t s, const Geom& g, const uint8_t* label, const uint8_t* rk,
                          const uint8_t* bk, uint8_t* rn, uint8_t* bn);
void launch_adj_ext_sweep(hipStream_t s, const Geom& g, const uint8_t* label, const uint8_t* rk,
                          const float* inu, const float* inv, float* outu, float* outv);
void launch_adj_pre(hipStream_t s, const Geom& g, const uint8_t* label, const uint8_t* m0,
                    const float* e0u, const float* e0v, const float* g4u, const float* g4v,
                    const float* g1u, const float* g1v, const float* prev_u,
                    const float* prev_v, const float4* acc, float gx_dt, float gy_dt,
                    int closed, float* qu, float* qv, float* nu, float* nv);
void launch_adj_p2g(hipStream_t s, const Geom& g, const float4* x0, int64_t np, const float* g1u,
                    const float* g1v, const float* qu, const float* qv, float4* pbar);
// rho_fx set (C int64, zero on entry and on return): fixed-point accumulation, and
// rho is overwritten with the sums; otherwise the terms are added to rho.
void launch_density(hipStream_t s, const Geom& g, const float4* pt, int64_t np, float* rho,
                    long long* rho_fx = nullptr);
void launch_density_adj(hipStream_t s, const Geom& g, const float4* pt, int64_t np,
                        const float* rb, float4* pbar);
// the semi-Lagrangian step's adjoint: particle advection through the grid (pbar in /
// out, the grid's adjoint added to su, sv), then gravity + the first Dirichlet pass and
// the velocity scatter (the front's adjoint at the step's start into fu, fv from the
// adjoint abu, abv of the pressure update's source and sbu, sbv of the scatter)
void launch_adj_advect_grid(hipStream_t s, const Geom& g, const float4* x0, int64_t np,
                            const float* uf, const float* vf, int integrator, float dt,
                            float4* pbar, float* su, float* sv);
void launch_adj_sl(hipStream_t s, const Geom& g, const uint8_t* label, const float* f0u,
                   const float* f0v, const float* abu, const float* abv, const float* sbu,
                   const float* sbv, float gx_dt, float gy_dt, int closed, int integrator,
                   float dt, float* fu, float* fv);

// a batch: every scene's own levels (labels / info once per solve, a cycle per iteration)
void launch_mg_scene_setup(hipStream_t s, const SceneCoarseArgs& a);
void launch_mg_scene_coarse(hipStream_t s, const SceneCoarseArgs& a, CgScalars* sc);
void launch_mg_up(hipStream_t s, const MgHost& m, int l, int it, CgScalars* sc);
void launch_mg_dir(hipStream_t s, const MgHost& m, int it, CgScalars* sc);
void launch_mg_xr(hipStream_t s, const MgHost& m, int it, CgScalars* sc);
// x at the unknowns -> xg (FP32) and m.x[0] (FP64, read by the pressure update)
void launch_mg_finish(hipStream_t s, const MgHost& m, int project_mean, CgScalars* sc,
                      float* xg, const MgBatchScale* batch = nullptr);

}  // namespace fsg
