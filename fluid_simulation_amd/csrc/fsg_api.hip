// fsg_api.hip -- the C ABI (include/fsg.h): device state of one FluidDomain +
// FluidSolver pair and the per-step orchestration of the gfx950 kernels.
//
// State carried between steps is exactly the reference's: front and back u, v
// (swapped by index, src/MacGrid.cpp:89-93), prev and diff grids, labels, the
// extension masks and the particles.  Every phase writes the same subset of
// faces as the reference does, so faces the reference leaves stale stay stale.
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <initializer_list>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/fsg.h"
#include "fsg_kernels.h"

using namespace fsg;

namespace {

const char kValidateMsg[] =
    "Memory pool and fluid domain does not match. Initialize the fluid solver with a "
    "memory pool corresponding to the right fluid domain!";

// Per-kernel HIP-event timer on the handle's stream.
struct Timer {
  int on = 0;  // 0 off, 1 every kernel, 2 only names starting with `prefix`
  std::string prefix = "cg_";
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long long>> acc;
  std::vector<std::string> names;  // stable storage for fsg_get_timing
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  void collect() {
    for (auto& p : pending) {
      float ms = 0.f;
      (void)hipEventSynchronize(p.b);
      (void)hipEventElapsedTime(&ms, p.a, p.b);
      auto& slot = acc[p.name];
      slot.first += ms;
      slot.second += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  ~Timer() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

// Packed batch of scenes (fsg_batch_create): scene s occupies the rectangle
// (ox, oy, nx, ny) of the packed grid; everything outside every scene is SOLID.
struct BatchInfo {
  int S = 0;
  std::vector<int> nx, ny, ox, oy;
  std::vector<int64_t> np, part_base;
  std::vector<uint8_t> solid;  // host copy of the packed grid's solid mask
  fsg::BatchScene* d_scene = nullptr;
  int2* d_ptile = nullptr;
  int n_ptile = 0;
  int n_buckets = 0;          // the splat's 16 x 16-cell buckets over all scenes
  int max_scene_buckets = 0;  // the most buckets of one scene
  int4* d_rect = nullptr;
  int* d_sid = nullptr;
  double* d_norm = nullptr;
  double* d_inv = nullptr;
  // multigrid: packed levels 0 .. la, then every scene's own levels (one workgroup each)
  int la = 1;
  std::vector<fsg::SceneCoarse> coarse;
  fsg::SceneCoarse* d_coarse = nullptr;
  uint8_t* d_sinf = nullptr;
  int stride = 0;
  // per-scene CG: 64 x 16 vector tiles inside the scenes, and the scalars (5 S doubles)
  int4* d_vtile = nullptr;
  int2* d_vend = nullptr;  // per vector tile: its scene's column and row ends
  int n_vtile = 0;
  double* d_cg = nullptr;
  uint8_t* d_vflag = nullptr;
  fsg::BatchTables tables() const { return fsg::BatchTables{d_scene, d_ptile, n_ptile}; }
  fsg::MgbArgs cg_args(int W, int H) const {
    fsg::MgbArgs a;
    a.tile = d_vtile;
    a.tend = d_vend;
    a.rect = d_rect;
    a.n_tiles = n_vtile;
    a.n_scenes = S;
    a.W = W;
    a.C = W * H;
    a.rz = d_cg;
    a.pq = d_cg + 2 * S;
    a.rr = d_cg + 3 * S;
    a.bb = d_cg + 4 * S;
    a.vflag = d_vflag;
    return a;
  }
  fsg::MgBatchScale scale() const {
    return fsg::MgBatchScale{S, d_rect, d_sid, d_norm, d_inv};
  }
};

// Reverse mode (fsg_adjoint_*): the adjoint state carried between steps (particles
// and the back buffer), the tape of one step, scratch grids; allocated on first use.
struct AdjState {
  int64_t pcap = 0;
  float4* pbar = nullptr;  // per particle: (x bar, y bar, vx bar, vy bar)
  float4* x0 = nullptr;    // tape: the particles at the step's start
  float *bu = nullptr, *bv = nullptr;  // adjoint of the back buffer
  float *fu = nullptr, *fv = nullptr;  // adjoint of the front buffer
  float *f0u = nullptr, *f0v = nullptr;  // tape: the front at a semi-Lagrangian step's start
  float *su = nullptr, *sv = nullptr, *eu = nullptr, *ev = nullptr, *e2u = nullptr,
        *e2v = nullptr, *g4u = nullptr, *g4v = nullptr, *g1u = nullptr, *g1v = nullptr,
        *qu = nullptr, *qv = nullptr;
  double* y = nullptr;     // adjoint of the pressure, then of the divergence
  uint8_t* zmask = nullptr;  // Dirichlet decisions of the taped pressure update
  std::vector<uint8_t*> rmask;  // extension read masks, one per sweep
  bool grids = false;
  bool taping = false;
  int last_iters = 0;
};

// The multigrid solve's iterations captured as HIP graphs, `iters` per graph
// (fsg_params.graph_iters): graph c runs iterations [c iters, (c + 1) iters) (the last
// one stops at the iteration cap); iters == 0: one graph per solve length (solve, unit
// below).  Their launch arguments depend only on the
// iteration index and on the hierarchy's buffers, fixed once the hierarchy is built,
// and on the tolerance, the cap and the smoothing settings (fsg_set_params drops the
// graphs when those change).
struct MgGraphs {
  int iters = 2;
  std::vector<hipGraphExec_t> exec;
  // iters == 0: solve[L] runs a whole solve of L iterations (through the stop test of
  // iteration L), unit[q] the rest of iteration q and the head of q + 1
  std::vector<hipGraphExec_t> solve, unit;
  void clear() {
    for (std::vector<hipGraphExec_t>* v : {&exec, &solve, &unit}) {
      for (hipGraphExec_t e : *v)
        if (e) (void)hipGraphExecDestroy(e);
      v->clear();
    }
  }
  ~MgGraphs() { clear(); }
};

// A checkpoint of the carried state: particles, front and back velocities.
struct Checkpoint {
  float4* pt = nullptr;
  int64_t np = 0, cap = 0;
  float* u[2] = {nullptr, nullptr};
  float* v[2] = {nullptr, nullptr};
};

struct fsg_domain {
  int device = 0;
  hipStream_t stream = nullptr;
  Geom g{};
  float density = 0.f, pic_ratio = 0.f;
  fsg_params prm{};
  // velocity double buffers + prev/diff
  float* u[2] = {nullptr, nullptr};
  float* v[2] = {nullptr, nullptr};
  int front = 0;
  float *u_prev = nullptr, *v_prev = nullptr, *u_diff = nullptr, *v_diff = nullptr;
  uint8_t* label = nullptr;
  uint8_t* solid = nullptr;  // optional (points at solid_buf when set)
  uint8_t* solid_buf = nullptr;
  float* phi = nullptr;      // optional
  float4* acc = nullptr;
  long long* acc_fx = nullptr;  // fixed-point P2G / density sums (fsg_params.deterministic)
  uint8_t* dec = nullptr;       // the last step's decisions (fsg_params.record_decisions)
  uint8_t* mask[4] = {nullptr, nullptr, nullptr, nullptr};  // u0 u1 v0 v1
  int mask_front = 0;
  uint8_t *info_u = nullptr, *info_v = nullptr;
  // particles
  float4* pt = nullptr;
  int64_t np = 0, cap = 0;
  uint8_t* reject = nullptr;
  // the splat's bucket order (one domain): slot per particle and the particles
  // grouped by bucket (cap each, with pt), particles per bucket and bucket offsets (on
  // first use)
  int* prank = nullptr;
  float4* psorted = nullptr;
  int* bkt_cnt = nullptr;
  int* bkt_off = nullptr;
  // pressure system
  int* counts = nullptr;
  int* offsets = nullptr;
  int* n_fluid = nullptr;
  int* fluid_index = nullptr;
  int* cell_of = nullptr;
  // true: the last pressure solve compacted the system (fluid_index, and b per unknown
  // index); false (multigrid form): b holds the reference's b per cell, and the fluid
  // indices follow from the labels
  bool compacted = false;
  float* b = nullptr;
  uint8_t* info = nullptr;
  float* r = nullptr;
  float* ap = nullptr;
  float* xg = nullptr;
  float* pg0 = nullptr;
  int2* updn = nullptr;
  float* exch = nullptr;
  double* x64 = nullptr;
  double* partials = nullptr;
  unsigned* bar = nullptr;
  CgScalars* sc = nullptr;
  HostScalars* hsc = nullptr;  // pinned host memory: the solve loop's readback
  bool mg_finish_pending = false;  // the next pressure update finishes the MG solve
  int max_coop_blocks = 0;  // co-resident 1024-thread blocks of the persistent solver
  size_t pcg_lds_max = 0;   // dynamic LDS the persistent solver may use per block
  int last_F = 0, last_grid = 0, last_kpt = 0;
  // multigrid form: hierarchy allocated on first use; iterations of the last solve
  MgHost mg{};
  BatchInfo* batch = nullptr;  // packed batch of scenes (fsg_batch_create), else null
  int* pcounts = nullptr;   // particles per cell (particle-pressure term), on first use
  float* phi2 = nullptr;    // second level-set buffer (advection), on first use
  float* emit_buf = nullptr;  // emitter coordinates
  size_t emit_cap = 0;
  AdjState adj;
  std::vector<Checkpoint> ckpt;
  MgGraphs mg_graphs;  // captured multigrid iterations (fsg_params.graph)
  int kclock = 0;  // fsg_set_kernel_clock: 1 k_mg_coarse, 2 level 0's down kernel
  bool mg_ready = false;
  bool mg_solved = false;  // the last solve was the multigrid form (its FP64 x is valid)
  int mg_last_iters = 0;      // iterations of the last forward multigrid solve
  int mg_last_iters_adj = 0;  // and of the last adjoint solve (a different system)
  int mg_prev_iters = 0;      // the same for the solves before those
  int mg_prev_iters_adj = 0;
  // host-side statistics of the plain solver
  fsg_stats host_stats{};
  std::string err;
  Timer timer;
  std::vector<void*> allocations;
};

namespace {

int fail(fsg_domain* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((h), FSG_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t dalloc(fsg_domain* h, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, n * sizeof(T) + 16);
  if (e != hipSuccess) return e;
  h->allocations.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return hipMemsetAsync(q, 0, n * sizeof(T) + 16, h->stream);
}

// Timed launch helper: wraps `body` in events when timing is on.
template <typename F>
void timed(fsg_domain* h, const char* name, F&& body) {
  if (!h->timer.on ||
      (h->timer.on == 2 && strncmp(name, h->timer.prefix.c_str(), h->timer.prefix.size()) != 0)) {
    body();
    return;
  }
  Timer::Pending p{name, h->timer.get(), h->timer.get()};
  (void)hipEventRecord(p.a, h->stream);
  body();
  (void)hipEventRecord(p.b, h->stream);
  h->timer.pending.push_back(p);
  if (h->timer.pending.size() > 4096) h->timer.collect();
}

inline size_t ncell(const fsg_domain* h) { return (size_t)h->g.nx * h->g.ny; }
inline float* UF(fsg_domain* h) { return h->u[h->front]; }
inline float* VF(fsg_domain* h) { return h->v[h->front]; }
inline float* UB(fsg_domain* h) { return h->u[1 - h->front]; }
inline float* VB(fsg_domain* h) { return h->v[1 - h->front]; }
inline void swap_vel(fsg_domain* h) { h->front ^= 1; }
inline uint8_t* MXF(fsg_domain* h) { return h->mask[h->mask_front ? 1 : 0]; }
inline uint8_t* MXB(fsg_domain* h) { return h->mask[h->mask_front ? 0 : 1]; }
inline uint8_t* MYF(fsg_domain* h) { return h->mask[h->mask_front ? 3 : 2]; }
inline uint8_t* MYB(fsg_domain* h) { return h->mask[h->mask_front ? 2 : 3]; }

CgParams make_cg_params(const fsg_domain* h, float tol) {
  CgParams cp{};
  cp.max_iter = h->prm.cg_max_iter;
  cp.tol = tol;
  const double dx = (double)h->g.dx;
  cp.off = (float)(1 / pow(dx, 2));  // src/FluidSolver.cpp:381-402
  for (int n = 0; n <= 4; ++n) {
    cp.diag_tab[n] = (float)(-n / pow(dx, 2));  // :409-410
    cp.invd_tab[n] = cp.diag_tab[n] != 0.f ? 1.0f / cp.diag_tab[n] : 1.0f;
  }
  cp.project_mean = h->prm.project_mean;
  return cp;
}

// ---------------------------------------------------------------- phases
int ph_classify(fsg_domain* h) {
  timed(h, "cell_init", [&] { launch_cell_init(h->stream, h->g, h->label, h->solid, h->acc, 1, 0); });
  timed(h, "particle_scatter", [&] {
    launch_particle_scatter(h->stream, h->g, h->pt, h->np, h->label, h->acc, 1, 0);
  });
  return FSG_OK;
}

// Fused classify + P2G splat (one read of the particles), then the per-face
// normalise [+ prev + gravity + Dirichlet + extension init] on the new front.
void classify_and_splat(fsg_domain* h) {
  // deterministic form: the splat sums in fixed point (zero between steps) and the
  // conversion writes every accumulator, so they are not cleared here
  long long* fx = h->prm.deterministic ? h->acc_fx : nullptr;
  // the splat reads the particles grouped by bucket (their storage order is kept); a
  // batch groups each scene's particles within the scene's own range
  const float4* src = h->pt;
  const bool bucketed = h->prm.splat_order == 1;
  bool single = bucketed && !h->batch && h->np > 0 && h->np < INT_MAX && h->prank &&
                h->psorted && fsg::bucket_count(h->g) <= fsg::kMaxBuckets;
  if (single && !h->bkt_cnt) {
    const size_t nb = (size_t)fsg::bucket_count(h->g);
    if (dalloc(h, &h->bkt_cnt, nb) != hipSuccess || dalloc(h, &h->bkt_off, nb) != hipSuccess)
      h->bkt_cnt = h->bkt_off = nullptr;
  }
  single = single && h->bkt_cnt;
  if (single) {
    // one domain: the bucket count kernel also does the cell set-up
    timed(h, "bucket_order", [&] {
      launch_bucket_order(h->stream, h->g, h->pt, h->np, h->bkt_cnt, h->bkt_off, h->prank,
                          h->psorted, fsg::CellInit{h->label, h->solid, fx ? nullptr : h->acc});
    });
    src = h->psorted;
  } else {
    timed(h, "cell_init", [&] {
      launch_cell_init(h->stream, h->g, h->label, h->solid, h->acc, 1, fx ? 0 : 1);
    });
  }
  if (bucketed && h->batch && h->np > 0 && h->np < INT_MAX && h->prank && h->psorted &&
      h->batch->max_scene_buckets <= fsg::kMaxBuckets) {
    BatchInfo& b = *h->batch;
    if (!h->bkt_cnt) {
      if (dalloc(h, &h->bkt_cnt, (size_t)b.n_buckets) != hipSuccess ||
          dalloc(h, &h->bkt_off, (size_t)b.n_buckets) != hipSuccess)
        h->bkt_cnt = h->bkt_off = nullptr;
    }
    if (h->bkt_cnt) {
      timed(h, "bucket_order", [&] {
        launch_bucket_order_batch(h->stream, h->pt, h->np, h->bkt_cnt, h->bkt_off, h->prank,
                                  h->psorted, b.tables(), b.n_buckets, b.max_scene_buckets);
      });
      src = h->psorted;
    }
  }
  timed(h, "particle_scatter", [&] {
    fsg::BatchTables bt;
    if (h->batch) bt = h->batch->tables();
    launch_particle_scatter(h->stream, h->g, src, h->np, h->label, h->acc, 1, 1,
                            h->batch ? &bt : nullptr, fx);
  });
  if (fx)
    timed(h, "fx_to_float", [&] {
      launch_fx_to_float(h->stream, 4 * (int64_t)ncell(h), fx, reinterpret_cast<float*>(h->acc));
    });
}

void face_update(fsg_domain* h, int flags, float dt, bool record_info) {
  if ((flags & FU_DIRICHLET) && h->prm.closed_walls) flags |= FU_CLOSED;
  const float gx_dt = h->prm.gravity_x * dt;
  const float gy_dt = h->prm.gravity_y * dt;
  timed(h, "face_update", [&] {
    launch_face_update(h->stream, h->g, flags, UF(h), VF(h), h->acc, h->u_prev, h->v_prev, h->label,
                       gx_dt, gy_dt, UB(h), VB(h), MXF(h), MXB(h), MYF(h), MYB(h),
                       record_info ? h->info_u : nullptr, record_info ? h->info_v : nullptr,
                       (flags & FU_DIRICHLET) && h->prm.record_decisions ? h->dec : nullptr);
  });
}

void extend_sweeps(fsg_domain* h, bool record_info) {
  for (int it = 0; it < h->prm.ext_layers; ++it) {
    timed(h, "extend_sweep", [&] {
      launch_extend_sweep(h->stream, h->g, h->label, UB(h), VB(h), MXF(h), MXB(h), MYF(h), MYB(h),
                          record_info ? h->info_u : nullptr, record_info ? h->info_v : nullptr,
                          it + 1);
    });
    h->mask_front ^= 1;  // FluidSolverMemoryPool::swapValidMaskBuffer
  }
  swap_vel(h);
}

// Multigrid hierarchy: halve (rounding up) until both sides are <= 4 cells; the
// levels from the first one of <= kMgCoarseCells cells run in one workgroup.
int mg_prepare(fsg_domain* h) {
  if (h->mg_ready) return FSG_OK;
  MgHost& m = h->mg;
  m.nx[0] = h->g.nx;
  m.ny[0] = h->g.ny;
  m.nlev = 1;
  // a batch: packed levels 0 .. la only, each scene's own levels below (per-scene kernel)
  const int last = h->batch ? h->batch->la : kMgMaxLevels;
  while ((m.nx[m.nlev - 1] > 4 || m.ny[m.nlev - 1] > 4) && m.nlev < kMgMaxLevels &&
         m.nlev <= last) {
    m.nx[m.nlev] = (m.nx[m.nlev - 1] + 1) / 2;
    m.ny[m.nlev] = (m.ny[m.nlev - 1] + 1) / 2;
    ++m.nlev;
  }
  if (h->batch) {
    if (m.nlev != last + 1)
      return fail(h, FSG_E_INVALID, "multigrid solver: packed grid smaller than its split level");
    m.lc = last;
    m.per_scene = true;
  } else {
    // the coarsest level is swept by one wavefront (a cell per lane)
    if (m.nx[m.nlev - 1] * m.ny[m.nlev - 1] > 64)
      return fail(h, FSG_E_INVALID, "multigrid solver: grid too elongated for the hierarchy");
    m.lc = 0;
    while ((size_t)m.nx[m.lc] * m.ny[m.lc] > (size_t)kMgCoarseCells) ++m.lc;
    if (m.lc >= m.nlev)
      return fail(h, FSG_E_INVALID, "multigrid solver: grid too elongated for the hierarchy");
    m.fused = m.lc >= 2;
  }
  // cell offsets of the one-workgroup levels (its LDS layout and its info array)
  MgCoarseArgs& a = m.coarse;
  int off = 0;
  for (int l = 0; l < m.nlev; ++l) {
    a.off[l] = l >= m.lc ? off : 0;
    if (l >= m.lc) off += m.nx[l] * m.ny[l];
  }
  a.total = (off + 3) & ~3;
  bool ok = true;
  uint8_t* coarse_inf = nullptr;
  ok &= dalloc(h, &coarse_inf, (size_t)a.total) == hipSuccess;
  for (int l = 0; l < m.nlev; ++l) {
    const size_t n = (size_t)m.nx[l] * m.ny[l];
    if (l > 0) ok &= dalloc(h, &m.lab[l], n) == hipSuccess;
    if (l < m.lc)
      ok &= dalloc(h, &m.inf[l], n) == hipSuccess;
    else if (coarse_inf)
      m.inf[l] = coarse_inf + a.off[l];
    ok &= dalloc(h, &m.rl[l], n) == hipSuccess;
    ok &= dalloc(h, &m.xs[l], n) == hipSuccess;
    ok &= dalloc(h, &m.xo[l], n) == hipSuccess;
  }
  const size_t C = ncell(h);
  ok &= dalloc(h, &m.b, C) == hipSuccess;
  ok &= dalloc(h, &m.x[0], C) == hipSuccess;
  ok &= dalloc(h, &m.x[1], C) == hipSuccess;
  ok &= dalloc(h, &m.p[0], C) == hipSuccess;
  ok &= dalloc(h, &m.p[1], C) == hipSuccess;
  // the set-up's partials: one per vector block, or per level-0 tile (k_mg_setup0t)
  const size_t n_setup = (size_t)std::max(kMgMaxVecBlocks, mg_up_blocks(m.nx[0], m.ny[0]));
  ok &= dalloc(h, &m.part_sum, n_setup) == hipSuccess;
  ok &= dalloc(h, &m.part_cnt, n_setup) == hipSuccess;
  ok &= dalloc(h, &m.part_liq, n_setup) == hipSuccess;
  ok &= dalloc(h, &m.part_bb, kMgMaxVecBlocks) == hipSuccess;
  ok &= dalloc(h, &m.part_rr, (size_t)mg_part_rr_size(m)) == hipSuccess;
  // k_mg_dir: one partial per vector block, or per level-0 tile where tiles are skipped
  ok &= dalloc(h, &m.part_pq, (size_t)std::max(kMgMaxVecBlocks, mg_up_blocks(m.nx[0], m.ny[0]))) ==
        hipSuccess;
  ok &= dalloc(h, &m.part_rz, (size_t)std::max(mg_up_blocks(m.nx[0], m.ny[0]), 1)) == hipSuccess;
  ok &= dalloc(h, &m.clk, 2) == hipSuccess;
  if (m.lc > 0) ok &= dalloc(h, &m.tflag, (size_t)mg_up_blocks(m.nx[0], m.ny[0])) == hipSuccess;
  if (m.lc > 0 && !h->batch)
    ok &= dalloc(h, &m.tmap_buf, (size_t)mg_up_blocks(m.nx[0], m.ny[0])) == hipSuccess;
  if (m.per_scene)
    for (int l = 1; l < m.lc; ++l)
      ok &= dalloc(h, &m.tflag_l[l], (size_t)mg_up_blocks(m.nx[l], m.ny[l])) == hipSuccess;
  if (!m.per_scene && m.lc >= 2) {
    ok &= dalloc(h, &m.ticket, 1) == hipSuccess;
    if (ok) ok &= hipMemsetAsync(m.ticket, 0, sizeof(unsigned), h->stream) == hipSuccess;
  }
  if (!ok) return fail(h, FSG_E_NOMEM, "multigrid solver: device allocation failed");
  m.lab[0] = h->label;
  a.lc = m.lc;
  a.nlev = m.nlev;
  a.omega = a.omega2 = 0.8f;  // set per solve from fsg_params.mg_smoother
  a.sweeps = 2 * std::max(m.nx[m.nlev - 1], m.ny[m.nlev - 1]) + 8;
  for (int l = 0; l < m.nlev; ++l) {
    a.nx[l] = m.nx[l];
    a.ny[l] = m.ny[l];
  }
  a.inf_all = coarse_inf;
  a.r_in = m.rl[m.lc];
  a.x_out = m.xo[m.lc];
  // The one-workgroup levels hold at most ~2 x kMgCoarseCells cells (13 bytes each);
  // the kernel's LDS limit is raised once, process-wide, to a fixed 128 KiB.
  constexpr size_t kCoarseLdsLimit = 128 * 1024;
  m.coarse_lds = (size_t)a.total * (3 * sizeof(float) + 1);
  if (h->batch) m.coarse_lds = (size_t)h->batch->stride * (3 * sizeof(float) + 1);
  if (m.coarse_lds > kCoarseLdsLimit)
    return fail(h, FSG_E_INVALID, "multigrid solver: coarse levels exceed the LDS budget");
  if (mg_set_lds_limit(kCoarseLdsLimit) != hipSuccess)
    return fail(h, FSG_E_HIP, "multigrid solver: cannot raise the LDS limit of the coarse kernel");
  h->mg_ready = true;
  return FSG_OK;
}

fsg::SceneCoarseArgs scene_coarse_args(const fsg_domain* h) {
  const BatchInfo& b = *h->batch;
  const MgHost& m = h->mg;
  fsg::SceneCoarseArgs a;
  a.tab = b.d_coarse;
  a.n_scenes = b.S;
  a.pitch = m.nx[b.la];
  a.stride = b.stride;
  a.inf_all = b.d_sinf;
  a.lab = m.lab[b.la];
  a.r_in = m.rl[b.la];
  a.x_out = m.xo[b.la];
  a.omega = m.coarse.omega;
  a.omega2 = m.coarse.omega2;
  a.trunc = m.coarse.trunc;
  return a;
}

// CG preconditioned by one V-cycle per iteration (FSG_SOLVER_MG).  Iterations are
// enqueued in batches (the first sized by the previous solve's count); every
// kernel returns at once after the device-side stop test has fired.
// Room for kSpanLaunches launches of level 0's down kernel in the block-span clock
// (held until destroy).
void alloc_clk_span(fsg_domain* h) {
  constexpr int kSpanLaunches = 4096;
  MgHost& m = h->mg;
  m.clk_span_nb = mg_down_blocks0(m);
  if (m.clk_span_nb > 0 &&
      dalloc(h, &m.clk_span, 2 * (size_t)m.clk_span_nb * kSpanLaunches) == hipSuccess)
    m.clk_span_cap = kSpanLaunches;
}

// F < 0: the number of unknowns is still on the device (h->n_fluid); it is read back
// with the stop-test copy, so the solve costs no extra host round trip.
// defer_finish: the caller's pressure update does the finish (MgFinishFuse) where it can
// (one domain, no mean projection); h->mg_finish_pending says whether it must.
int mg_solve(fsg_domain* h, int F, const double* adj_rhs = nullptr, bool defer_finish = false) {
  h->mg_finish_pending = false;
  if (int rc = mg_prepare(h)) return rc;
  MgHost& m = h->mg;
  // smoothing weights: damped Jacobi (0.8, 0.8) or the two-step Chebyshev smoother on
  // [0.4, 2]: 1 / (1.2 + 0.8 cos(pi / 4)), 1 / (1.2 - 0.8 cos(pi / 4))
  if (h->prm.mg_smoother == 1) {
    m.coarse.omega = (float)(1.0 / (1.2 + 0.8 * 0.70710678118654752));
    m.coarse.omega2 = (float)(1.0 / (1.2 - 0.8 * 0.70710678118654752));
  } else {
    m.coarse.omega = m.coarse.omega2 = 0.8f;
  }
  m.coarse.trunc = h->prm.mg_coarse_sweeps;
  m.fused = !h->batch && m.lc >= 2 && h->prm.mg_fuse_update != 0;
  m.l0_threads = h->prm.mg_tile_threads;
  m.level_sweeps = h->prm.mg_level_sweeps;
  m.level0_sweeps = h->prm.mg_level0_sweeps;
  m.tmap = h->prm.mg_spread_tiles ? m.tmap_buf : nullptr;
  m.clk_on = h->kclock == 1;
  m.clk_span_on = h->kclock == 2;
  if (m.clk_span_on && !m.clk_span) alloc_clk_span(h);
  timed(h, "cg_mg_setup", [&] {
    fsg::MgBatchScale bs;
    if (h->batch) bs = h->batch->scale();
    launch_mg_setup(h->stream, h->g, m, h->label, UF(h), VF(h), h->prm.project_mean,
                    h->prm.warm_start != 0 && !h->batch && !adj_rhs, h->sc,
                    h->batch ? &bs : nullptr, adj_rhs, adj_rhs ? nullptr : h->b,
                    adj_rhs ? nullptr : h->n_fluid);
    if (h->batch) {
      launch_mg_scene_setup(h->stream, scene_coarse_args(h));
      launch_mgb_init(h->stream, h->batch->cg_args(h->g.nx, h->g.ny), h->batch->d_norm, m);
    }
  });
  const int max_iter = h->prm.cg_max_iter;
  // a batch runs one CG per scene in lock step, each stopping at |r_s| < tol |b'_s|
  const double tol = h->prm.cg_tol;
  // iterations are enqueued in batches between stop-test reads: first the previous
  // solve's count (of the same kind) + 1, then doubling batches
  int& last_iters = adj_rhs ? h->mg_last_iters_adj : h->mg_last_iters;
  int& prev_iters = adj_rhs ? h->mg_prev_iters_adj : h->mg_prev_iters;
  int it = 0, batch = std::max(last_iters + 1, 4), more = 2;
  CgScalars hs{};
  // one timer name per level (the last tile level's down kernel also runs the
  // one-workgroup stage)
  static const char* const kDownNames[8] = {"cg_mg_down0", "cg_mg_down1", "cg_mg_down2",
                                            "cg_mg_down3", "cg_mg_down4", "cg_mg_down5",
                                            "cg_mg_down6", "cg_mg_down7+"};
  static const char* const kUpNames[8] = {"cg_mg_up0", "cg_mg_up1", "cg_mg_up2", "cg_mg_up3",
                                          "cg_mg_up4", "cg_mg_up5", "cg_mg_up6", "cg_mg_up7+"};
  // one CG iteration: the V-cycle's kernels, the direction (and the x / residual
  // update where it is not fused into the next iteration's level-0 kernel).  Its head
  // is its kernels up to and including the stop test (a batch's check kernel; level 0's
  // down kernel, or levels 0-1 in the fused form; the whole iteration where the
  // hierarchy has no tile level), its tail the rest: once the solve has stopped in
  // iteration q's head, the kernels after it return at once.
  const int head_levels = h->batch ? 0 : m.fused ? 2 : 1;
  const bool whole_head = !h->batch && m.lc < head_levels;
  auto enqueue_part = [&](int it, bool head) {
    if (h->batch && head)
      timed(h, "cg_mg_check", [&] {
        launch_mgb_check(h->stream, h->batch->cg_args(h->g.nx, h->g.ny), it, max_iter, tol, h->sc);
      });
    if (!whole_head)
      for (int l = 0; l < m.lc; ++l)
        if ((l < head_levels) == head)
          timed(h, kDownNames[std::min(l, 7)],
                [&] { launch_mg_down(h->stream, m, l, it, max_iter, tol, h->sc); });
    if (head != whole_head) return;
    if (whole_head)
      for (int l = 0; l < m.lc; ++l)
        timed(h, kDownNames[std::min(l, 7)],
              [&] { launch_mg_down(h->stream, m, l, it, max_iter, tol, h->sc); });
    if (h->batch)
      timed(h, "cg_mg_coarse", [&] { launch_mg_scene_coarse(h->stream, scene_coarse_args(h), h->sc); });
    else if (!mg_tail_coarse(m))
      timed(h, "cg_mg_coarse", [&] { launch_mg_coarse(h->stream, m, it, max_iter, tol, h->sc); });
    for (int l = m.lc - 1; l >= 0; --l)
      timed(h, kUpNames[std::min(l, 7)], [&] { launch_mg_up(h->stream, m, l, it, h->sc); });
    if (h->batch) {
      const fsg::MgbArgs ba = h->batch->cg_args(h->g.nx, h->g.ny);
      timed(h, "cg_mg_rz", [&] { launch_mgb_vectors(h->stream, ba, m, it, h->sc, 0); });
      timed(h, "cg_mg_dir", [&] { launch_mgb_vectors(h->stream, ba, m, it, h->sc, 1); });
      timed(h, "cg_mg_xr", [&] { launch_mgb_vectors(h->stream, ba, m, it, h->sc, 2); });
    } else {
      timed(h, "cg_mg_dir", [&] { launch_mg_dir(h->stream, m, it, h->sc); });
      // fused form: the update runs in the next iteration's level-0 down kernel
      if (!m.fused) timed(h, "cg_mg_xr", [&] { launch_mg_xr(h->stream, m, it, h->sc); });
    }
  };
  auto enqueue_iteration = [&](int it) {
    enqueue_part(it, true);
    enqueue_part(it, false);
  };
  // Graph replay: graph c holds iterations [c K, (c + 1) K) (up to the cap), captured
  // on its first use.  Per-kernel timing and the kernel clocks need every launch.
  const bool use_graph = h->prm.graph && !h->timer.on && !m.clk_on && !m.clk_span_on;
  if (h->mg_graphs.iters != h->prm.graph_iters) {
    h->mg_graphs.clear();
    h->mg_graphs.iters = h->prm.graph_iters;
  }
  const int K = std::max(1, h->mg_graphs.iters);
  // captures what body enqueues into *slot once
  auto capture = [&](hipGraphExec_t* slot, auto&& body) -> int {
    if (*slot) return FSG_OK;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    body();
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    const hipError_t e = hipGraphInstantiate(slot, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      *slot = nullptr;
      HIPCHK(h, e);
    }
    return FSG_OK;
  };
  auto graph_of = [&](int c, hipGraphExec_t* out) -> int {
    std::vector<hipGraphExec_t>& ex = h->mg_graphs.exec;
    if ((int)ex.size() <= c) ex.resize(c + 1, nullptr);
    if (int rc = capture(&ex[c], [&] {
          for (int q = c * K; q < std::min((c + 1) * K, max_iter + 1); ++q) enqueue_iteration(q);
        }))
      return rc;
    *out = ex[c];
    return FSG_OK;
  };
  auto read_back = [&]() -> int {
    if (!h->hsc)
      HIPCHK(h, hipHostMalloc((void**)&h->hsc, sizeof(HostScalars), hipHostMallocCoherent));
    launch_read_scalars(h->stream, h->sc, F < 0 ? h->n_fluid : nullptr, h->hsc);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    memcpy(&hs, &h->hsc->sc, sizeof(hs));
    if (F < 0) F = h->hsc->n_fluid;
    return FSG_OK;
  };
  if (use_graph && h->mg_graphs.iters == 0) {
    // one graph for a solve as long as the longer of the previous two (L iterations,
    // through the stop test of iteration L), then one graph per further iteration until
    // the test passes (a further graph costs a host round trip, a shorter solve than L
    // the kernels of its remaining iterations, which return at once)
    const int L = std::min(last_iters > 0 ? std::max(last_iters, prev_iters) : 4, max_iter);
    std::vector<hipGraphExec_t>& sv = h->mg_graphs.solve;
    if ((int)sv.size() <= L) sv.resize(L + 1, nullptr);
    if (int rc = capture(&sv[L], [&] {
          enqueue_part(0, true);
          for (int q = 0; q < L; ++q) {
            enqueue_part(q, false);
            enqueue_part(q + 1, true);
          }
        }))
      return rc;
    HIPCHK(h, hipGraphLaunch(sv[L], h->stream));
    it = L;
    for (;;) {
      if (int rc = read_back()) return rc;
      if (hs.done || it >= max_iter) break;
      std::vector<hipGraphExec_t>& uv = h->mg_graphs.unit;
      if ((int)uv.size() <= it) uv.resize(it + 1, nullptr);
      const int q = it;
      if (int rc = capture(&uv[q], [&] {
            enqueue_part(q, false);
            enqueue_part(q + 1, true);
          }))
        return rc;
      HIPCHK(h, hipGraphLaunch(uv[q], h->stream));
      ++it;
    }
  }
  while (!(use_graph && h->mg_graphs.iters == 0)) {
    const int target = std::min(it + batch, max_iter + 1);
    if (use_graph) {
      while (it < target) {
        const int c = it / K;
        hipGraphExec_t e = nullptr;
        if (int rc = graph_of(c, &e)) return rc;
        HIPCHK(h, hipGraphLaunch(e, h->stream));
        it = std::min((c + 1) * K, max_iter + 1);
      }
    } else {
      for (; it < target; ++it) enqueue_iteration(it);
    }
    if (int rc = read_back()) return rc;
    if (hs.done || it > max_iter) break;
    batch = more;
    more = std::min(2 * more, 32);
  }
  h->last_F = F;
  h->host_stats.n_fluid = F;
  if (defer_finish && !adj_rhs && !h->batch && !h->prm.project_mean)
    h->mg_finish_pending = true;
  else
    timed(h, "cg_mg_finish",
          [&] {
            fsg::MgBatchScale bs;
            if (h->batch) bs = h->batch->scale();
            launch_mg_finish(h->stream, m, h->prm.project_mean, h->sc, h->xg,
                             h->batch ? &bs : nullptr);
          });
  prev_iters = last_iters;
  last_iters = hs.iter;
  h->mg_solved = true;
  h->host_stats.cg_iterations = hs.iter;
  h->host_stats.cg_error = hs.err;
  h->host_stats.mg_live_tiles = hs.live_tiles;
  h->host_stats.total_unknown_iterations += (int64_t)F * hs.iter;
  h->host_stats.total_solves += 1;
  return FSG_OK;
}

int solve_pressure(fsg_domain* h) {
  h->mg_solved = false;
  h->host_stats.cg_iterations = 0;
  h->host_stats.cg_error = 0.0;
  h->host_stats.refine_sweeps = 0;
  const int solver = h->prm.solver;
  // the multigrid form's launches do not depend on the number of unknowns (a system
  // without any stops at its first test, and the pressure update reads the count on
  // the device); it reads the count back with its stop test.  Its set-up forms the
  // LIQUID count and the reference's b per cell itself: no compaction.
  h->compacted = solver != FSG_SOLVER_MG;
  if (solver == FSG_SOLVER_MG) return mg_solve(h, -1, nullptr, true);
  const int nb = (int)((ncell(h) + kScanChunk - 1) / kScanChunk);
  timed(h, "count_liquid", [&] { launch_count_liquid(h->stream, h->g, h->label, h->counts); });
  timed(h, "scan_counts", [&] { launch_scan_counts(h->stream, nb, h->counts, h->offsets, h->n_fluid); });
  timed(h, "compact", [&] {
    launch_compact(h->stream, h->g, h->label, h->offsets, h->fluid_index, h->cell_of, UF(h), VF(h),
                   h->b, h->info);
  });
  // the launch geometry of the other forms depends on the number of unknowns
  int F = 0;
  HIPCHK(h, hipMemcpyAsync(&F, h->n_fluid, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->host_stats.n_fluid = F;
  h->last_F = F;
  if (F == 0) return FSG_OK;
  if (solver == FSG_SOLVER_PLAIN) {
    const CgParams cp = make_cg_params(h, (float)h->prm.cg_tol);
    timed(h, "cg_init", [&] {
      launch_cg_init(h->stream, F, cp, h->cell_of, h->info, h->b, h->r, h->pg0, h->xg, h->partials,
                     h->sc);
    });
    const int check = h->prm.check_every > 0 ? h->prm.check_every : 32;
    int launched = 0;
    CgScalars hs{};
    while (launched < cp.max_iter) {
      const int n = std::min(check, cp.max_iter - launched);
      for (int q = 0; q < n; ++q) {
        const int it = launched + q;
        timed(h, "cg_iteration", [&] {
          launch_cg_iteration(h->stream, F, h->g.nx, it, cp, h->cell_of, h->info, h->r, h->ap,
                              h->pg0, h->xg, h->partials, h->sc);
        });
      }
      launched += n;
      HIPCHK(h, hipMemcpyAsync(&hs, h->sc, sizeof(hs), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      if (hs.done) break;
    }
    if (cp.max_iter == 0) {
      HIPCHK(h, hipMemcpyAsync(&hs, h->sc, sizeof(hs), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    h->host_stats.cg_iterations = hs.iter;
    h->host_stats.cg_error = hs.err;
    h->host_stats.total_unknown_iterations += (int64_t)F * hs.spmv;
    h->host_stats.total_solves += 1;
    return FSG_OK;
  }
  // persistent / refined: one cooperative launch, every iteration on the device
  // Launch shape: G blocks of 1024 threads, KPT unknowns per thread.  Fewer blocks
  // make the grid barrier cheaper (measured ~1 us at 32 blocks, ~3.6 us at 256);
  // more unknowns per thread cost registers and LDS (a window of per + 2 nx values
  // of three arrays).  The default takes the largest KPT whose LDS window fits;
  // fsg_params.pcg_kpt / pcg_blocks override.
  int kpt = 0, grid = 0;
  const int gmax = h->max_coop_blocks;
  const int kfix = h->prm.pcg_kpt;
  // spill-free register budgets at 1024 threads: KPT <= 4 (FP32 form), KPT <= 2
  // (refined), the only instances built; measured fastest at 1024^2: KPT 2 for both
  const bool refined = solver == FSG_SOLVER_REFINED;
  const std::initializer_list<int> order_plain = {2, 4, 1}, order_ref = {2, 1};
  for (int k : refined ? order_ref : order_plain) {
    if (kfix > 0 && k != kfix) continue;
    const int gneed = (F + 1024 * k - 1) / (1024 * k);
    if (gneed > gmax) continue;
    int g = gneed;
    if (h->prm.pcg_blocks > 0) g = std::min(std::max(gneed, h->prm.pcg_blocks), gmax);
    if (pcg_lds_bytes(F, g, h->g.nx) > h->pcg_lds_max) continue;
    kpt = k;
    grid = g;
    break;
  }
  if (kpt == 0)
    return fail(h, FSG_E_INVALID,
                "pressure system too large for the register-resident persistent solver "
                "(use FSG_SOLVER_PLAIN)");
  timed(h, "updown", [&] {
    launch_updown(h->stream, h->g.nx, F, h->n_fluid, h->cell_of, h->info, h->fluid_index, h->updn);
  });
  PersistentArgs a{};
  a.nx = h->g.nx;
  a.F_max = (int)ncell(h);
  a.F_actual = F;
  a.n_fluid = h->n_fluid;
  a.cell_of = h->cell_of;
  a.info = h->info;
  a.updn = h->updn;
  a.b = h->b;
  a.xg = h->xg;
  a.exch = h->exch;
  a.x64 = h->x64;
  a.partials = h->partials;
  a.bar = h->bar;
  a.sc = h->sc;
  a.refined = solver == FSG_SOLVER_REFINED;
  a.cp = make_cg_params(h, a.refined ? (float)h->prm.refine_inner_tol : (float)h->prm.cg_tol);
  a.max_outer = h->prm.refine_max_outer;
  a.inner_tol = (float)h->prm.refine_inner_tol;
  a.outer_tol = h->prm.cg_tol;
  a.warm_start = h->prm.warm_start != 0;
  timed(h, "zero_barrier", [&] { launch_fill(h->stream, reinterpret_cast<float*>(h->bar), 4, 0.f); });
  hipError_t e = hipSuccess;
  timed(h, a.refined ? "cg_refined" : "cg_persistent", [&] { e = launch_pcg(h->stream, a, grid, kpt); });
  HIPCHK(h, e);
  h->last_grid = grid;
  h->last_kpt = kpt;
  return FSG_OK;
}

int ph_pressure(fsg_domain* h, float dt, bool dirichlet, bool diff) {
  int rc = solve_pressure(h);
  if (rc != FSG_OK) return rc;
  // particle-count pressure term (the reference multiplies it by k = 0)
  const float ppk = h->prm.particle_pressure_k;
  int* counts = nullptr;
  if (ppk != 0.f) {
    if (!h->pcounts && dalloc(h, &h->pcounts, ncell(h)) != hipSuccess)
      return fail(h, FSG_E_NOMEM, "particle counts: device allocation failed");
    counts = h->pcounts;
    timed(h, "count_particles",
          [&] { launch_count_particles(h->stream, h->g, h->pt, h->np, counts); });
  }
  const float dt_over_density = dt / h->density;
  fsg::MgFinishFuse fin;
  if (h->mg_finish_pending) fin = fsg::MgFinishFuse{h->mg.inf[0], h->mg.x[1], h->sc};
  h->mg_finish_pending = false;
  timed(h, "pressure_update", [&] {
    launch_pressure_update(h->stream, h->g, h->n_fluid, h->label, h->xg,
                           h->mg_solved ? h->mg.x[0] : nullptr, counts, ppk, UF(h), VF(h), UB(h),
                           VB(h),
                           dt_over_density, dirichlet ? (h->prm.closed_walls ? 2 : 1) : 0,
                           diff ? 1 : 0, h->u_prev, h->v_prev,
                           h->u_diff, h->v_diff, h->adj.taping ? h->adj.zmask : nullptr,
                           dirichlet && h->prm.record_decisions ? h->dec : nullptr, fin);
  });
  swap_vel(h);
  return FSG_OK;
}

int check_validate(fsg_domain* h) {
  // FluidSolver::validate against the pool the solver copied: its copy constructor
  // passes deltaX twice (src/FluidSolver.cpp:56-63), so the check reduces to
  // |dx - dx| < 1e-7 and |dy - dx| < 1e-7 (in float, compared with a double).
  float d = h->g.dy - h->g.dx;
  if (d < 0) d = -d;
  if (!((double)d < 0.0000001)) return fail(h, FSG_E_VALIDATE, kValidateMsg);
  return FSG_OK;
}

int step_flip_family(fsg_domain* h, int scheme, float dt) {
  const bool flip = scheme != FSG_SCHEME_PIC;
  const bool rec = h->prm.record_ext_info != 0;
  classify_and_splat(h);
  // P2G writes the back buffer, then swaps (src/FluidSolver.cpp:898-918)
  swap_vel(h);
  int flags = FU_P2G | FU_GRAVITY | FU_DIRICHLET | FU_EXTINIT;
  if (flip || h->adj.taping) flags |= FU_PREV;  // the adjoint reads the P2G grid from prev
  // the P2G normalise must read the stale values of the buffer it writes: after the
  // swap that buffer is the front, and the extension init writes the new back.
  face_update(h, flags, dt, rec);
  extend_sweeps(h, rec);
  int rc = ph_pressure(h, dt, /*dirichlet=*/true, /*diff=*/flip);
  if (rc != FSG_OK) return rc;
  const int g2p = scheme == FSG_SCHEME_PIC ? 1 : scheme == FSG_SCHEME_FLIP ? 2 : 3;
  const int move = scheme == FSG_SCHEME_PICFLIP ? 2 : 1;
  timed(h, "particle_gather", [&] {
    fsg::BatchTables bt;
    if (h->batch) bt = h->batch->tables();
    launch_particle_gather(h->stream, h->g, h->pt, h->np, UF(h), VF(h), h->u_diff, h->v_diff,
                           h->pic_ratio, g2p, move, dt, h->label, h->reject,
                           h->batch ? &bt : nullptr);
  });
  return FSG_OK;
}

// advectVelocitySemiLagrangian (src/FluidSolver.cpp:709-772): back := 0, then the
// scatter (fixed-point sums written whole to the back buffer when deterministic)
void ph_sl_scatter(fsg_domain* h, float dt) {
  long long* fx = h->prm.deterministic ? h->acc_fx : nullptr;
  if (!fx)
    timed(h, "fill", [&] {
      launch_fill(h->stream, UB(h), (int64_t)ncell(h), 0.f);
      launch_fill(h->stream, VB(h), (int64_t)ncell(h), 0.f);
    });
  timed(h, "sl_scatter", [&] {
    launch_sl_scatter(h->stream, h->g, h->label, UF(h), VF(h), UB(h), VB(h), h->prm.integrator, dt,
                      fx);
  });
}

int step_semi_lagrangian(fsg_domain* h, float dt) {
  int rc = ph_classify(h);
  if (rc) return rc;
  ph_sl_scatter(h, dt);
  face_update(h, FU_GRAVITY | FU_DIRICHLET, dt, false);
  rc = ph_pressure(h, dt, true, false);
  if (rc) return rc;
  timed(h, "advect_grid", [&] {
    launch_advect_grid(h->stream, h->g, h->pt, h->np, UF(h), VF(h), h->prm.integrator, dt);
  });
  return FSG_OK;
}

int step_grid_advect(fsg_domain* h, float dt) {
  if (!h->phi) return fail(h, FSG_E_INVALID, "grid-only scheme needs a level set (fsg_set_level_set)");
  timed(h, "classify_levelset",
        [&] { launch_classify_levelset(h->stream, h->g, h->phi, h->solid, h->label); });
  timed(h, "sl_gather", [&] {
    launch_sl_gather(h->stream, h->g, h->label, UF(h), VF(h), UB(h), VB(h), h->prm.integrator, dt);
  });
  swap_vel(h);
  face_update(h, FU_GRAVITY | FU_DIRICHLET, dt, false);
  return ph_pressure(h, dt, true, false);
}

int check_grid_ptr(fsg_domain* h, const void* p) {
  if (!p) return fail(h, FSG_E_INVALID, "null array");
  return FSG_OK;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* fsg_version(void) { return "fsg 0.1 (gfx950)"; }

void fsg_params_default(fsg_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->gravity_x = 0.0f;
  p->gravity_y = (float)-9.82;
  p->cg_tol = (double)FLT_EPSILON;
  p->cg_max_iter = 100;
  p->ext_layers = 2;
  p->integrator = FSG_RK3;
  p->solver = FSG_SOLVER_PLAIN;
  p->project_mean = 0;
  p->check_every = 32;
  p->refine_inner_tol = 1e-3;
  p->refine_max_outer = 20;
  p->record_ext_info = 0;
  p->graph = 1;
  p->mg_smoother = 1;
  p->mg_coarse_sweeps = 0;
  p->graph_iters = 0;
  p->mg_spread_tiles = 1;
  p->mg_fuse_update = 1;
  p->splat_order = 1;
  p->mg_tile_threads = 512;
  p->mg_level_sweeps = 5;
  p->mg_level0_sweeps = 3;
}

int fsg_create(fsg_domain** out, int32_t nx, int32_t ny, float length_x, float length_y,
               float density, float pic_ratio) {
  if (!out) return FSG_E_INVALID;
  *out = nullptr;
  if (nx < 3 || ny < 3 || (int64_t)nx * ny > (int64_t)1 << 30) return FSG_E_INVALID;
  if (!(length_x > 0.f) || !(length_y > 0.f)) return FSG_E_INVALID;
  fsg_domain* h = new (std::nothrow) fsg_domain();
  if (!h) return FSG_E_NOMEM;
  auto bail = [&](int code) {
    fsg_destroy(h);
    return code;
  };
  if (hipGetDevice(&h->device) != hipSuccess) return bail(FSG_E_HIP);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bail(FSG_E_HIP);
  fsg_params_default(&h->prm);
  h->g = make_geom(nx, ny, length_x, length_y);
  h->density = density;
  h->pic_ratio = pic_ratio;     // the constructor stores it unclamped (src/FluidDomain.cpp:54-68)
  const size_t C = (size_t)nx * ny;
  const int nb = (int)((C + kScanChunk - 1) / kScanChunk);
  bool ok = true;
  ok &= dalloc(h, &h->u[0], C) == hipSuccess;
  ok &= dalloc(h, &h->u[1], C) == hipSuccess;
  ok &= dalloc(h, &h->v[0], C) == hipSuccess;
  ok &= dalloc(h, &h->v[1], C) == hipSuccess;
  ok &= dalloc(h, &h->u_prev, C) == hipSuccess;
  ok &= dalloc(h, &h->v_prev, C) == hipSuccess;
  ok &= dalloc(h, &h->u_diff, C) == hipSuccess;
  ok &= dalloc(h, &h->v_diff, C) == hipSuccess;
  ok &= dalloc(h, &h->label, C) == hipSuccess;
  ok &= dalloc(h, &h->acc, C) == hipSuccess;
  for (int q = 0; q < 4; ++q) ok &= dalloc(h, &h->mask[q], C) == hipSuccess;
  ok &= dalloc(h, &h->info_u, C) == hipSuccess;
  ok &= dalloc(h, &h->info_v, C) == hipSuccess;
  ok &= dalloc(h, &h->counts, (size_t)nb) == hipSuccess;
  ok &= dalloc(h, &h->offsets, (size_t)nb) == hipSuccess;
  ok &= dalloc(h, &h->n_fluid, 4) == hipSuccess;
  ok &= dalloc(h, &h->fluid_index, C) == hipSuccess;
  ok &= dalloc(h, &h->cell_of, C) == hipSuccess;
  ok &= dalloc(h, &h->b, C) == hipSuccess;
  ok &= dalloc(h, &h->info, C) == hipSuccess;
  ok &= dalloc(h, &h->r, C) == hipSuccess;
  ok &= dalloc(h, &h->ap, C) == hipSuccess;
  ok &= dalloc(h, &h->xg, C) == hipSuccess;
  ok &= dalloc(h, &h->pg0, C) == hipSuccess;
  ok &= dalloc(h, &h->updn, C) == hipSuccess;
  ok &= dalloc(h, &h->exch, 3 * C) == hipSuccess;
  ok &= dalloc(h, &h->x64, C) == hipSuccess;
  ok &= dalloc(h, &h->partials, (size_t)4 * 4096) == hipSuccess;
  ok &= dalloc(h, &h->bar, 16) == hipSuccess;
  ok &= dalloc(h, &h->sc, 1) == hipSuccess;
  if (!ok) return bail(FSG_E_NOMEM);
  // labels start as clearCellTypeBuffer leaves them (MacGrid ctor, src/MacGrid.cpp:24)
  launch_cell_init(h->stream, h->g, h->label, nullptr, h->acc, 1, 0);
  {
    // the persistent solver's LDS window: all of a CU's LDS but the static part
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) {
      size_t lds = prop.maxSharedMemoryPerMultiProcessor;
      if (lds == 0 || lds > 160 * 1024) lds = 160 * 1024;
      h->pcg_lds_max = lds - 4096;
      if (pcg_set_lds_limit(h->pcg_lds_max) == hipSuccess)
        h->max_coop_blocks = std::min(prop.multiProcessorCount, 1024);
    }
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(FSG_E_HIP);
  *out = h;
  return FSG_OK;
}

int fsg_destroy(fsg_domain* h) {
  if (!h) return FSG_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->mg_graphs.clear();
  delete h->batch;  // its device arrays are in h->allocations
  h->timer.collect();
  for (void* p : h->allocations) (void)hipFree(p);
  if (h->pt) (void)hipFree(h->pt);
  if (h->emit_buf) (void)hipFree(h->emit_buf);
  if (h->reject) (void)hipFree(h->reject);
  if (h->prank) (void)hipFree(h->prank);
  if (h->psorted) (void)hipFree(h->psorted);
  if (h->hsc) (void)hipHostFree(h->hsc);
  if (h->adj.pbar) (void)hipFree(h->adj.pbar);
  if (h->adj.x0) (void)hipFree(h->adj.x0);
  for (Checkpoint& c : h->ckpt)
    if (c.pt) (void)hipFree(c.pt);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FSG_OK;
}

const char* fsg_last_error(const fsg_domain* h) { return h ? h->err.c_str() : "null handle"; }

int fsg_set_params(fsg_domain* h, const fsg_params* p) {
  if (!h || !p) return FSG_E_INVALID;
  if (p->cg_max_iter < 0 || p->ext_layers < 0 || p->ext_layers > 14 || p->integrator < 0 ||
      p->integrator > 3 || p->solver < 0 || p->solver > 3 || !(p->cg_tol >= 0.0) ||
      !std::isfinite(p->particle_pressure_k) || p->closed_walls < 0 || p->closed_walls > 1 ||
      p->graph < 0 || p->graph > 1 || p->deterministic < 0 || p->deterministic > 1 ||
      p->record_decisions < 0 || p->record_decisions > 1 || p->mg_smoother < 0 ||
      p->mg_smoother > 1 || p->mg_coarse_sweeps < 0 || p->mg_coarse_sweeps > 64 ||
      p->graph_iters < 0 || p->graph_iters > 16 || p->mg_fuse_update < 0 ||
      p->mg_fuse_update > 1 || p->splat_order < 0 || p->splat_order > 1 ||
      (p->mg_tile_threads != 256 && p->mg_tile_threads != 512 && p->mg_tile_threads != 1024) ||
      p->mg_level_sweeps < 2 || p->mg_level_sweeps > 6 || p->mg_level0_sweeps < 2 ||
      p->mg_level0_sweeps > 3 || p->mg_spread_tiles < 0 || p->mg_spread_tiles > 1)
    return fail(h, FSG_E_INVALID, "fsg_set_params: value out of range");
  const size_t C = ncell(h);
  // fixed-point sums: 4 per cell (the P2G accumulators; the density uses the first C),
  // zero between uses
  if (p->deterministic && !h->acc_fx) {
    long long* fx = nullptr;
    if (dalloc(h, &fx, 4 * C) != hipSuccess)
      return fail(h, FSG_E_NOMEM, "fsg_set_params: deterministic sums: allocation failed");
    HIPCHK(h, hipMemsetAsync(fx, 0, 4 * C * sizeof(long long), h->stream));
    h->acc_fx = fx;
  }
  if (p->record_decisions && !h->dec) {
    uint8_t* d = nullptr;
    if (dalloc(h, &d, C) != hipSuccess)
      return fail(h, FSG_E_NOMEM, "fsg_set_params: decision record: allocation failed");
    HIPCHK(h, hipMemsetAsync(d, 0, C, h->stream));
    h->dec = d;
  }
  if ((p->solver == FSG_SOLVER_PERSISTENT || p->solver == FSG_SOLVER_REFINED) &&
      h->max_coop_blocks <= 0)
    return fail(h, FSG_E_INVALID, "persistent solver unavailable on this device");
  // the captured iterations hold the tolerance and the iteration cap
  if (p->cg_tol != h->prm.cg_tol || p->cg_max_iter != h->prm.cg_max_iter ||
      p->mg_smoother != h->prm.mg_smoother || p->mg_coarse_sweeps != h->prm.mg_coarse_sweeps ||
      p->mg_fuse_update != h->prm.mg_fuse_update || p->mg_tile_threads != h->prm.mg_tile_threads ||
      p->mg_level_sweeps != h->prm.mg_level_sweeps ||
      p->mg_level0_sweeps != h->prm.mg_level0_sweeps ||
      p->mg_spread_tiles != h->prm.mg_spread_tiles)
    h->mg_graphs.clear();
  h->prm = *p;
  return FSG_OK;
}

int fsg_get_params(const fsg_domain* h, fsg_params* p) {
  if (!h || !p) return FSG_E_INVALID;
  *p = h->prm;
  return FSG_OK;
}

int fsg_set_pic_ratio(fsg_domain* h, float r) {
  if (!h) return FSG_E_INVALID;
  // CLAMP(pic_ratio, 0, 1) (src/FluidDomain.cpp:99-102)
  const float t = r < 0.f ? 0.f : r;
  h->pic_ratio = t > 1.f ? 1.f : t;
  return FSG_OK;
}

int fsg_get_geometry(const fsg_domain* h, int32_t* nx, int32_t* ny, float* dx, float* dy,
                     float* density, float* pic_ratio) {
  if (!h) return FSG_E_INVALID;
  if (nx) *nx = h->g.nx;
  if (ny) *ny = h->g.ny;
  if (dx) *dx = h->g.dx;
  if (dy) *dy = h->g.dy;
  if (density) *density = h->density;
  if (pic_ratio) *pic_ratio = h->pic_ratio;
  return FSG_OK;
}

// The persistent solver raises bar[1] when a block abandons a grid-barrier wait
// (bounded spin); every block then leaves the kernel and the solution is invalid.
static int check_solver_abort(fsg_domain* h) {
  unsigned flag = 0;
  HIPCHK(h, hipMemcpyAsync(&flag, h->bar + 1, sizeof(flag), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (flag)
    return fail(h, FSG_E_HIP,
                "persistent pressure solver abandoned a grid-barrier wait: results are invalid");
  return FSG_OK;
}

int fsg_synchronize(fsg_domain* h) {
  if (!h) return FSG_E_INVALID;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return check_solver_abort(h);
}

static int reserve_particles(fsg_domain* h, int64_t n) {
  if (n <= h->cap) return FSG_OK;
  int64_t cap = h->cap > 0 ? h->cap : 1024;
  while (cap < n) cap *= 2;
  float4* np = nullptr;
  uint8_t* nr = nullptr;
  int* nrank = nullptr;
  float4* nsorted = nullptr;
  HIPCHK(h, hipMalloc(&np, (size_t)cap * sizeof(float4)));
  HIPCHK(h, hipMalloc(&nr, (size_t)cap));
  HIPCHK(h, hipMalloc(&nrank, (size_t)cap * sizeof(int)));
  HIPCHK(h, hipMalloc(&nsorted, (size_t)cap * sizeof(float4)));
  HIPCHK(h, hipMemsetAsync(nr, 0, (size_t)cap, h->stream));
  if (h->np > 0) {
    HIPCHK(h, hipMemcpyAsync(np, h->pt, (size_t)h->np * sizeof(float4), hipMemcpyDeviceToDevice,
                             h->stream));
    HIPCHK(h, hipMemcpyAsync(nr, h->reject, (size_t)h->np, hipMemcpyDeviceToDevice, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->pt) (void)hipFree(h->pt);
  if (h->reject) (void)hipFree(h->reject);
  if (h->prank) (void)hipFree(h->prank);
  if (h->psorted) (void)hipFree(h->psorted);
  h->pt = np;
  h->reject = nr;
  h->prank = nrank;
  h->psorted = nsorted;
  h->cap = cap;
  return FSG_OK;
}

int fsg_set_particles(fsg_domain* h, const float* xyuv, int64_t n) {
  if (!h || n < 0 || (n > 0 && !xyuv)) return FSG_E_INVALID;
  h->np = 0;
  return fsg_add_particles(h, xyuv, n);
}

int fsg_add_particles(fsg_domain* h, const float* xyuv, int64_t n) {
  if (!h || n < 0 || (n > 0 && !xyuv)) return FSG_E_INVALID;
  if (n == 0) return FSG_OK;
  int rc = reserve_particles(h, h->np + n);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(h->pt + h->np, xyuv, (size_t)n * sizeof(float4), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->np += n;
  return FSG_OK;
}

int fsg_emit_lattice(fsg_domain* h, const float* xs, int32_t nxp, const float* ys, int32_t nyp,
                     float vx, float vy) {
  if (!h || nxp < 0 || nyp < 0 || (nxp > 0 && !xs) || (nyp > 0 && !ys)) return FSG_E_INVALID;
  const int64_t n = (int64_t)nxp * nyp;
  if (n == 0) return FSG_OK;
  int rc = reserve_particles(h, h->np + n);
  if (rc) return rc;
  const size_t need = (size_t)nxp + (size_t)nyp;
  if (need > h->emit_cap) {
    if (h->emit_buf) HIPCHK(h, hipFree(h->emit_buf));
    h->emit_buf = nullptr;
    h->emit_cap = 0;
    HIPCHK(h, hipMalloc(&h->emit_buf, need * sizeof(float)));
    h->emit_cap = need;
  }
  HIPCHK(h, hipMemcpyAsync(h->emit_buf, xs, (size_t)nxp * sizeof(float), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(h->emit_buf + nxp, ys, (size_t)nyp * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
  timed(h, "emit_lattice", [&] {
    launch_emit_lattice(h->stream, h->pt, h->np, h->emit_buf, nxp, h->emit_buf + nxp, nyp, vx, vy);
  });
  // the host arrays may be released once the copies are done
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->np += n;
  return FSG_OK;
}

int fsg_advect_level_set(fsg_domain* h, float dt) {
  if (!h) return FSG_E_INVALID;
  if (!h->phi) return fail(h, FSG_E_INVALID, "fsg_advect_level_set: no level set (fsg_set_level_set)");
  if (!h->phi2 && dalloc(h, &h->phi2, ncell(h)) != hipSuccess)
    return fail(h, FSG_E_NOMEM, "level set: device allocation failed");
  timed(h, "advect_level_set",
        [&] { launch_advect_level_set(h->stream, h->g, h->phi, h->phi2, UF(h), VF(h), dt); });
  std::swap(h->phi, h->phi2);
  return FSG_OK;
}

int fsg_get_level_set(fsg_domain* h, float* phi) {
  if (!h || check_grid_ptr(h, phi)) return FSG_E_INVALID;
  if (!h->phi) return fail(h, FSG_E_INVALID, "fsg_get_level_set: no level set");
  HIPCHK(h, hipMemcpyAsync(phi, h->phi, ncell(h) * sizeof(float), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_num_particles(const fsg_domain* h, int64_t* n) {
  if (!h || !n) return FSG_E_INVALID;
  *n = h->np;
  return FSG_OK;
}

int fsg_get_particles(fsg_domain* h, float* xyuv, int64_t capacity) {
  if (!h || (h->np > 0 && !xyuv)) return FSG_E_INVALID;
  if (capacity < h->np) return fail(h, FSG_E_INVALID, "fsg_get_particles: capacity too small");
  if (h->np == 0) return FSG_OK;
  HIPCHK(h, hipMemcpyAsync(xyuv, h->pt, (size_t)h->np * sizeof(float4), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

static float* vel_ptr(fsg_domain* h, int which, int comp) {
  switch (which) {
    case FSG_VEL_FRONT: return comp ? VF(h) : UF(h);
    case FSG_VEL_BACK: return comp ? VB(h) : UB(h);
    case FSG_VEL_PREV: return comp ? h->v_prev : h->u_prev;
    case FSG_VEL_DIFF: return comp ? h->v_diff : h->u_diff;
    default: return nullptr;
  }
}

int fsg_set_velocity(fsg_domain* h, int32_t which, const float* u, const float* v) {
  if (!h) return FSG_E_INVALID;
  float* du = vel_ptr(h, which, 0);
  float* dv = vel_ptr(h, which, 1);
  if (!du) return fail(h, FSG_E_INVALID, "fsg_set_velocity: bad buffer id");
  const size_t bytes = ncell(h) * sizeof(float);
  if (u) HIPCHK(h, hipMemcpyAsync(du, u, bytes, hipMemcpyHostToDevice, h->stream));
  if (v) HIPCHK(h, hipMemcpyAsync(dv, v, bytes, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_get_velocity(fsg_domain* h, int32_t which, float* u, float* v) {
  if (!h) return FSG_E_INVALID;
  float* du = vel_ptr(h, which, 0);
  float* dv = vel_ptr(h, which, 1);
  if (!du) return fail(h, FSG_E_INVALID, "fsg_get_velocity: bad buffer id");
  const size_t bytes = ncell(h) * sizeof(float);
  if (u) HIPCHK(h, hipMemcpyAsync(u, du, bytes, hipMemcpyDeviceToHost, h->stream));
  if (v) HIPCHK(h, hipMemcpyAsync(v, dv, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_swap_velocity(fsg_domain* h) {
  if (!h) return FSG_E_INVALID;
  swap_vel(h);
  return FSG_OK;
}

int fsg_set_labels(fsg_domain* h, const uint8_t* labels) {
  if (!h || check_grid_ptr(h, labels)) return FSG_E_INVALID;
  const int nx = h->g.nx, ny = h->g.ny;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const uint8_t l = labels[(size_t)i + (size_t)j * nx];
      if (l > FSG_SOLID) return fail(h, FSG_E_INVALID, "fsg_set_labels: label out of range");
      const bool border = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
      if (border && l != FSG_SOLID)
        return fail(h, FSG_E_INVALID, "fsg_set_labels: border cells must be SOLID");
    }
  HIPCHK(h, hipMemcpyAsync(h->label, labels, ncell(h), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_get_labels(fsg_domain* h, uint8_t* labels) {
  if (!h || check_grid_ptr(h, labels)) return FSG_E_INVALID;
  HIPCHK(h, hipMemcpyAsync(labels, h->label, ncell(h), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_set_solid_mask(fsg_domain* h, const uint8_t* mask) {
  if (!h) return FSG_E_INVALID;
  if (!mask) {
    h->solid = nullptr;  // the allocation stays owned by the handle
    return FSG_OK;
  }
  if (!h->solid_buf) HIPCHK(h, dalloc(h, &h->solid_buf, ncell(h)));
  HIPCHK(h, hipMemcpyAsync(h->solid_buf, mask, ncell(h), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->solid = h->solid_buf;
  return FSG_OK;
}

int fsg_set_level_set(fsg_domain* h, const float* phi) {
  if (!h || check_grid_ptr(h, phi)) return FSG_E_INVALID;
  if (!h->phi) HIPCHK(h, dalloc(h, &h->phi, ncell(h)));
  HIPCHK(h, hipMemcpyAsync(h->phi, phi, ncell(h) * sizeof(float), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_validate_pool(fsg_domain* h, int32_t pool_nx, int32_t pool_ny, float pool_dx,
                      float pool_dy) {
  if (!h) return FSG_E_INVALID;
  float ddx = h->g.dx - pool_dx, ddy = h->g.dy - pool_dy;
  if (ddx < 0) ddx = -ddx;
  if (ddy < 0) ddy = -ddy;
  const bool ok = h->g.nx == pool_nx && h->g.ny == pool_ny && (double)ddx < 0.0000001 &&
                  (double)ddy < 0.0000001;
  return ok ? FSG_OK : fail(h, FSG_E_VALIDATE, kValidateMsg);
}

int fsg_step(fsg_domain* h, int32_t scheme, float dt) {
  if (!h) return FSG_E_INVALID;
  if (h->batch) {
    if (scheme != FSG_SCHEME_PIC && scheme != FSG_SCHEME_FLIP && scheme != FSG_SCHEME_PICFLIP)
      return fail(h, FSG_E_INVALID, "fsg_step: a batch steps PIC, FLIP or PICFLIP");
    if (h->prm.solver != FSG_SOLVER_MG || h->prm.particle_pressure_k != 0.f)
      return fail(h, FSG_E_INVALID,
                  "fsg_step: a batch needs the multigrid solver and particle_pressure_k = 0");
  }
  if (scheme != FSG_SCHEME_GRID_ADVECT) {
    int rc = check_validate(h);
    if (rc) return rc;
  }
  int rc;
  switch (scheme) {
    case FSG_SCHEME_SEMI_LAGRANGIAN: rc = step_semi_lagrangian(h, dt); break;
    case FSG_SCHEME_PIC:
    case FSG_SCHEME_FLIP:
    case FSG_SCHEME_PICFLIP: rc = step_flip_family(h, scheme, dt); break;
    case FSG_SCHEME_GRID_ADVECT: rc = step_grid_advect(h, dt); break;
    default: return fail(h, FSG_E_INVALID, "fsg_step: unknown scheme");
  }
  if (rc) return rc;
  HIPCHK(h, hipGetLastError());
  return FSG_OK;
}

int fsg_run_phase(fsg_domain* h, int32_t phase, float dt) {
  if (!h) return FSG_E_INVALID;
  if (h->batch) return fail(h, FSG_E_INVALID, "fsg_run_phase: not available for a batch");
  const bool rec = true;
  switch (phase) {
    case FSG_PH_CLASSIFY: ph_classify(h); break;
    case FSG_PH_P2G: {
      long long* fx = h->prm.deterministic ? h->acc_fx : nullptr;
      timed(h, "cell_init", [&] {
        launch_cell_init(h->stream, h->g, h->label, h->solid, h->acc, 0, fx ? 0 : 1);
      });
      timed(h, "particle_scatter", [&] {
        launch_particle_scatter(h->stream, h->g, h->pt, h->np, h->label, h->acc, 0, 1, nullptr,
                                fx);
      });
      if (fx)
        launch_fx_to_float(h->stream, 4 * (int64_t)ncell(h), fx, reinterpret_cast<float*>(h->acc));
      swap_vel(h);  // the normalise writes the (old) back, which is now the front
      face_update(h, FU_P2G, dt, false);
      break;
    }
    case FSG_PH_SAVE_PREV: face_update(h, FU_PREV, dt, false); break;
    case FSG_PH_GRAVITY: face_update(h, FU_GRAVITY, dt, false); break;
    case FSG_PH_DIRICHLET: face_update(h, FU_DIRICHLET, dt, false); break;
    case FSG_PH_EXTEND:
      face_update(h, FU_EXTINIT, dt, rec);
      extend_sweeps(h, rec);
      break;
    case FSG_PH_EXTEND_AVERAGE:
      timed(h, "extend_average_init", [&] {
        launch_extend_average(h->stream, h->g, h->label, UF(h), VF(h), UB(h), VB(h), MXF(h),
                              MXB(h), 1);
      });
      for (int it = 0; it < h->prm.ext_layers; ++it) {
        timed(h, "extend_average_sweep", [&] {
          launch_extend_average(h->stream, h->g, h->label, UF(h), VF(h), UB(h), VB(h), MXF(h),
                                MXB(h), 0);
        });
        h->mask_front ^= 1;  // FluidSolverMemoryPool::swapValidMaskBuffer
      }
      swap_vel(h);
      break;
    case FSG_PH_PRESSURE: {
      int rc = ph_pressure(h, dt, false, false);
      if (rc) return rc;
      break;
    }
    case FSG_PH_DIFF:
      timed(h, "diff", [&] {
        launch_diff(h->stream, h->g, UF(h), VF(h), h->u_prev, h->v_prev, h->u_diff, h->v_diff);
      });
      break;
    case FSG_PH_G2P_PIC:
    case FSG_PH_G2P_FLIP:
    case FSG_PH_G2P_PICFLIP: {
      const int mode = phase == FSG_PH_G2P_PIC ? 1 : phase == FSG_PH_G2P_FLIP ? 2 : 3;
      timed(h, "particle_gather", [&] {
        launch_particle_gather(h->stream, h->g, h->pt, h->np, UF(h), VF(h), h->u_diff, h->v_diff,
                               h->pic_ratio, mode, 0, dt, h->label, h->reject);
      });
      break;
    }
    case FSG_PH_ADVECT:
    case FSG_PH_ADVECT_REJECT: {
      const int move = phase == FSG_PH_ADVECT ? 1 : 2;
      timed(h, "particle_gather", [&] {
        launch_particle_gather(h->stream, h->g, h->pt, h->np, UF(h), VF(h), h->u_diff, h->v_diff,
                               h->pic_ratio, 0, move, dt, h->label, h->reject);
      });
      break;
    }
    case FSG_PH_SL_SCATTER:
      ph_sl_scatter(h, dt);
      break;
    case FSG_PH_SL_GATHER:
      timed(h, "sl_gather", [&] {
        launch_sl_gather(h->stream, h->g, h->label, UF(h), VF(h), UB(h), VB(h), h->prm.integrator, dt);
      });
      swap_vel(h);
      break;
    case FSG_PH_ADVECT_GRID:
      timed(h, "advect_grid", [&] {
        launch_advect_grid(h->stream, h->g, h->pt, h->np, UF(h), VF(h), h->prm.integrator, dt);
      });
      break;
    case FSG_PH_CLASSIFY_LEVELSET:
      if (!h->phi) return fail(h, FSG_E_INVALID, "no level set (fsg_set_level_set)");
      timed(h, "classify_levelset",
            [&] { launch_classify_levelset(h->stream, h->g, h->phi, h->solid, h->label); });
      break;
    default: return fail(h, FSG_E_INVALID, "fsg_run_phase: unknown phase");
  }
  HIPCHK(h, hipGetLastError());
  return FSG_OK;
}

// The fluid index of every cell (src/FluidSolver.cpp:331-346): the compaction's, or,
// where the last solve did not compact (the multigrid form), the row-major running
// count of the LIQUID cells of the labels, which is what the compaction forms.
static int host_fluid_indices(fsg_domain* h, int32_t* idx) {
  const size_t C = ncell(h);
  if (h->compacted) {
    HIPCHK(h, hipMemcpyAsync(idx, h->fluid_index, C * sizeof(int), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FSG_OK;
  }
  std::vector<uint8_t> lab(C);
  HIPCHK(h, hipMemcpyAsync(lab.data(), h->label, C, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  int32_t k = 0;
  for (size_t c = 0; c < C; ++c) idx[c] = lab[c] == kLiquid ? k++ : -1;
  return FSG_OK;
}

int fsg_get_pressure(fsg_domain* h, float* p) {
  if (!h || check_grid_ptr(h, p)) return FSG_E_INVALID;
  const size_t C = ncell(h);
  std::vector<int> idx(C);
  if (int rc = host_fluid_indices(h, idx.data())) return rc;
  HIPCHK(h, hipMemcpyAsync(p, h->xg, C * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t c = 0; c < C; ++c)
    if (idx[c] < 0) p[c] = 0.f;
  return FSG_OK;
}

int fsg_get_rhs(fsg_domain* h, float* bout) {
  if (!h || check_grid_ptr(h, bout)) return FSG_E_INVALID;
  const size_t C = ncell(h);
  if (!h->compacted) {  // b per cell, 0 off the LIQUID cells
    HIPCHK(h, hipMemcpyAsync(bout, h->b, C * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FSG_OK;
  }
  std::vector<int> idx(C);
  std::vector<float> bc(C);
  HIPCHK(h, hipMemcpyAsync(idx.data(), h->fluid_index, C * sizeof(int), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(bc.data(), h->b, C * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t c = 0; c < C; ++c) bout[c] = idx[c] >= 0 ? bc[(size_t)idx[c]] : 0.f;
  return FSG_OK;
}

int fsg_get_fluid_indices(fsg_domain* h, int32_t* idx) {
  if (!h || check_grid_ptr(h, idx)) return FSG_E_INVALID;
  return host_fluid_indices(h, idx);
}

int fsg_get_valid_masks(fsg_domain* h, uint8_t* mu, uint8_t* mv) {
  if (!h) return FSG_E_INVALID;
  if (mu) HIPCHK(h, hipMemcpyAsync(mu, MXF(h), ncell(h), hipMemcpyDeviceToHost, h->stream));
  if (mv) HIPCHK(h, hipMemcpyAsync(mv, MYF(h), ncell(h), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_get_ext_info(fsg_domain* h, uint8_t* iu, uint8_t* iv) {
  if (!h) return FSG_E_INVALID;
  if (iu) HIPCHK(h, hipMemcpyAsync(iu, h->info_u, ncell(h), hipMemcpyDeviceToHost, h->stream));
  if (iv) HIPCHK(h, hipMemcpyAsync(iv, h->info_v, ncell(h), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_get_step_decisions(fsg_domain* h, uint8_t* dec) {
  if (!h || !dec) return FSG_E_INVALID;
  if (!h->prm.record_decisions || !h->dec)
    return fail(h, FSG_E_INVALID, "fsg_get_step_decisions: fsg_params.record_decisions is off");
  HIPCHK(h, hipMemcpyAsync(dec, h->dec, ncell(h), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_get_reject_flags(fsg_domain* h, uint8_t* flags) {
  if (!h || (h->np > 0 && !flags)) return FSG_E_INVALID;
  if (h->np == 0) return FSG_OK;
  HIPCHK(h, hipMemcpyAsync(flags, h->reject, (size_t)h->np, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_get_stats(fsg_domain* h, fsg_stats* s) {
  if (!h || !s) return FSG_E_INVALID;
  // the plain and multigrid forms keep their statistics on the host (read with their
  // stop tests): no device access, no synchronisation
  if (h->prm.solver == FSG_SOLVER_PLAIN || h->prm.solver == FSG_SOLVER_MG) {
    *s = h->host_stats;
    return FSG_OK;
  }
  if (int rc = check_solver_abort(h)) return rc;
  CgScalars hs{};
  int F = 0;
  HIPCHK(h, hipMemcpyAsync(&hs, h->sc, sizeof(hs), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(&F, h->n_fluid, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  memset(s, 0, sizeof(*s));
  s->n_fluid = F;
  s->cg_iterations = hs.iter;
  s->cg_error = hs.err;
  s->refine_sweeps = hs.refine_sweeps;
  s->total_unknown_iterations = (int64_t)hs.total_unknown_iters + h->host_stats.total_unknown_iterations;
  s->total_solves = hs.total_solves + h->host_stats.total_solves;
  return FSG_OK;
}

int fsg_reset_stats(fsg_domain* h) {
  if (!h) return FSG_E_INVALID;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  CgScalars hs{};
  HIPCHK(h, hipMemcpyAsync(&hs, h->sc, sizeof(hs), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  hs.total_unknown_iters = 0;
  hs.total_solves = 0;
  HIPCHK(h, hipMemcpyAsync(h->sc, &hs, sizeof(hs), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->host_stats.total_unknown_iterations = 0;
  h->host_stats.total_solves = 0;
  return FSG_OK;
}

int fsg_set_timing(fsg_domain* h, int32_t enable) {
  if (!h) return FSG_E_INVALID;
  if (enable < 0 || enable > 2) return fail(h, FSG_E_INVALID, "fsg_set_timing: mode 0, 1 or 2");
  h->timer.on = enable;
  return FSG_OK;
}

int fsg_set_timing_prefix(fsg_domain* h, const char* prefix) {
  if (!h || !prefix || strlen(prefix) > 63)
    return fail(h, FSG_E_INVALID, "fsg_set_timing_prefix: prefix of at most 63 characters");
  h->timer.prefix = prefix;
  return FSG_OK;
}

int fsg_get_timing(fsg_domain* h, int32_t cap, const char** names, double* ms, int64_t* launches) {
  if (!h) return FSG_E_INVALID;
  h->timer.collect();
  h->timer.names.clear();
  for (auto& kv : h->timer.acc) h->timer.names.push_back(kv.first);
  int q = 0;
  for (auto& kv : h->timer.acc) {
    if (q >= cap) break;
    if (names) names[q] = h->timer.names[(size_t)q].c_str();
    if (ms) ms[q] = kv.second.first;
    if (launches) launches[q] = kv.second.second;
    ++q;
  }
  return (int)h->timer.acc.size();
}

int fsg_reset_timing(fsg_domain* h) {
  if (!h) return FSG_E_INVALID;
  h->timer.collect();
  h->timer.acc.clear();
  return FSG_OK;
}

// ============================================================= packed batch
// Scenes packed side by side into one grid ("shelf" rows, origins on multiples of
// 8 cells so that the first multigrid levels never mix scenes); the rest of the
// grid is SOLID.  Grid phases and the pressure solve run on the packed grid (one
// launch per phase for all scenes); particle phases use each scene's own geometry
// and scene-local coordinates.
int fsg_batch_create(fsg_domain** out, int32_t n_scenes, const int32_t* nx, const int32_t* ny,
                     float delta, float density, float pic_ratio) {
  if (!out || n_scenes < 1 || !nx || !ny || !(delta > 0.f)) return FSG_E_INVALID;
  *out = nullptr;
  auto up = [](int a, int m) { return (a + m - 1) / m * m; };
  for (int s = 0; s < n_scenes; ++s)
    if (nx[s] < 3 || ny[s] < 3 || nx[s] > (1 << 15) || ny[s] > (1 << 15)) return FSG_E_INVALID;
  // Multigrid split: packed levels 0 .. la, then each scene's own hierarchy from its
  // rectangle of level la (at most kSceneCoarseCells cells, one workgroup).  Slots are
  // aligned to 2^la (and to 8 cells) and leave at least one all-SOLID cell of level la
  // after every scene in x and y, so no level of the packed hierarchy couples scenes.
  int la = 1;
  for (bool fits = false; !fits; ) {
    fits = true;
    for (int s = 0; s < n_scenes && fits; ++s)
      fits = (int64_t)((nx[s] + (1 << la) - 1) >> la) * ((ny[s] + (1 << la) - 1) >> la) <=
             fsg::kSceneCoarseCells;
    if (!fits) ++la;
  }
  const int A = 1 << la, U = std::max(8, A);
  auto slot = [&](int n) { return up(A * ((n + A - 1) / A + 1), U); };
  int64_t area = 0;
  int wmax = 0;
  for (int s = 0; s < n_scenes; ++s) {
    area += (int64_t)slot(nx[s]) * slot(ny[s]);
    wmax = std::max(wmax, slot(nx[s]));
  }
  const int W = std::max(wmax, up((int)ceil(sqrt(1.1 * (double)area)), U));
  std::vector<int> order(n_scenes);
  for (int s = 0; s < n_scenes; ++s) order[s] = s;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ny[a] > ny[b]; });
  std::vector<int> ox(n_scenes), oy(n_scenes);
  int x = 0, y = 0, shelf = 0;
  for (int s : order) {
    const int w = slot(nx[s]), hh = slot(ny[s]);
    if (x + w > W) {
      y += shelf;
      x = 0;
      shelf = 0;
    }
    ox[s] = x;
    oy[s] = y;
    x += w;
    shelf = std::max(shelf, hh);
  }
  const int H = y + shelf;
  if ((int64_t)W * H > ((int64_t)1 << 30)) return FSG_E_INVALID;
  fsg_domain* h = nullptr;
  int rc = fsg_create(&h, W, H, (float)W * delta, (float)H * delta, density, pic_ratio);
  if (rc) return rc;
  auto bail = [&](int code, const char* msg) {
    fsg_destroy(h);
    (void)msg;
    return code;
  };
  BatchInfo* b = new (std::nothrow) BatchInfo();
  if (!b) return bail(FSG_E_NOMEM, "");
  h->batch = b;
  b->S = n_scenes;
  b->la = la;
  b->nx.assign(nx, nx + n_scenes);
  b->ny.assign(ny, ny + n_scenes);
  b->ox = ox;
  b->oy = oy;
  b->np.assign(n_scenes, 0);
  b->part_base.assign(n_scenes, 0);
  b->solid.assign((size_t)W * H, 1);
  std::vector<int> sid((size_t)W * H, -1);
  std::vector<int4> rect(n_scenes);
  std::vector<fsg::BatchScene> table(n_scenes);
  for (int s = 0; s < n_scenes; ++s) {
    const Geom gs = make_geom(nx[s], ny[s], (float)nx[s] * delta, (float)ny[s] * delta);
    if (gs.dx != h->g.dx || gs.dy != h->g.dy)
      return bail(FSG_E_INVALID, "scene spacing differs from the packed grid's");
    for (int j = 0; j < ny[s]; ++j)
      for (int i = 0; i < nx[s]; ++i) {
        const size_t c = (size_t)(ox[s] + i) + (size_t)(oy[s] + j) * W;
        sid[c] = s;
        const bool border = i == 0 || j == 0 || i == nx[s] - 1 || j == ny[s] - 1;
        b->solid[c] = border ? 1 : 0;
      }
    rect[s] = make_int4(ox[s], oy[s], nx[s], ny[s]);
    const int bbx = (nx[s] + 15) >> 4, bby = (ny[s] + 15) >> 4;
    table[s] = fsg::BatchScene{gs, 0, 0, ox[s] + oy[s] * W, W, b->n_buckets, bbx};
    b->n_buckets += bbx * bby;
    b->max_scene_buckets = std::max(b->max_scene_buckets, bbx * bby);
    // the scene's own multigrid levels: its rectangle of level la, halved until 4 x 4
    fsg::SceneCoarse c{};
    c.cx = ox[s] >> la;
    c.cy = oy[s] >> la;
    int w = (nx[s] + A - 1) >> la, hh = (ny[s] + A - 1) >> la, off = 0;
    for (;;) {
      c.nx[c.nl] = w;
      c.ny[c.nl] = hh;
      c.off[c.nl] = off;
      off += w * hh;
      ++c.nl;
      if ((w <= 4 && hh <= 4) || c.nl == fsg::kSceneMaxLevels) break;
      w = (w + 1) / 2;
      hh = (hh + 1) / 2;
    }
    if (c.nx[c.nl - 1] * c.ny[c.nl - 1] > 64) return bail(FSG_E_INVALID, "scene hierarchy");
    c.total = (off + 3) & ~3;
    b->stride = std::max(b->stride, c.total);
    b->coarse.push_back(c);
  }
  bool ok = true;
  ok &= dalloc(h, &b->d_scene, (size_t)n_scenes) == hipSuccess;
  ok &= dalloc(h, &b->d_rect, (size_t)n_scenes) == hipSuccess;
  ok &= dalloc(h, &b->d_sid, (size_t)W * H) == hipSuccess;
  ok &= dalloc(h, &b->d_norm, (size_t)n_scenes) == hipSuccess;
  ok &= dalloc(h, &b->d_inv, (size_t)n_scenes) == hipSuccess;
  ok &= dalloc(h, &b->d_coarse, (size_t)n_scenes) == hipSuccess;
  ok &= dalloc(h, &b->d_sinf, (size_t)n_scenes * b->stride) == hipSuccess;
  std::vector<int4> vtile;
  std::vector<int2> vend;
  for (int s = 0; s < n_scenes; ++s)
    for (int y0 = 0; y0 < ny[s]; y0 += 16)
      for (int x0 = 0; x0 < nx[s]; x0 += 64) {
        vtile.push_back(make_int4(ox[s] + x0, oy[s] + y0, s, 0));
        vend.push_back(make_int2(ox[s] + nx[s], oy[s] + ny[s]));
      }
  b->n_vtile = (int)vtile.size();
  ok &= dalloc(h, &b->d_vtile, vtile.size()) == hipSuccess;
  ok &= dalloc(h, &b->d_vend, vend.size()) == hipSuccess;
  ok &= dalloc(h, &b->d_cg, (size_t)5 * n_scenes) == hipSuccess;
  ok &= dalloc(h, &b->d_vflag, vtile.size()) == hipSuccess;
  if (!ok) return bail(FSG_E_NOMEM, "");
  if (hipMemcpyAsync(b->d_scene, table.data(), sizeof(fsg::BatchScene) * n_scenes,
                     hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(b->d_rect, rect.data(), sizeof(int4) * n_scenes, hipMemcpyHostToDevice,
                     h->stream) != hipSuccess ||
      hipMemcpyAsync(b->d_sid, sid.data(), sizeof(int) * sid.size(), hipMemcpyHostToDevice,
                     h->stream) != hipSuccess ||
      hipMemcpyAsync(b->d_coarse, b->coarse.data(), sizeof(fsg::SceneCoarse) * n_scenes,
                     hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(b->d_vtile, vtile.data(), sizeof(int4) * vtile.size(),
                     hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(b->d_vend, vend.data(), sizeof(int2) * vend.size(),
                     hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return bail(FSG_E_HIP, "");
  if ((rc = fsg_set_solid_mask(h, b->solid.data())) != FSG_OK) return bail(rc, "");
  // labels as MacGrid's constructor leaves them, with the scene borders SOLID
  launch_cell_init(h->stream, h->g, h->label, h->solid, h->acc, 1, 0);
  if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(FSG_E_HIP, "");
  h->prm.solver = FSG_SOLVER_MG;
  *out = h;
  return FSG_OK;
}

int fsg_batch_num_scenes(const fsg_domain* h, int32_t* n) {
  if (!h || !h->batch || !n) return FSG_E_INVALID;
  *n = h->batch->S;
  return FSG_OK;
}

int fsg_batch_scene_rect(const fsg_domain* h, int32_t scene, int32_t* rect4) {
  if (!h || !h->batch || !rect4 || scene < 0 || scene >= h->batch->S) return FSG_E_INVALID;
  const BatchInfo& b = *h->batch;
  rect4[0] = b.ox[scene];
  rect4[1] = b.oy[scene];
  rect4[2] = b.nx[scene];
  rect4[3] = b.ny[scene];
  return FSG_OK;
}

// All scenes' particles at once: counts[s] particles of scene s, concatenated, in
// scene-local coordinates.
int fsg_batch_set_particles(fsg_domain* h, const int64_t* counts, const float* xyuv) {
  if (!h || !h->batch || !counts) return FSG_E_INVALID;
  BatchInfo& b = *h->batch;
  int64_t total = 0;
  std::vector<int2> ptile;
  std::vector<fsg::BatchScene> table(b.S);
  HIPCHK(h, hipMemcpy(table.data(), b.d_scene, sizeof(fsg::BatchScene) * b.S,
                      hipMemcpyDeviceToHost));
  for (int s = 0; s < b.S; ++s) {
    if (counts[s] < 0) return fail(h, FSG_E_INVALID, "fsg_batch_set_particles: negative count");
    b.part_base[s] = total;
    b.np[s] = counts[s];
    table[s].part_base = total;
    table[s].np = counts[s];
    for (int64_t t = 0; t < (counts[s] + 255) / 256; ++t) ptile.push_back(make_int2(s, (int)t));
    total += counts[s];
  }
  if (total > 0 && !xyuv) return FSG_E_INVALID;
  if (int rc = fsg_set_particles(h, xyuv, total)) return rc;
  if ((int)ptile.size() > b.n_ptile) {
    HIPCHK(h, dalloc(h, &b.d_ptile, ptile.size()));
  }
  b.n_ptile = (int)ptile.size();
  if (b.n_ptile > 0)
    HIPCHK(h, hipMemcpyAsync(b.d_ptile, ptile.data(), sizeof(int2) * ptile.size(),
                             hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(b.d_scene, table.data(), sizeof(fsg::BatchScene) * b.S,
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

// Interior SOLID cells of one scene (its ny x nx mask, row-major); its border stays SOLID.
int fsg_batch_set_solid_mask(fsg_domain* h, int32_t scene, const uint8_t* mask) {
  if (!h || !h->batch || !mask || scene < 0 || scene >= h->batch->S) return FSG_E_INVALID;
  BatchInfo& b = *h->batch;
  const int W = h->g.nx, nx = b.nx[scene], ny = b.ny[scene];
  for (int j = 1; j < ny - 1; ++j)
    for (int i = 1; i < nx - 1; ++i)
      b.solid[(size_t)(b.ox[scene] + i) + (size_t)(b.oy[scene] + j) * W] =
          mask[(size_t)i + (size_t)j * nx] ? 1 : 0;
  return fsg_set_solid_mask(h, b.solid.data());
}

int fsg_batch_get_particles(fsg_domain* h, int32_t scene, float* xyuv, int64_t capacity) {
  if (!h || !h->batch || scene < 0 || scene >= h->batch->S) return FSG_E_INVALID;
  const BatchInfo& b = *h->batch;
  if (capacity < b.np[scene]) return fail(h, FSG_E_INVALID, "fsg_batch_get_particles: capacity");
  if (b.np[scene] == 0) return FSG_OK;
  HIPCHK(h, hipMemcpyAsync(xyuv, h->pt + b.part_base[scene], sizeof(float4) * b.np[scene],
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_batch_num_particles(const fsg_domain* h, int32_t scene, int64_t* n) {
  if (!h || !h->batch || !n || scene < 0 || scene >= h->batch->S) return FSG_E_INVALID;
  *n = h->batch->np[scene];
  return FSG_OK;
}

// One scene's rectangle of a packed grid into a dense nx * ny host array.
static int batch_copy_rect(fsg_domain* h, int scene, const void* src, size_t elem, void* dst) {
  const BatchInfo& b = *h->batch;
  const size_t W = (size_t)h->g.nx;
  const char* base =
      static_cast<const char*>(src) + ((size_t)b.ox[scene] + (size_t)b.oy[scene] * W) * elem;
  HIPCHK(h, hipMemcpy2DAsync(dst, (size_t)b.nx[scene] * elem, base, W * elem,
                             (size_t)b.nx[scene] * elem, (size_t)b.ny[scene],
                             hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_batch_get_velocity(fsg_domain* h, int32_t scene, int32_t which, float* u, float* v) {
  if (!h || !h->batch || scene < 0 || scene >= h->batch->S) return FSG_E_INVALID;
  float* du = vel_ptr(h, which, 0);
  float* dv = vel_ptr(h, which, 1);
  if (!du) return fail(h, FSG_E_INVALID, "fsg_batch_get_velocity: bad buffer id");
  if (u)
    if (int rc = batch_copy_rect(h, scene, du, sizeof(float), u)) return rc;
  if (v)
    if (int rc = batch_copy_rect(h, scene, dv, sizeof(float), v)) return rc;
  return FSG_OK;
}

int fsg_batch_get_labels(fsg_domain* h, int32_t scene, uint8_t* labels) {
  if (!h || !h->batch || !labels || scene < 0 || scene >= h->batch->S) return FSG_E_INVALID;
  return batch_copy_rect(h, scene, h->label, 1, labels);
}

// ---------------------------------------------------------------- checkpoints
int fsg_checkpoint_save(fsg_domain* h, int32_t slot) {
  if (!h || h->batch || slot < 0 || slot >= (1 << 16)) return FSG_E_INVALID;
  if ((size_t)slot >= h->ckpt.size()) h->ckpt.resize((size_t)slot + 1);
  Checkpoint& c = h->ckpt[(size_t)slot];
  const size_t C = ncell(h);
  if (!c.u[0]) {
    HIPCHK(h, dalloc(h, &c.u[0], C));
    HIPCHK(h, dalloc(h, &c.u[1], C));
    HIPCHK(h, dalloc(h, &c.v[0], C));
    HIPCHK(h, dalloc(h, &c.v[1], C));
  }
  if (c.cap < h->np) {
    HIPCHK(h, hipStreamSynchronize(h->stream));  // earlier copies may still use the slot
    if (c.pt) (void)hipFree(c.pt);
    c.pt = nullptr;
    c.cap = 0;
    HIPCHK(h, hipMalloc(&c.pt, (size_t)h->np * sizeof(float4)));
    c.cap = h->np;
  }
  c.np = h->np;
  if (h->np > 0)
    HIPCHK(h, hipMemcpyAsync(c.pt, h->pt, (size_t)h->np * sizeof(float4),
                             hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(c.u[0], UF(h), C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(c.v[0], VF(h), C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(c.u[1], UB(h), C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(c.v[1], VB(h), C * 4, hipMemcpyDeviceToDevice, h->stream));
  return FSG_OK;  // stream-ordered: later work on the handle sees the copy
}

int fsg_checkpoint_load(fsg_domain* h, int32_t slot) {
  if (!h || h->batch || slot < 0 || (size_t)slot >= h->ckpt.size() || !h->ckpt[slot].u[0])
    return FSG_E_INVALID;
  const Checkpoint& c = h->ckpt[(size_t)slot];
  const size_t C = ncell(h);
  if (int rc = reserve_particles(h, c.np)) return rc;
  h->np = c.np;
  if (c.np > 0)
    HIPCHK(h, hipMemcpyAsync(h->pt, c.pt, (size_t)c.np * sizeof(float4),
                             hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(UF(h), c.u[0], C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(VF(h), c.v[0], C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(UB(h), c.u[1], C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(VB(h), c.v[1], C * 4, hipMemcpyDeviceToDevice, h->stream));
  return FSG_OK;  // stream-ordered
}

int fsg_checkpoint_clear(fsg_domain* h) {
  if (!h) return FSG_E_INVALID;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // the slots' grids stay with the handle's allocations; particle copies are freed
  for (Checkpoint& c : h->ckpt) {
    if (c.pt) (void)hipFree(c.pt);
    c.pt = nullptr;
    c.np = c.cap = 0;
  }
  return FSG_OK;
}

// ---------------------------------------------------------------- autotuner
// The multigrid solve's settings (smoother, coarse solve, fused update, iterations per
// graph) chosen by timing whole steps from one saved state: every candidate runs one
// untimed step (graph capture), then `steps` timed steps, each from the saved state,
// which is restored at the end (particles, the four velocity grids, labels, reject
// flags, statistics), so the handle continues as if the call had not been made,
// with the fastest candidate's settings.
struct TuneState {
  float4* pt = nullptr;
  float* g[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  uint8_t* label = nullptr;
  uint8_t* reject = nullptr;
  int64_t np = 0;
  std::vector<void*> mem;
  ~TuneState() {
    for (void* p : mem) (void)hipFree(p);
  }
};

static int tune_save(fsg_domain* h, TuneState& t) {
  const size_t C = ncell(h);
  t.np = h->np;
  auto get = [&](size_t bytes, void** out) -> hipError_t {
    const hipError_t e = hipMalloc(out, bytes ? bytes : 16);
    if (e == hipSuccess) t.mem.push_back(*out);
    return e;
  };
  HIPCHK(h, get((size_t)h->np * sizeof(float4), reinterpret_cast<void**>(&t.pt)));
  for (float*& q : t.g) HIPCHK(h, get(C * 4, reinterpret_cast<void**>(&q)));
  HIPCHK(h, get(C, reinterpret_cast<void**>(&t.label)));
  HIPCHK(h, get((size_t)h->np, reinterpret_cast<void**>(&t.reject)));
  const float* src[6] = {UF(h), VF(h), UB(h), VB(h), h->u_prev, h->v_prev};
  if (h->np > 0) {
    HIPCHK(h, hipMemcpyAsync(t.pt, h->pt, (size_t)h->np * sizeof(float4),
                             hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(t.reject, h->reject, (size_t)h->np, hipMemcpyDeviceToDevice,
                             h->stream));
  }
  for (int k = 0; k < 6; ++k)
    HIPCHK(h, hipMemcpyAsync(t.g[k], src[k], C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(t.label, h->label, C, hipMemcpyDeviceToDevice, h->stream));
  return FSG_OK;
}

static int tune_load(fsg_domain* h, const TuneState& t, bool with_diff) {
  const size_t C = ncell(h);
  if (t.np > 0) {
    HIPCHK(h, hipMemcpyAsync(h->pt, t.pt, (size_t)t.np * sizeof(float4),
                             hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->reject, t.reject, (size_t)t.np, hipMemcpyDeviceToDevice,
                             h->stream));
  }
  float* dst[6] = {UF(h), VF(h), UB(h), VB(h), h->u_prev, h->v_prev};
  for (int k = 0; k < (with_diff ? 6 : 4); ++k)
    HIPCHK(h, hipMemcpyAsync(dst[k], t.g[k], C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->label, t.label, C, hipMemcpyDeviceToDevice, h->stream));
  return FSG_OK;
}

int fsg_autotune(fsg_domain* h, int32_t scheme, float dt, int32_t steps, fsg_tune* out) {
  if (!h || steps < 1 || steps > 64) return FSG_E_INVALID;
  if (h->prm.solver != FSG_SOLVER_MG)
    return fail(h, FSG_E_INVALID, "fsg_autotune: tunes the multigrid solver (FSG_SOLVER_MG)");
  if (h->np == 0 && scheme != FSG_SCHEME_GRID_ADVECT)
    return fail(h, FSG_E_INVALID, "fsg_autotune: no particles to step");
  struct Cand {
    int sweeps, sweeps0, coarse, fuse, iters, order, threads;
  };
  std::vector<Cand> cands;
  const fsg_params p0 = h->prm;
  const bool can_fuse = !h->batch;
  const bool has_particles = h->np > 0;
  // smoothing sweeps (levels below 0, level 0): three presets; the smoother weights of
  // the two-sweep levels stay the handle's
  static const int kSweeps[3][2] = {{5, 3}, {5, 2}, {2, 2}};
  for (const auto& ls : kSweeps)
    for (int co : {0, 8, 16})
      for (int fu : can_fuse ? std::initializer_list<int>{1, 0} : std::initializer_list<int>{1})
        for (int gi : {0, 2})
          for (int so : has_particles ? std::initializer_list<int>{1, 0}
                                      : std::initializer_list<int>{p0.splat_order})
            for (int nt : {256, 512}) cands.push_back({ls[0], ls[1], co, fu, gi, so, nt});
  TuneState saved;
  if (int rc = tune_save(h, saved)) return rc;
  const fsg_stats stats0 = h->host_stats;
  const int last0 = h->mg_last_iters, prev0 = h->mg_prev_iters;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double best = 1e300, cur = -1.0;
  int best_k = 0, best_iters = 0, cur_k = -1, cur_iters = 0;
  for (size_t k = 0; k < cands.size(); ++k) {
    const Cand& c = cands[k];
    fsg_params q = p0;
    q.mg_level_sweeps = c.sweeps;
    q.mg_level0_sweeps = c.sweeps0;
    q.mg_coarse_sweeps = c.coarse;
    q.mg_fuse_update = c.fuse;
    q.graph_iters = c.iters;
    q.splat_order = c.order;
    q.mg_tile_threads = c.threads;
    if (int rc = fsg_set_params(h, &q)) return rc;
    double t_sum = 0.0;
    int its = 0;
    for (int s = -1; s < steps; ++s) {  // s = -1: the untimed step (graph capture)
      if (int rc = tune_load(h, saved, false)) return rc;
      HIPCHK(h, hipStreamSynchronize(h->stream));
      const auto t0 = std::chrono::steady_clock::now();
      if (int rc = fsg_step(h, scheme, dt)) return rc;
      HIPCHK(h, hipStreamSynchronize(h->stream));
      const auto t1 = std::chrono::steady_clock::now();
      if (s >= 0) t_sum += std::chrono::duration<double, std::milli>(t1 - t0).count();
      its = h->host_stats.cg_iterations;
    }
    const double ms = t_sum / steps;
    const bool is_cur = c.sweeps == p0.mg_level_sweeps && c.sweeps0 == p0.mg_level0_sweeps &&
                        c.coarse == p0.mg_coarse_sweeps &&
                        c.fuse == p0.mg_fuse_update && c.iters == p0.graph_iters &&
                        c.order == p0.splat_order && c.threads == p0.mg_tile_threads;
    if (is_cur) {
      cur = ms;
      cur_k = (int)k;
      cur_iters = its;
    }
    if (ms < best) {
      best = ms;
      best_k = (int)k;
      best_iters = its;
    }
  }
  // the handle's own configuration stays unless a candidate beats it by more than
  // kTuneMargin (the spread of a few timed steps)
  constexpr double kTuneMargin = 0.02;
  if (cur_k >= 0 && best > cur * (1.0 - kTuneMargin)) {
    best = cur;
    best_k = cur_k;
    best_iters = cur_iters;
  }
  fsg_params q = p0;
  q.mg_level_sweeps = cands[best_k].sweeps;
  q.mg_level0_sweeps = cands[best_k].sweeps0;
  q.mg_coarse_sweeps = cands[best_k].coarse;
  q.mg_fuse_update = cands[best_k].fuse;
  q.graph_iters = cands[best_k].iters;
  q.splat_order = cands[best_k].order;
  q.mg_tile_threads = cands[best_k].threads;
  if (int rc = fsg_set_params(h, &q)) return rc;
  if (int rc = tune_load(h, saved, true)) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->host_stats = stats0;
  h->mg_last_iters = last0;
  h->mg_prev_iters = prev0;
  if (out) {
    memset(out, 0, sizeof(*out));
    out->mg_smoother = q.mg_smoother;
    out->mg_coarse_sweeps = q.mg_coarse_sweeps;
    out->mg_fuse_update = q.mg_fuse_update;
    out->graph_iters = q.graph_iters;
    out->splat_order = q.splat_order;
    out->mg_tile_threads = q.mg_tile_threads;
    out->mg_level_sweeps = q.mg_level_sweeps;
    out->mg_level0_sweeps = q.mg_level0_sweeps;
    out->candidates = (int32_t)cands.size();
    out->cg_iterations = best_iters;
    out->ms_per_step = best;
    out->ms_per_step_before = cur;
  }
  return FSG_OK;
}

// ---------------------------------------------------------------- reverse mode
static int adj_prepare(fsg_domain* h) {
  AdjState& a = h->adj;
  const size_t C = ncell(h);
  if (!a.grids) {
    float** grids[] = {&a.bu, &a.bv, &a.fu, &a.fv, &a.f0u, &a.f0v, &a.su, &a.sv, &a.eu,
                       &a.ev, &a.e2u, &a.e2v, &a.g4u, &a.g4v, &a.g1u, &a.g1v, &a.qu, &a.qv};
    for (float** g : grids) HIPCHK(h, dalloc(h, g, C));
    HIPCHK(h, dalloc(h, &a.y, C));
    HIPCHK(h, dalloc(h, &a.zmask, C));
    a.grids = true;
  }
  while ((int)a.rmask.size() < std::max(h->prm.ext_layers, 1)) {
    uint8_t* m = nullptr;
    HIPCHK(h, dalloc(h, &m, C));
    a.rmask.push_back(m);
  }
  const int64_t need = std::max<int64_t>(h->np, 1);
  if (a.pcap < need) {
    int64_t cap = a.pcap > 0 ? a.pcap : 1024;
    while (cap < need) cap *= 2;
    float4 *pb = nullptr, *x0 = nullptr;
    HIPCHK(h, hipMalloc(&pb, (size_t)cap * sizeof(float4)));
    HIPCHK(h, hipMalloc(&x0, (size_t)cap * sizeof(float4)));
    HIPCHK(h, hipMemsetAsync(pb, 0, (size_t)cap * sizeof(float4), h->stream));
    if (a.pcap > 0)
      HIPCHK(h, hipMemcpyAsync(pb, a.pbar, (size_t)a.pcap * sizeof(float4),
                               hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (a.pbar) (void)hipFree(a.pbar);
    if (a.x0) (void)hipFree(a.x0);
    a.pbar = pb;
    a.x0 = x0;
    a.pcap = cap;
  }
  return FSG_OK;
}

int fsg_adjoint_set(fsg_domain* h, const float* pbar_xyuv, const float* bbar_u,
                    const float* bbar_v) {
  if (!h || h->batch) return FSG_E_INVALID;
  if (int rc = adj_prepare(h)) return rc;
  AdjState& a = h->adj;
  const size_t C = ncell(h);
  if (h->np > 0) {
    if (pbar_xyuv)
      HIPCHK(h, hipMemcpyAsync(a.pbar, pbar_xyuv, (size_t)h->np * sizeof(float4),
                               hipMemcpyHostToDevice, h->stream));
    else
      launch_fill(h->stream, reinterpret_cast<float*>(a.pbar), 4 * h->np, 0.f);
  }
  if (bbar_u)
    HIPCHK(h, hipMemcpyAsync(a.bu, bbar_u, C * 4, hipMemcpyHostToDevice, h->stream));
  else
    launch_fill(h->stream, a.bu, (int64_t)C, 0.f);
  if (bbar_v)
    HIPCHK(h, hipMemcpyAsync(a.bv, bbar_v, C * 4, hipMemcpyHostToDevice, h->stream));
  else
    launch_fill(h->stream, a.bv, (int64_t)C, 0.f);
  launch_fill(h->stream, a.fu, (int64_t)C, 0.f);
  launch_fill(h->stream, a.fv, (int64_t)C, 0.f);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_adjoint_set_front(fsg_domain* h, const float* fbar_u, const float* fbar_v) {
  if (!h || h->batch) return FSG_E_INVALID;
  if (int rc = adj_prepare(h)) return rc;
  AdjState& a = h->adj;
  const size_t C = ncell(h);
  if (fbar_u)
    HIPCHK(h, hipMemcpyAsync(a.fu, fbar_u, C * 4, hipMemcpyHostToDevice, h->stream));
  else
    launch_fill(h->stream, a.fu, (int64_t)C, 0.f);
  if (fbar_v)
    HIPCHK(h, hipMemcpyAsync(a.fv, fbar_v, C * 4, hipMemcpyHostToDevice, h->stream));
  else
    launch_fill(h->stream, a.fv, (int64_t)C, 0.f);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_adjoint_get_front(fsg_domain* h, float* fbar_u, float* fbar_v) {
  if (!h || h->batch || !h->adj.grids) return FSG_E_INVALID;
  const AdjState& a = h->adj;
  const size_t C = ncell(h);
  if (fbar_u) HIPCHK(h, hipMemcpyAsync(fbar_u, a.fu, C * 4, hipMemcpyDeviceToHost, h->stream));
  if (fbar_v) HIPCHK(h, hipMemcpyAsync(fbar_v, a.fv, C * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_adjoint_get(fsg_domain* h, float* pbar_xyuv, float* bbar_u, float* bbar_v) {
  if (!h || h->batch || !h->adj.grids) return FSG_E_INVALID;
  const AdjState& a = h->adj;
  const size_t C = ncell(h);
  if (pbar_xyuv && h->np > a.pcap)
    return fail(h, FSG_E_INVALID, "fsg_adjoint_get: more particles than the adjoint state holds");
  if (pbar_xyuv && h->np > 0)
    HIPCHK(h, hipMemcpyAsync(pbar_xyuv, a.pbar, (size_t)h->np * sizeof(float4),
                             hipMemcpyDeviceToHost, h->stream));
  if (bbar_u) HIPCHK(h, hipMemcpyAsync(bbar_u, a.bu, C * 4, hipMemcpyDeviceToHost, h->stream));
  if (bbar_v) HIPCHK(h, hipMemcpyAsync(bbar_v, a.bv, C * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_particle_density(fsg_domain* h, float* rho) {
  if (!h || h->batch || !rho) return FSG_E_INVALID;
  if (int rc = adj_prepare(h)) return rc;
  const size_t C = ncell(h);
  long long* fx = h->prm.deterministic ? h->acc_fx : nullptr;
  if (!fx) launch_fill(h->stream, h->adj.su, (int64_t)C, 0.f);
  launch_density(h->stream, h->g, h->pt, h->np, h->adj.su, fx);
  HIPCHK(h, hipMemcpyAsync(rho, h->adj.su, C * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

int fsg_adjoint_seed_density(fsg_domain* h, const float* rho_bar) {
  if (!h || h->batch || !rho_bar) return FSG_E_INVALID;
  if (int rc = adj_prepare(h)) return rc;
  AdjState& a = h->adj;
  const size_t C = ncell(h);
  HIPCHK(h, hipMemcpyAsync(a.su, rho_bar, C * 4, hipMemcpyHostToDevice, h->stream));
  launch_density_adj(h->stream, h->g, h->pt, h->np, a.su, a.pbar);
  for (float* z : {a.bu, a.bv, a.fu, a.fv}) launch_fill(h->stream, z, (int64_t)C, 0.f);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSG_OK;
}

// The semi-Lagrangian step's adjoint (the particles' start copy already taped): the
// front at the step's start taped, the step run with the second Dirichlet pass's
// decisions recorded, then in reverse: the particles' advection through the new front,
// the pressure update and its solve (as for the PIC / FLIP family, with no FLIP
// difference), gravity + the first Dirichlet pass and the velocity scatter.  The
// front's adjoint is the carried grid state; the back is zeroed before the step reads
// it, so its adjoint at the start is zero.
static int adjoint_step_sl(fsg_domain* h, float dt) {
  AdjState& a = h->adj;
  const int64_t np = h->np;
  const size_t C = ncell(h);
  HIPCHK(h, hipMemcpyAsync(a.f0u, UF(h), C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(a.f0v, VF(h), C * 4, hipMemcpyDeviceToDevice, h->stream));
  a.taping = true;
  int rc = step_semi_lagrangian(h, dt);
  a.taping = false;
  if (rc) return rc;
  if (h->last_F == 0)
    return fail(h, FSG_E_INVALID,
                "fsg_adjoint_step: no LIQUID cell (the no-fluid step has no adjoint here)");
  const fsg_stats fwd = h->host_stats;
  const float dt_over_density = dt / h->density;
  HIPCHK(h, hipMemcpyAsync(a.su, a.fu, C * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(a.sv, a.fv, C * 4, hipMemcpyDeviceToDevice, h->stream));
  timed(h, "adj_advect_grid", [&] {
    launch_adj_advect_grid(h->stream, h->g, a.x0, np, UF(h), VF(h), h->prm.integrator, dt, a.pbar,
                           a.su, a.sv);
  });
  // the pressure update: eu = the source's adjoint through the direct term, g4u = the
  // scatter's (faces the update keeps)
  timed(h, "adj_post", [&] {
    launch_adj_post(h->stream, h->g, h->label, a.zmask, 1.0f, 0, a.su, a.sv, nullptr, nullptr,
                    a.eu, a.ev, a.g4u, a.g4v, a.g1u, a.g1v);
    launch_adj_pgrad(h->stream, h->g, h->label, a.eu, a.ev, dt_over_density, a.y);
  });
  rc = mg_solve(h, h->last_F, a.y);
  if (rc) return rc;
  a.last_iters = h->host_stats.cg_iterations;
  h->host_stats = fwd;
  // divergence adjoint, plus the back buffer after the step (the source itself)
  timed(h, "adj_div", [&] {
    if (h->prm.project_mean) launch_mg_remove_mean(h->stream, h->mg);
    launch_adj_div(h->stream, h->g, h->mg.inf[0], h->mg.x[0], a.bu, a.bv, a.eu, a.ev);
  });
  timed(h, "adj_sl", [&] {
    launch_adj_sl(h->stream, h->g, h->label, a.f0u, a.f0v, a.eu, a.ev, a.g4u, a.g4v,
                  h->prm.gravity_x * dt, h->prm.gravity_y * dt, h->prm.closed_walls,
                  h->prm.integrator, dt, a.fu, a.fv);
    launch_fill(h->stream, a.bu, (int64_t)C, 0.f);
    launch_fill(h->stream, a.bv, (int64_t)C, 0.f);
  });
  HIPCHK(h, hipGetLastError());
  return FSG_OK;
}

// One step forward from the handle's state, taped, then its adjoint: the adjoint
// state (particles, front and back buffers) is carried from the step's end to its
// start.
int fsg_adjoint_step(fsg_domain* h, int32_t scheme, float dt) {
  if (!h) return FSG_E_INVALID;
  if (h->batch) return fail(h, FSG_E_INVALID, "fsg_adjoint_step: not available for a batch");
  if (scheme != FSG_SCHEME_PIC && scheme != FSG_SCHEME_FLIP && scheme != FSG_SCHEME_PICFLIP &&
      scheme != FSG_SCHEME_SEMI_LAGRANGIAN)
    return fail(h, FSG_E_INVALID,
                "fsg_adjoint_step: SEMI_LAGRANGIAN, PIC, FLIP or PICFLIP only");
  if (h->prm.solver != FSG_SOLVER_MG || h->prm.warm_start != 0)
    return fail(h, FSG_E_INVALID,
                "fsg_adjoint_step: needs FSG_SOLVER_MG with warm_start = 0 (a replayable solve)");
  if (int rc = check_validate(h)) return rc;
  if (int rc = adj_prepare(h)) return rc;
  AdjState& a = h->adj;
  const int64_t np = h->np;
  const size_t C = ncell(h);
  if (np > 0)
    HIPCHK(h, hipMemcpyAsync(a.x0, h->pt, (size_t)np * sizeof(float4), hipMemcpyDeviceToDevice,
                             h->stream));
  if (scheme == FSG_SCHEME_SEMI_LAGRANGIAN) return adjoint_step_sl(h, dt);
  a.taping = true;
  int rc = step_flip_family(h, scheme, dt);
  a.taping = false;
  if (rc) return rc;
  if (h->last_F == 0)
    return fail(h, FSG_E_INVALID,
                "fsg_adjoint_step: no LIQUID cell (the no-fluid step has no adjoint here)");
  const fsg_stats fwd = h->host_stats;
  const bool flip = scheme != FSG_SCHEME_PIC;
  const float r = scheme == FSG_SCHEME_PIC ? 1.0f : scheme == FSG_SCHEME_FLIP ? 0.0f : h->pic_ratio;
  const float dt_over_density = dt / h->density;
  // G2P + advect
  timed(h, "adj_g2p", [&] {
    launch_fill(h->stream, a.su, (int64_t)C, 0.f);
    launch_fill(h->stream, a.sv, (int64_t)C, 0.f);
    launch_adj_g2p(h->stream, h->g, a.x0, np, UF(h), VF(h), h->u_diff, h->v_diff, r, dt,
                   h->reject, a.pbar, a.su, a.sv);
  });
  // pressure update: direct term, pressure adjoint, adjoint solve, divergence adjoint
  timed(h, "adj_post", [&] {
    // the front after the step (H') also carries the front adjoint of the step after
    // it (non-zero when a semi-Lagrangian step follows)
    launch_adj_post(h->stream, h->g, h->label, a.zmask, r, flip ? 1 : 0, a.su, a.sv, a.fu, a.fv,
                    a.eu, a.ev, a.g4u, a.g4v, a.g1u, a.g1v);
    launch_adj_pgrad(h->stream, h->g, h->label, a.eu, a.ev, dt_over_density, a.y);
  });
  rc = mg_solve(h, h->last_F, a.y);
  if (rc) return rc;
  a.last_iters = h->host_stats.cg_iterations;
  h->host_stats = fwd;
  timed(h, "adj_div", [&] {
    if (h->prm.project_mean) launch_mg_remove_mean(h->stream, h->mg);
    launch_adj_div(h->stream, h->g, h->mg.inf[0], h->mg.x[0], a.bu, a.bv, a.eu, a.ev);
  });
  // extension: replay the read masks, then the sweeps in reverse
  const int n = h->prm.ext_layers;
  float *cu = a.eu, *cv = a.ev, *ou = a.e2u, *ov = a.e2v;
  timed(h, "adj_extend", [&] {
    launch_adj_mask_init(h->stream, h->g, h->label, a.rmask[0], nullptr);
    for (int k = 1; k < n; ++k)
      launch_adj_mask_step(h->stream, h->g, h->label, a.rmask[k - 1],
                           a.rmask[k >= 2 ? k - 2 : 0], a.rmask[k], nullptr);
    for (int k = n; k >= 1; --k) {
      launch_adj_ext_sweep(h->stream, h->g, h->label, a.rmask[k - 1], cu, cv, ou, ov);
      std::swap(cu, ou);
      std::swap(cv, ov);
    }
  });
  // extension init, Dirichlet, gravity, the P2G split, then P2G per particle
  timed(h, "adj_p2g", [&] {
    launch_adj_pre(h->stream, h->g, h->label, a.rmask[0], cu, cv, a.g4u, a.g4v, a.g1u, a.g1v,
                   h->u_prev, h->v_prev, h->acc, h->prm.gravity_x * dt, h->prm.gravity_y * dt,
                   h->prm.closed_walls, a.qu, a.qv, a.bu, a.bv);
    launch_adj_p2g(h->stream, h->g, a.x0, np, h->u_prev, h->v_prev, a.qu, a.qv, a.pbar);
    // the step overwrites its start front before reading it
    launch_fill(h->stream, a.fu, (int64_t)C, 0.f);
    launch_fill(h->stream, a.fv, (int64_t)C, 0.f);
  });
  HIPCHK(h, hipGetLastError());
  return FSG_OK;  // stream-ordered; fsg_adjoint_get synchronises
}

int fsg_set_kernel_clock(fsg_domain* h, int32_t enable) {
  if (!h) return FSG_E_INVALID;
  if (enable < 0 || enable > 2)
    return fail(h, FSG_E_INVALID, "fsg_set_kernel_clock: enable is 0, 1 or 2");
  h->kclock = enable;
  h->mg.clk_span_n = 0;
  // allocated here when the hierarchy exists, so that the first clocked solve does not
  if (enable == 2 && h->mg_ready && !h->mg.clk_span) alloc_clk_span(h);
  if (h->mg.clk) {
    HIPCHK(h, hipMemsetAsync(h->mg.clk, 0, 2 * sizeof(unsigned long long), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return FSG_OK;
}

int fsg_get_kernel_clock(fsg_domain* h, double* us_total, int64_t* launches) {
  if (!h || !us_total || !launches) return FSG_E_INVALID;
  *us_total = 0.0;
  *launches = 0;
  int khz = 0;
  HIPCHK(h, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device));
  if (khz <= 0) return fail(h, FSG_E_HIP, "fsg_get_kernel_clock: no device wall clock rate");
  unsigned long long ticks = 0;
  int64_t n = 0;
  if (h->kclock == 2) {
    // per launch: latest block end minus earliest block start
    const MgHost& m = h->mg;
    if (m.clk_span && m.clk_span_n > 0) {
      std::vector<unsigned long long> v(2 * (size_t)m.clk_span_nb * m.clk_span_n);
      HIPCHK(h, hipMemcpyAsync(v.data(), m.clk_span, v.size() * sizeof(v[0]),
                               hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      for (int k = 0; k < m.clk_span_n; ++k) {
        const unsigned long long* b = v.data() + 2 * (size_t)m.clk_span_nb * k;
        unsigned long long t0 = ~0ull, t1 = 0;
        for (int q = 0; q < m.clk_span_nb; ++q) {
          t0 = std::min(t0, b[2 * q]);
          t1 = std::max(t1, b[2 * q + 1]);
        }
        ticks += t1 - t0;
      }
      n = m.clk_span_n;
    }
  } else if (h->mg.clk) {
    unsigned long long v[2] = {0, 0};
    HIPCHK(h, hipMemcpyAsync(v, h->mg.clk, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    ticks = v[0];
    n = (int64_t)v[1];
  }
  *us_total = (double)ticks * 1000.0 / (double)khz;
  *launches = n;
  return FSG_OK;
}

int fsg_adjoint_last_iterations(const fsg_domain* h, int32_t* iters) {
  if (!h || !iters) return FSG_E_INVALID;
  *iters = h->adj.last_iters;
  return FSG_OK;
}

}  // extern "C"
