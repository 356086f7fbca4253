// kernels_grid.hip -- gfx950 kernels for every per-step phase except the CG solve.
//
// Layout in HBM: face velocities are row-major float32 grids of nx*ny (u(i,j) is
// the left face of cell (i,j), v(i,j) the bottom face); labels and extension masks
// are uint8 grids; the P2G accumulators are interleaved (sum, weight) float2 grids;
// particles are AoS float4 (x, y, vx, vy).  Grid kernels run one thread per cell
// (both faces of the cell), 256 threads per block over a 1-D tile order remapped
// so that each XCD owns a contiguous band of rows.
#include <climits>

#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

// ---------------------------------------------------------------- cell init
// MacGrid::clearCellTypeBuffer (src/MacGrid.cpp:32-50) with optional static solids,
// fused with zeroing the P2G accumulators (src/FluidSolver.cpp:877-886).
__device__ __forceinline__ void cell_init_one(const Geom& g, const CellInit& ci, int c) {
  if (ci.label) {
    const int i = c % g.nx, j = c / g.nx;
    const bool border = (i == 0 || j == 0 || i == g.nx - 1 || j == g.ny - 1);
    uint8_t l = border ? kSolid : kAir;
    if (ci.solid && ci.solid[c]) l = kSolid;
    ci.label[c] = l;
  }
  if (ci.acc) ci.acc[c] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void __launch_bounds__(256) k_cell_init(Geom g, CellInit ci) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int c = t * 256 + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  cell_init_one(g, ci, c);
}

// ------------------------------------------------------- particle scatter pass
// classify: FluidDomain::classifyCells(MarkerParticleSet&) (src/FluidDomain.cpp:
// 150-180): cell = (int)((x / lengthX) * sizeX), clamped, := LIQUID unless the
// cell is SOLID (the reference re-marks border / solid cells afterwards).
// p2g:      transferVelocityToGridSpread splat (src/FluidSolver.cpp:887-897): u and
// weight at (x, y - dy/2), v and weight at (x - dx/2, y), on the pool grids whose
// y spacing is dx (src/FluidSolver.cpp:56-63).  acc = (u sum, u weight, v sum,
// v weight) per face.
//
// The 16 splat terms of a particle go to an LDS copy of the block's footprint
// when the bounding box of the block's 256 particles' splat cells fits in
// kSplatTile cells (particles of one block are usually neighbours: emitters and
// the initial fill are row-major); the tile is then added to the grid with one
// atomic per touched component.  Otherwise every term is a global atomic.  Both
// give the reference's terms; only the order of the float additions differs,
// which atomic accumulation leaves unspecified anyway.  For one domain the particles
// keep their storage order, and the splat reads a copy of them grouped by 16 x 16-cell
// bucket (k_bucket_count / k_bucket_scan / k_bucket_fill, a counting sort each step):
// a block's particles then lie in one or two neighbouring buckets and its footprint
// fits the tile however far the particles have mixed.
// The LDS copy holds 64-bit fixed-point sums (fx_term: rint(t * 2^40)) wherever the
// footprint fits kSplatTileFx cells: LDS integer atomics issue about twice as fast as
// float ones on gfx950 (the developed 1024^2 dam break's splat 99 -> ~50 us).  The
// float form then adds each fixed-point sum to the grid as one float (one rounding
// instead of one per term); the deterministic form adds the integers.  The integer
// adds go to LDS one per lane: summing a quad's equal addresses first (DPP, as the
// float sums do) cost more than it saved on bucket-ordered particles.  A block whose
// terms could overflow the fixed point (|velocity| > kFxMaxVel: 256 terms of at most
// 2^14 stay below 2^23 = 2^63 / 2^40), or is not finite, sums in float instead.
constexpr int kSplatTile = 2048;  // cells (float4 each, 32 KiB of LDS)
constexpr int kSplatTileFx = 1024;  // cells in fixed point (4 x int64 each, 32 KiB)
constexpr float kFxMaxVel = 16384.f;
constexpr int kBucketShift = 4;   // buckets of 16 x 16 cells

// Adds the 8 terms of one splat (value and weight, Grid.h:152-184 order of
// operations) through `add(cell_i, cell_j, component, term)`.
template <typename Add>
__device__ __forceinline__ void splat_terms(const SplatCorners& s, float value, int comp,
                                            Add&& add) {
  const float a0 = (1.0f - s.fj) * value, a1 = s.fj * value;
  const float w0 = (1.0f - s.fj) * 1.0f, w1 = s.fj * 1.0f;
  add(s.i, s.j, comp, (1.0f - s.fi) * a0);
  add(s.i1, s.j, comp, s.fi * a0);
  add(s.i, s.j1, comp, (1.0f - s.fi) * a1);
  add(s.i1, s.j1, comp, s.fi * a1);
  add(s.i, s.j, comp + 1, (1.0f - s.fi) * w0);
  add(s.i1, s.j, comp + 1, s.fi * w0);
  add(s.i, s.j1, comp + 1, (1.0f - s.fi) * w1);
  add(s.i1, s.j1, comp + 1, s.fi * w1);
}

// Particle block -> (geometry, particles, grid rectangle): the whole grid for one
// domain; the block's scene (tiles never straddle scenes) for the packed batch.
struct PartTile {
  Geom g;
  int64_t p;       // this thread's particle
  int64_t np_end;  // end of the block's particle range
  int off, pitch;  // cell (i, j) of the block's grid is off + i + j * pitch
};

template <bool BATCH>
__device__ __forceinline__ PartTile part_tile(const Geom& g0, int64_t np, const BatchTables& tb) {
  PartTile pt;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (BATCH) {
    const int2 e = tb.ptile[t];
    const BatchScene sc = tb.scene[e.x];
    pt.g = sc.g;
    pt.p = sc.part_base + (int64_t)e.y * 256 + threadIdx.x;
    pt.np_end = sc.part_base + sc.np;
    pt.off = sc.cell_off;
    pt.pitch = sc.pitch;
  } else {
    pt.g = g0;
    pt.p = (int64_t)t * 256 + threadIdx.x;
    pt.np_end = np;
    pt.off = 0;
    pt.pitch = g0.nx;
  }
  return pt;
}

// FX (fsg_params.deterministic): every term t is accumulated as the integer
// fx_term(t) = rint(t * 2^40) with 64-bit integer adds (LDS and global), which are
// associative: the sums, and everything computed from them, do not depend on the
// order in which the particles' terms arrive.  Terms of magnitude >= 2^-17 are
// represented exactly (a float has 24 significant bits), so each sum is the exact
// sum of the float terms, rounded once when it is converted back (k_fx_to_float);
// smaller terms are rounded to 2^-40.  Range: |sum| < 2^23 per face component.
template <bool BATCH, bool FX>
__global__ void __launch_bounds__(256) k_particle_scatter(Geom g0, const float4* __restrict__ pt,
                                                          int64_t np, uint8_t* __restrict__ label,
                                                          void* __restrict__ acc_out,
                                                          int do_classify, int do_p2g,
                                                          BatchTables tb) {
  const PartTile T = part_tile<BATCH>(g0, np, tb);
  const Geom& g = T.g;
  const int64_t p = T.p;
  const bool live = p < T.np_end;
  label += T.off;
  float* acc = reinterpret_cast<float*>(acc_out) + 4 * (int64_t)T.off;
  unsigned long long* accx = reinterpret_cast<unsigned long long*>(acc_out) + 4 * (int64_t)T.off;
  const int pitch = T.pitch;
  const float4 q = live ? pt[p] : make_float4(0.f, 0.f, 0.f, 0.f);
  if (do_classify && live) {
    int x = (int)((q.x / g.len_x) * (float)g.nx);
    int y = (int)((q.y / g.len_y) * (float)g.ny);
    x = clampi(x, 0, g.nx - 1);
    y = clampi(y, 0, g.ny - 1);
    const int c = x + y * pitch;
    if (label[c] != kSolid) label[c] = kLiquid;
  }
  if (!do_p2g) return;
  const int nx = g.nx, ny = g.ny;
  // pool spacing: (dx, dx)
  const float yu = q.y - g.dy * 0.5f;
  const float xv = q.x - g.dx * 0.5f;
  const SplatCorners su = splat_corners(div_x(q.x, g), div_x(yu, g), nx, ny);
  const SplatCorners sv = splat_corners(div_x(xv, g), div_x(q.y, g), nx, ny);
  // footprint of the block: min / max corner over live particles
  int i0 = live ? min(su.i, sv.i) : INT_MAX, j0 = live ? min(su.j, sv.j) : INT_MAX;
  int i1 = live ? max(su.i1, sv.i1) : INT_MIN, j1 = live ? max(su.j1, sv.j1) : INT_MIN;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    i0 = min(i0, __shfl_xor(i0, off, 64));
    j0 = min(j0, __shfl_xor(j0, off, 64));
    i1 = max(i1, __shfl_xor(i1, off, 64));
    j1 = max(j1, __shfl_xor(j1, off, 64));
  }
  __shared__ int s_box[4][4];
  // the block's LDS copy of its footprint: float4 per cell, or 4 x int64 (FX)
  __shared__ float4 s_tile[kSplatTile];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    s_box[0][wave] = i0;
    s_box[1][wave] = j0;
    s_box[2][wave] = i1;
    s_box[3][wave] = j1;
  }
  __syncthreads();
  i0 = min(min(s_box[0][0], s_box[0][1]), min(s_box[0][2], s_box[0][3]));
  j0 = min(min(s_box[1][0], s_box[1][1]), min(s_box[1][2], s_box[1][3]));
  i1 = max(max(s_box[2][0], s_box[2][1]), max(s_box[2][2], s_box[2][3]));
  j1 = max(max(s_box[3][0], s_box[3][1]), max(s_box[3][2], s_box[3][3]));
  const int bw = i1 - i0 + 1, bh = j1 - j0 + 1;
  const bool tiled = bw > 0 && bh > 0 && (int64_t)bw * bh <= (FX ? kSplatTileFx : kSplatTile);
  // fixed-point LDS sums (the float form: where the footprint fits and the block's
  // velocities are in range; block-uniform)
  bool fixed = FX;
  if (!FX && tiled && (int64_t)bw * bh <= kSplatTileFx) {
    const bool bad = live && !(fabsf(q.z) <= kFxMaxVel && fabsf(q.w) <= kFxMaxVel);
    fixed = !__syncthreads_or(bad);
  }
  if (!tiled) {
    if (!live) return;
    auto gadd = [&](int i, int j, int comp, float v) {
      if (FX)
        atomicAdd(accx + 4 * (i + j * pitch) + comp, (unsigned long long)fx_term(v));
      else
        unsafeAtomicAdd(acc + 4 * (i + j * pitch) + comp, v);
    };
    splat_terms(su, q.z, 0, gadd);
    splat_terms(sv, q.w, 2, gadd);
    return;
  }
  const int area = bw * bh;
  if (fixed) {
    unsigned long long* tile = reinterpret_cast<unsigned long long*>(s_tile);
    for (int c = threadIdx.x; c < 4 * area; c += 256) tile[c] = 0ull;
    __syncthreads();
    auto ladd = [&](int i, int j, int comp, float v) {
      const long long t = fx_term(v);
      if (live && t != 0ll)
        atomicAdd(tile + 4 * ((i - i0) + (j - j0) * bw) + comp, (unsigned long long)t);
    };
    splat_terms(su, q.z, 0, ladd);
    splat_terms(sv, q.w, 2, ladd);
    __syncthreads();
    for (int c = threadIdx.x; c < 4 * area; c += 256) {
      const unsigned long long v = tile[c];
      const int cell = c >> 2;
      const int64_t dst = 4 * ((i0 + cell % bw) + (int64_t)(j0 + cell / bw) * pitch) + (c & 3);
      if (v != 0ull) {
        if (FX)
          atomicAdd(accx + dst, v);
        else
          unsafeAtomicAdd(acc + dst, fx_value((long long)v));
      }
    }
    return;
  }
  for (int c = threadIdx.x; c < area; c += 256) s_tile[c] = make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();
  {
    // every lane takes part (the quad exchange needs the whole wave); a lane past the
    // end adds nothing
    float* tile = reinterpret_cast<float*>(s_tile);
    auto ladd = [&](int i, int j, int comp, float v) {
      lds_add_quad(tile, live ? 4 * ((i - i0) + (j - j0) * bw) + comp : -1, live ? v : 0.f);
    };
    splat_terms(su, q.z, 0, ladd);
    splat_terms(sv, q.w, 2, ladd);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < area; c += 256) {
    const float4 v = s_tile[c];
    float* dst = acc + 4 * ((i0 + c % bw) + (j0 + c / bw) * pitch);
    if (v.x != 0.f) unsafeAtomicAdd(dst + 0, v.x);
    if (v.y != 0.f) unsafeAtomicAdd(dst + 1, v.y);
    if (v.z != 0.f) unsafeAtomicAdd(dst + 2, v.z);
    if (v.w != 0.f) unsafeAtomicAdd(dst + 3, v.w);
  }
}

// -------------------------------------------------------------- face update
// One thread per cell (its u and v face), applying in order the per-face phases
// selected by `flags` to the buffer pair (u, v) in place:
//   FU_P2G      back := sum / weight where weight > 1e-6 (src/FluidSolver.cpp:898-917)
//   FU_PREV     prev := value (src/MacGrid.cpp:52-56)
//   FU_GRAVITY  value += a*dt at LIQUID cells (src/FluidSolver.cpp:286-291)
//   FU_DIRICHLET zero flow into SOLID (src/FluidSolver.cpp:300-320); with FU_CLOSED
//               zero every face next to SOLID (fsg_params.closed_walls)
//   FU_EXTINIT  extension init (src/FluidSolver.cpp:490-530): ext := value on valid
//               faces else 0; masks; value.u := 0 where u- OR v-invalid (the v
//               branch writes the x buffer, :527)
// The weight test `w > 0.000001` compares the float weight with a double constant;
// the float test w > 1e-6f selects the same floats (no float lies in between).
// dec (optional, fsg_params.record_decisions): the step's discrete decisions per cell,
// bit 0 / 1 u / v normalised by P2G, bit 2 / 3 u / v zeroed by this Dirichlet pass.
__global__ void __launch_bounds__(256) k_face_update(
    Geom g, int flags, float* __restrict__ u, float* __restrict__ v,
    const float4* __restrict__ acc, float* __restrict__ prev_u, float* __restrict__ prev_v,
    const uint8_t* __restrict__ label, float gx_dt, float gy_dt, float* __restrict__ ext_u,
    float* __restrict__ ext_v, uint8_t* __restrict__ mxf, uint8_t* __restrict__ mxb,
    uint8_t* __restrict__ myf, uint8_t* __restrict__ myb, uint8_t* __restrict__ info_u,
    uint8_t* __restrict__ info_v, uint8_t* __restrict__ dec) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int c = t * 256 + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int i = c % g.nx, j = c / g.nx;
  float uu = u[c], vv = v[c];
  unsigned dbits = 0;
  if (flags & FU_P2G) {
    const float4 a = acc[c];
    if (a.y > 0.000001f) {
      uu = a.x / a.y;
      dbits |= 1;
    }
    if (a.w > 0.000001f) {
      vv = a.z / a.w;
      dbits |= 2;
    }
  }
  if (flags & FU_PREV) {
    prev_u[c] = uu;
    prev_v[c] = vv;
  }
  const uint8_t L = label[c];
  const uint8_t Lim = label[clampi(i - 1, 0, g.nx - 1) + j * g.nx];
  const uint8_t Ljm = label[i + clampi(j - 1, 0, g.ny - 1) * g.nx];
  if ((flags & FU_GRAVITY) && L == kLiquid) {
    uu = uu + gx_dt;
    vv = vv + gy_dt;
  }
  if (flags & FU_DIRICHLET) {
    bool zu, zv;
    if (flags & FU_CLOSED) {
      zu = Lim == kSolid || L == kSolid;
      zv = Ljm == kSolid || L == kSolid;
    } else {
      zu = (Lim == kSolid && uu < 0.f) || (L == kSolid && uu > 0.f);
      zv = (Ljm == kSolid && vv < 0.f) || (L == kSolid && vv > 0.f);
    }
    if (zu) uu = 0.f;
    if (zv) vv = 0.f;
    dbits |= (zu ? 4u : 0u) | (zv ? 8u : 0u);
  }
  if (dec) dec[c] = (uint8_t)dbits;
  if (flags & FU_EXTINIT) {
    const bool vu = (L == kLiquid || Lim == kLiquid);
    const bool vv_ok = (L == kLiquid || Ljm == kLiquid);
    ext_u[c] = vu ? uu : 0.f;
    ext_v[c] = vv_ok ? vv : 0.f;
    mxf[c] = vu; mxb[c] = vu;
    myf[c] = vv_ok; myb[c] = vv_ok;
    if (info_u) info_u[c] = vu ? 0 : 0xF0;
    if (info_v) info_v[c] = vv_ok ? 0 : 0xF0;
    if (!vu || !vv_ok) uu = 0.f;
  }
  u[c] = uu;
  v[c] = vv;
}

// ------------------------------------------------------------ extension sweep
// One sweep of extendVelocityIndividual (src/FluidSolver.cpp:533-617): an invalid
// face whose two cells are not SOLID takes the mean of the back-buffer values of
// its valid neighbours, summed (i-1,j), (i,j-1), (i,j+1), (i+1,j), and is marked
// valid in the back mask.  Reads touch only front-valid faces and writes only
// front-invalid ones, so the sweep is race-free in place.  The interior guard only
// matters for labels whose border is not SOLID (the reference would assert).
__global__ void __launch_bounds__(256) k_extend_sweep(
    Geom g, const uint8_t* __restrict__ label, float* __restrict__ ub, float* __restrict__ vb,
    const uint8_t* __restrict__ mxf, uint8_t* __restrict__ mxb, const uint8_t* __restrict__ myf,
    uint8_t* __restrict__ myb, uint8_t* __restrict__ info_u, uint8_t* __restrict__ info_v,
    int layer) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int c = t * 256 + threadIdx.x;
  const int nx = g.nx;
  if (c >= nx * g.ny) return;
  const int i = c % nx, j = c / nx;
  if (i < 1 || j < 1 || i > nx - 2 || j > g.ny - 2) return;
  // every load first (one memory round trip): the neighbours' values are used only where
  // their front mask is 1, and this sweep writes only faces whose front mask is 0
  const uint8_t L = label[c], Lim = label[c - 1], Ljm = label[c - nx];
  const uint8_t mx0 = mxf[c], mx1 = mxf[c - 1], mx2 = mxf[c - nx], mx4 = mxf[c + nx],
                mx8 = mxf[c + 1];
  const uint8_t my0 = myf[c], my1 = myf[c - 1], my2 = myf[c - nx], my4 = myf[c + nx],
                my8 = myf[c + 1];
  const float u1 = ub[c - 1], u2 = ub[c - nx], u4 = ub[c + nx], u8 = ub[c + 1];
  const float v1 = vb[c - 1], v2 = vb[c - nx], v4 = vb[c + nx], v8 = vb[c + 1];
  if (L == kSolid) return;
  if (mx0 == 0 && Lim != kSolid) {
    float acc = 0.f;
    int n = 0, bits = 0;
    if (mx1 == 1) { acc += u1; n++; bits |= 1; }
    if (mx2 == 1) { acc += u2; n++; bits |= 2; }
    if (mx4 == 1) { acc += u4; n++; bits |= 4; }
    if (mx8 == 1) { acc += u8; n++; bits |= 8; }
    if (n > 0) {
      acc /= (float)n;
      ub[c] = acc;
      mxb[c] = 1;
      if (info_u) info_u[c] = (uint8_t)((layer << 4) | bits);
    }
  }
  if (my0 == 0 && Ljm != kSolid) {
    float acc = 0.f;
    int n = 0, bits = 0;
    if (my1 == 1) { acc += v1; n++; bits |= 1; }
    if (my2 == 1) { acc += v2; n++; bits |= 2; }
    if (my4 == 1) { acc += v4; n++; bits |= 4; }
    if (my8 == 1) { acc += v8; n++; bits |= 8; }
    if (n > 0) {
      acc /= (float)n;
      vb[c] = acc;
      myb[c] = 1;
      if (info_v) info_v[c] = (uint8_t)((layer << 4) | bits);
    }
  }
}

// -------------------------------------------------------- level-set classify
// FluidDomain::classifyCells(LevelSet&) (src/FluidDomain.cpp:129-148).
__global__ void __launch_bounds__(256) k_classify_levelset(Geom g, const float* __restrict__ phi,
                                                           const uint8_t* __restrict__ solid,
                                                           uint8_t* __restrict__ label) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int c = t * 256 + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int i = c % g.nx, j = c / g.nx;
  uint8_t l = phi[c] < 0.f ? kLiquid : kAir;
  const bool border = (i == 0 || j == 0 || i == g.nx - 1 || j == g.ny - 1);
  if (border || (solid && solid[c])) l = kSolid;
  label[c] = l;
}

// ------------------------------------------------------------ pressure update
// pressureSolve's gradient update (src/FluidSolver.cpp:427-482) fused with the
// following enforceDirichlet (:240) and updateVelocityDiffBuffer (src/MacGrid.cpp:
// 58-70).  A = front (extended) velocity, B = back; the host swaps afterwards.
//  F > 0: B := A - dt/density * (p - p_left) / dx at faces where (i,j), (i-1,j)
//         or (i,j-1) is LIQUID, p = x for LIQUID cells and 0 otherwise; other
//         faces keep B.  Then Dirichlet on B (the front after the swap):
//         do_dirichlet 1 the reference's one-sided rule, 2 closed walls.
//  F = 0: the reference returns before solving and does not swap (:347-350); the
//         same state is produced by exchanging A and B here (B := Dirichlet(A),
//         A := B) ahead of the host's swap.
// xg is zeroed at cells that are not LIQUID, so that it can seed the next solve
// (each thread writes only its own cell, and only where no thread reads it).
// x64 != null (the multigrid form's FP64 solution): the pressure differences are
// formed in FP64 and the update rounded once to float.  The pressure of a large
// domain is ~1e6 while its gradient is O(1), so differencing FP32 pressures would
// cost ~1e-4 relative accuracy that the FP64 solve has.
__global__ void __launch_bounds__(256) k_pressure_update(
    Geom g, const int* __restrict__ n_fluid, const uint8_t* __restrict__ label,
    float* __restrict__ xg, double* __restrict__ x64, const int* __restrict__ counts,
    float ppk, float* __restrict__ au, float* __restrict__ av,
    float* __restrict__ bu, float* __restrict__ bv, float dt_over_density, int do_dirichlet,
    int do_diff, const float* __restrict__ prev_u, const float* __restrict__ prev_v,
    float* __restrict__ diff_u, float* __restrict__ diff_v, uint8_t* __restrict__ zmask,
    uint8_t* __restrict__ dec, MgFinishFuse fin) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int c = t * 256 + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int i = c % g.nx, j = c / g.nx;
  const int cim = clampi(i - 1, 0, g.nx - 1) + j * g.nx;
  const int cjm = i + clampi(j - 1, 0, g.ny - 1) * g.nx;
  const uint8_t L = label[c], Lim = label[cim], Ljm = label[cjm];
  const float a_u = au[c], a_v = av[c];
  // fused finish: the solution is x[iter & 1] at the unknowns.  x64 may be that array
  // itself: a thread writes only its own cell, and reads its neighbours' only where they
  // are unknowns, whose values it leaves unchanged.
  const double* xs = x64;
  if (fin.inf0) {
    xs = (fin.sc->iter & 1) ? fin.x1 : x64;
    const double v = (fin.inf0[c] & kMgUnknown) ? xs[c] : 0.0;
    x64[c] = v;
    xg[c] = (float)v;
  }
  auto pval = [&](int cc) -> double {
    return !fin.inf0 || (fin.inf0[cc] & kMgUnknown) ? xs[cc] : 0.0;
  };
  float nu, nv;
  if (*n_fluid > 0) {
    nu = bu[c];
    nv = bv[c];
    // particle pressure k * (n > 9 ? n - 8 : 0) (src/FluidSolver.cpp:441-453); the
    // reference's k is 0 (then the term is + 0, as there)
    float pp = 0.f, pp_im = 0.f, pp_jm = 0.f;
    if (counts) {
      const int n0 = counts[c], n1 = counts[cim], n2 = counts[cjm];
      pp = ppk * (float)(n0 > 9 ? n0 - 8 : 0);
      pp_im = ppk * (float)(n1 > 9 ? n1 - 8 : 0);
      pp_jm = ppk * (float)(n2 > 9 ? n2 - 8 : 0);
    }
    if (x64 && (L == kLiquid || Lim == kLiquid || Ljm == kLiquid)) {
      const double p = L == kLiquid ? pval(c) + (double)pp : 0.0;
      const double p_im = Lim == kLiquid ? pval(cim) + (double)pp_im : 0.0;
      const double p_jm = Ljm == kLiquid ? pval(cjm) + (double)pp_jm : 0.0;
      nu = (float)((double)a_u - (double)dt_over_density * (p - p_im) / (double)g.dx);
      nv = (float)((double)a_v - (double)dt_over_density * (p - p_jm) / (double)g.dy);
    } else if (L == kLiquid || Lim == kLiquid || Ljm == kLiquid) {
      const float p = L == kLiquid ? xg[c] + pp : 0.f;
      const float p_im = Lim == kLiquid ? xg[cim] + pp_im : 0.f;
      const float p_jm = Ljm == kLiquid ? xg[cjm] + pp_jm : 0.f;
      const float pdx = p - p_im;
      const float pdy = p - p_jm;
      nu = a_u - dt_over_density * pdx / g.dx;
      nv = a_v - dt_over_density * pdy / g.dy;
    }
  } else {
    nu = a_u;
    nv = a_v;
    au[c] = bu[c];
    av[c] = bv[c];
  }
  if (L != kLiquid) xg[c] = 0.f;
  bool zu = false, zv = false;
  if (do_dirichlet == 2) {  // closed walls: no flow through any face next to SOLID
    zu = Lim == kSolid || L == kSolid;
    zv = Ljm == kSolid || L == kSolid;
  } else if (do_dirichlet) {
    zu = (Lim == kSolid && nu < 0.f) || (L == kSolid && nu > 0.f);
    zv = (Ljm == kSolid && nv < 0.f) || (L == kSolid && nv > 0.f);
  }
  if (zu) nu = 0.f;
  if (zv) nv = 0.f;
  // the adjoint's record of the Dirichlet decisions (bit 0: u zeroed, bit 1: v)
  if (zmask) zmask[c] = (uint8_t)((zu ? 1 : 0) | (zv ? 2 : 0));
  // the step's decision record: bit 4 / 5 u / v zeroed by this (second) Dirichlet pass
  if (dec) dec[c] = (uint8_t)(dec[c] | (zu ? 16 : 0) | (zv ? 32 : 0));
  bu[c] = nu;
  bv[c] = nv;
  if (do_diff) {
    diff_u[c] = nu - prev_u[c];
    diff_v[c] = nv - prev_v[c];
  }
}

// ------------------------------------------------ averaging extension
// FluidSolver::extendVelocityAvarageing (src/FluidSolver.cpp:625-707), which the
// reference's steps do not call.  Init: one mask per cell (the x pair), 1 at LIQUID
// cells in both; back := front.
__global__ void __launch_bounds__(256) k_extavg_init(Geom g, const uint8_t* __restrict__ label,
                                                     const float* __restrict__ uf,
                                                     const float* __restrict__ vf,
                                                     float* __restrict__ ub, float* __restrict__ vb,
                                                     uint8_t* __restrict__ mf,
                                                     uint8_t* __restrict__ mb) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const uint8_t m = label[c] == kLiquid ? 1 : 0;
  mf[c] = m;
  mb[c] = m;
  ub[c] = uf[c];
  vb[c] = vf[c];
}

// One sweep.  The reference visits cells in raster order and writes into the back
// buffer it reads, so a cell sees the writes of cells before it.  Cell (i, j) reads
// faces that only (i-2, j), (i-1, j-1), (i+1, j-1) and (i, j-2) write earlier, and
// no two cells on a line i + 2j = t touch each other's faces, so sweeping the lines
// t = 0, 1, ... in order, the cells of a line in parallel, reproduces the raster
// result exactly.  One workgroup; the barrier between lines orders the global
// memory writes within it.
__global__ void __launch_bounds__(1024) k_extavg_sweep(Geom g, const uint8_t* __restrict__ label,
                                                       float* __restrict__ ub,
                                                       float* __restrict__ vb,
                                                       const uint8_t* __restrict__ mf,
                                                       uint8_t* __restrict__ mb) {
  const int nx = g.nx, ny = g.ny;
  const int tmax = (nx - 1) + 2 * (ny - 1);
  for (int t = 0; t <= tmax; ++t) {
    for (int j = threadIdx.x; j < ny; j += 1024) {
      const int i = t - 2 * j;
      if (i < 0 || i >= nx) continue;
      const int c = i + j * nx;
      if (mf[c] != 0 || label[c] == kSolid) continue;
      float su = 0.f, sv = 0.f;
      int n = 0;
      const int ni[4] = {i - 1, i, i, i + 1}, nj[4] = {j, j - 1, j + 1, j};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int a = ni[q], b = nj[q];
        if (a < 0 || a >= nx || b < 0 || b >= ny) continue;
        const int k = a + b * nx;
        if (mf[k] != 1) continue;
        const float u1 = a + 1 < nx ? ub[k + 1] : ub[k];
        const float v1 = b + 1 < ny ? vb[k + nx] : vb[k];
        su += (ub[k] + u1) / 2.0f;
        sv += (vb[k] + v1) / 2.0f;
        ++n;
      }
      if (n > 0) {
        su /= (float)n;
        sv /= (float)n;
        ub[c] = su;
        if (i + 1 < nx) ub[c + 1] = su;
        vb[c] = sv;
        if (j + 1 < ny) vb[c + nx] = sv;
        mb[c] = 1;
      }
    }
    __syncthreads();
  }
}

void launch_extend_average(hipStream_t s, const Geom& g, const uint8_t* label, const float* uf,
                           const float* vf, float* ub, float* vb, uint8_t* mf, uint8_t* mb,
                           int init) {
  if (init) {
    k_extavg_init<<<(g.nx * g.ny + 255) / 256, 256, 0, s>>>(g, label, uf, vf, ub, vb, mf, mb);
    return;
  }
  k_extavg_sweep<<<1, 1024, 0, s>>>(g, label, ub, vb, mf, mb);
}

// ----------------------------------------------- particle-pressure counts
// Particles per cell for the volume term of the pressure update
// (src/FluidSolver.cpp:352-360): cell = (int)((x / lengthX) * sizeX), the
// reference's unclamped index clamped into the grid.
__global__ void __launch_bounds__(256) k_count_particles(Geom g, const float4* __restrict__ pt,
                                                         int64_t np, int* __restrict__ counts) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= np) return;
  const float4 q = pt[p];
  int x = (int)((q.x / g.len_x) * (float)g.nx);
  int y = (int)((q.y / g.len_y) * (float)g.ny);
  x = clampi(x, 0, g.nx - 1);
  y = clampi(y, 0, g.ny - 1);
  atomicAdd(counts + x + y * g.nx, 1);
}

// ---------------------------------------------------- FluidSource lattice
// FluidSource::update (src/FluidDomain.cpp:29-52) appended on the device: particle
// base + j*nxp + i = (xs[i], ys[j], vx, vy).  xs and ys are the emitter's MyFloat
// accumulations (x += delta/2.5 while x < max), formed on the host because each
// value is the result of a sequence of float additions.
__global__ void __launch_bounds__(256) k_emit_lattice(float4* __restrict__ pt, int64_t base,
                                                      const float* __restrict__ xs, int nxp,
                                                      const float* __restrict__ ys, int nyp,
                                                      float vx, float vy) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)nxp * nyp) return;
  const int i = (int)(t % nxp), j = (int)(t / nxp);
  pt[base + t] = make_float4(xs[i], ys[j], vx, vy);
}

// ------------------------------------------------------- level-set advection
// FluidDomain::advectLevelSet (src/FluidDomain.cpp:104-127): forward Euler with the
// reference's factor 5, first-order upwind gradients (LevelSet::
// computeUpwindGradientX/Y, src/LevelSet.cpp:23-59) and the cell-centred velocity
// (u(i,j) + u(i+1,j)) / 2 (MacGrid::velX, include/MacGrid.h:29-38).  Differences:
// the face past the last column / row, which the reference reads out of range (an
// assert in its -g build), is clamped; the reference's loops run j over sizeX and
// i over sizeY, so on a non-square grid only the min(nx, ny) square is computed and
// the rest of the new level set keeps its fresh value 0.
__global__ void __launch_bounds__(256) k_advect_level_set(Geom g, const float* __restrict__ phi,
                                                          float* __restrict__ out,
                                                          const float* __restrict__ u,
                                                          const float* __restrict__ v, float dt) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  const int m = nx < ny ? nx : ny;
  if (i >= m || j >= m) {
    out[c] = 0.f;
    return;
  }
  const int i1 = i + 1 < nx ? i + 1 : nx - 1;
  const int j1 = j + 1 < ny ? j + 1 : ny - 1;
  const float vel_x = (u[c] + u[i1 + j * nx]) / 2.0f;
  const float vel_y = (v[c] + v[i + j1 * nx]) / 2.0f;
  const int ip = clampi(i + 1, 0, nx - 1), im = clampi(i - 1, 0, nx - 1);
  const int jp = clampi(j + 1, 0, ny - 1), jm = clampi(j - 1, 0, ny - 1);
  const float p0 = phi[c];
  const float gx = vel_x < 0.f ? phi[ip + j * nx] - p0 : p0 - phi[im + j * nx];
  const float gy = vel_y < 0.f ? phi[i + jp * nx] - p0 : p0 - phi[i + jm * nx];
  const float rate = -(vel_x * gx + vel_y * gy);
  out[c] = p0 + rate * dt * 5.0f;
}

// ------------------------------------------------------------ diff (alone)
__global__ void __launch_bounds__(256) k_diff(Geom g, const float* __restrict__ u,
                                              const float* __restrict__ v,
                                              const float* __restrict__ pu,
                                              const float* __restrict__ pv,
                                              float* __restrict__ du, float* __restrict__ dv) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int c = t * 256 + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  du[c] = u[c] - pu[c];
  dv[c] = v[c] - pv[c];
}

// ------------------------------------------------------ particle gather pass
// G2P (src/FluidSolver.cpp:921-963) + particle motion (src/MarkerParticleSet.cpp:
// 40-62).  g2p: 0 none, 1 PIC, 2 FLIP, 3 PICFLIP; move: 0 none, 1 advect,
// 2 advect and step back out of SOLID cells, tested at (int)(pos / delta).
template <bool BATCH>
__global__ void __launch_bounds__(256) k_particle_gather(
    Geom g0, float4* __restrict__ pt, int64_t np, const float* __restrict__ u,
    const float* __restrict__ v, const float* __restrict__ du, const float* __restrict__ dv,
    float pic_ratio, int g2p, int move, float dt, const uint8_t* __restrict__ label,
    uint8_t* __restrict__ reject, BatchTables tb) {
  const PartTile T = part_tile<BATCH>(g0, np, tb);
  const Geom& g = T.g;
  const int64_t p = T.p;
  if (p >= T.np_end) return;
  const int pitch = T.pitch;
  u += T.off;
  v += T.off;
  du += T.off;
  dv += T.off;
  label += T.off;
  float4 q = pt[p];
  if (g2p) {
    const float xs_u = div_x(q.x, g), ys_u = div_y(q.y - g.dy * 0.5f, g);
    const float xs_v = div_x(q.x - g.dx * 0.5f, g), ys_v = div_y(q.y, g);
    if (g2p == 1) {
      q.z = interp_pitched(u, g.nx, g.ny, pitch, xs_u, ys_u);
      q.w = interp_pitched(v, g.nx, g.ny, pitch, xs_v, ys_v);
    } else if (g2p == 2) {
      q.z = q.z + interp_pitched(du, g.nx, g.ny, pitch, xs_u, ys_u);
      q.w = q.w + interp_pitched(dv, g.nx, g.ny, pitch, xs_v, ys_v);
    } else {
      const float pu = interp_pitched(u, g.nx, g.ny, pitch, xs_u, ys_u);
      const float pv = interp_pitched(v, g.nx, g.ny, pitch, xs_v, ys_v);
      const float fu = q.z + interp_pitched(du, g.nx, g.ny, pitch, xs_u, ys_u);
      const float fv = q.w + interp_pitched(dv, g.nx, g.ny, pitch, xs_v, ys_v);
      q.z = pu * pic_ratio + fu * (1.0f - pic_ratio);
      q.w = pv * pic_ratio + fv * (1.0f - pic_ratio);
    }
  }
  if (move) {
    q.x += q.z * dt;
    q.y += q.w * dt;
    uint8_t flag = 0;
    if (move == 2) {
      int x = (int)div_x(q.x, g);
      int y = (int)div_y(q.y, g);
      x = clampi(x, 0, g.nx - 1);
      y = clampi(y, 0, g.ny - 1);
      if (label[x + y * pitch] == kSolid) {
        q.x += q.z * -dt;
        q.y += q.w * -dt;
        flag = 1;
      }
    }
    if (reject) reject[p] = flag;
  }
  pt[p] = q;
}

// ---------------------------------------------------- semi-Lagrangian advection
// Reference scatter form, advectVelocitySemiLagrangian (src/FluidSolver.cpp:
// 709-772): for each face next to LIQUID, backtrace its position by -dt through the
// front field, sample the face's own velocity at its own position and splat it
// into the back buffer (zeroed beforehand) at the backtraced point.  No swap.
// FX (fsg_params.deterministic): the terms are summed as 2^-40 fixed point in fx
// (u sums at [0, C), v sums at [C, 2C)), converted to the back buffer afterwards, so
// the sums do not depend on the order of the atomics.
template <bool FX>
__device__ __forceinline__ void sl_splat(float* g, unsigned long long* fx, int nx, int ny,
                                         float xs, float ys, float value) {
  if (!FX) {
    splat_scaled(g, nx, ny, xs, ys, value);
    return;
  }
  const SplatCorners s = splat_corners(xs, ys, nx, ny);
  const float v0 = (1.0f - s.fj) * value, v1 = s.fj * value;
  atomicAdd(fx + s.i + s.j * nx, (unsigned long long)fx_term((1.0f - s.fi) * v0));
  atomicAdd(fx + s.i1 + s.j * nx, (unsigned long long)fx_term(s.fi * v0));
  atomicAdd(fx + s.i + s.j1 * nx, (unsigned long long)fx_term((1.0f - s.fi) * v1));
  atomicAdd(fx + s.i1 + s.j1 * nx, (unsigned long long)fx_term(s.fi * v1));
}

template <bool FX>
__global__ void __launch_bounds__(256) k_sl_scatter(Geom g, const uint8_t* __restrict__ label,
                                                    const float* __restrict__ uf,
                                                    const float* __restrict__ vf,
                                                    float* __restrict__ ub, float* __restrict__ vb,
                                                    unsigned long long* __restrict__ fx,
                                                    int integrator, float dt) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int c = t * 256 + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int i = c % g.nx, j = c / g.nx;
  const uint8_t L = label[c];
  const uint8_t Lim = label[clampi(i - 1, 0, g.nx - 1) + j * g.nx];
  const uint8_t Ljm = label[i + clampi(j - 1, 0, g.ny - 1) * g.nx];
  const VelField f{uf, vf, g};
  if (L == kLiquid || Lim == kLiquid) {
    const float xp = (float)i * g.dx;
    const float yp = ((float)j + 0.5f) * g.dy;
    float ddx, ddy;
    ode_disp(f, integrator, xp, yp, -dt, ddx, ddy);
    const float xb = xp + ddx, yb = yp + ddy;
    const float val = vel_x_at(uf, g, xp, yp);
    sl_splat<FX>(ub, fx, g.nx, g.ny, div_x(xb, g), div_y(yb - g.dy * 0.5f, g), val);
  }
  if (L == kLiquid || Ljm == kLiquid) {
    const float xp = ((float)i + 0.5f) * g.dx;
    const float yp = (float)j * g.dy;
    float ddx, ddy;
    ode_disp(f, integrator, xp, yp, -dt, ddx, ddy);
    const float xb = xp + ddx, yb = yp + ddy;
    const float val = vel_y_at(vf, g, xp, yp);
    sl_splat<FX>(vb, fx ? fx + (int64_t)g.nx * g.ny : nullptr, g.nx, g.ny,
                 div_x(xb - g.dx * 0.5f, g), div_y(yb, g), val);
  }
}

// Gather form (new): back := front component sampled at the backtraced face
// position for faces next to LIQUID, back := front elsewhere; the host swaps.
__global__ void __launch_bounds__(256) k_sl_gather(Geom g, const uint8_t* __restrict__ label,
                                                   const float* __restrict__ uf,
                                                   const float* __restrict__ vf,
                                                   float* __restrict__ ub, float* __restrict__ vb,
                                                   int integrator, float dt) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int c = t * 256 + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int i = c % g.nx, j = c / g.nx;
  const uint8_t L = label[c];
  const uint8_t Lim = label[clampi(i - 1, 0, g.nx - 1) + j * g.nx];
  const uint8_t Ljm = label[i + clampi(j - 1, 0, g.ny - 1) * g.nx];
  const VelField f{uf, vf, g};
  float nu = uf[c], nv = vf[c];
  if (L == kLiquid || Lim == kLiquid) {
    const float xp = (float)i * g.dx;
    const float yp = ((float)j + 0.5f) * g.dy;
    float ddx, ddy;
    ode_disp(f, integrator, xp, yp, -dt, ddx, ddy);
    nu = vel_x_at(uf, g, xp + ddx, yp + ddy);
  }
  if (L == kLiquid || Ljm == kLiquid) {
    const float xp = ((float)i + 0.5f) * g.dx;
    const float yp = (float)j * g.dy;
    float ddx, ddy;
    ode_disp(f, integrator, xp, yp, -dt, ddx, ddy);
    nv = vel_y_at(vf, g, xp + ddx, yp + ddy);
  }
  ub[c] = nu;
  vb[c] = nv;
}

// FluidSolver::advectParticlesWithGrid (src/FluidSolver.cpp:774-791).
__global__ void __launch_bounds__(256) k_advect_grid(Geom g, float4* __restrict__ pt, int64_t np,
                                                     const float* __restrict__ uf,
                                                     const float* __restrict__ vf,
                                                     int integrator, float dt) {
  const int64_t t = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t p = t * 256 + threadIdx.x;
  if (p >= np) return;
  float4 q = pt[p];
  const VelField f{uf, vf, g};
  float ddx, ddy;
  ode_disp(f, integrator, q.x, q.y, dt, ddx, ddy);
  q.x = q.x + ddx;
  q.y = q.y + ddy;
  pt[p] = q;
}

// ---------------------------------------------------------------- P2G bucket order
__device__ __forceinline__ int particle_bucket(const Geom& g, float4 q, int bx) {
  int x = (int)((q.x / g.len_x) * (float)g.nx);
  int y = (int)((q.y / g.len_y) * (float)g.ny);
  x = clampi(x, 0, g.nx - 1);
  y = clampi(y, 0, g.ny - 1);
  return (x >> kBucketShift) + (y >> kBucketShift) * bx;
}

// Particles per bucket; each particle keeps its slot in its bucket.  The block counts
// its particles in an LDS histogram of every bucket first (slots within the block),
// then takes one range per bucket it touches from the global count: one global atomic
// per (block, bucket) instead of one per particle on a few hot counters.
// The slots within a block follow the particles' storage order (a stable rank): each
// wave groups its lanes by bucket with ballots (one round per distinct bucket among its
// lanes), and the block's four waves take their ranges of each bucket in wave order.
// Particles that were neighbours in storage stay neighbours in the copy, which keeps
// the splat's quad aggregation effective on particles that have not mixed yet.
// BATCH: a block is one particle tile of one scene (its own geometry); the histogram
// holds that scene's buckets, which start at its bucket_base in cnt.  A single domain's
// block takes PPT rounds of 256 consecutive particles (the histogram's clearing and
// flush shared by PPT x 256 particles), and with ci set also the cell set-up of
// k_cell_init, its cells shared out over the blocks (the splat reads both afterwards).
template <bool BATCH, int PPT>
__global__ void __launch_bounds__(256) k_bucket_count(Geom g0, const float4* __restrict__ pt,
                                                      int np, int bx0, int nb0,
                                                      int* __restrict__ cnt,
                                                      int* __restrict__ rank, BatchTables tb,
                                                      CellInit ci) {
  static_assert(!BATCH || PPT == 1, "a batch block is one particle tile");
  extern __shared__ int s_h[];  // nb ints
  int bx = bx0, nb = nb0, base = 0;
  Geom g = g0;
  int64_t p0, np_end;
  if (BATCH) {
    const PartTile T = part_tile<true>(g0, np, tb);
    const BatchScene sc = tb.scene[tb.ptile[xcd_tile(blockIdx.x, gridDim.x)].x];
    bx = sc.bucket_bx;
    nb = sc.bucket_bx * ((sc.g.ny + (1 << kBucketShift) - 1) >> kBucketShift);
    base = sc.bucket_base;
    g = T.g;
    p0 = T.p - threadIdx.x;
    np_end = T.np_end;
  } else {
    p0 = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * (256 * PPT);
    np_end = np;
    if (ci.label || ci.acc) {
      const int C = g0.nx * g0.ny;
      for (int c = blockIdx.x * 256 + threadIdx.x; c < C; c += gridDim.x * 256)
        cell_init_one(g0, ci, c);
    }
  }
  for (int e = threadIdx.x; e < nb; e += 256) s_h[e] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  int b[PPT], local[PPT];
#pragma unroll
  for (int r = 0; r < PPT; ++r) {
    const int64_t p = p0 + r * 256 + threadIdx.x;
    const bool live = p < np_end;
    b[r] = live ? particle_bucket(g, pt[p], bx) : 0;
  }
#pragma unroll
  for (int r = 0; r < PPT; ++r) {
    // round r: this lane's rank among the wave's lanes of its bucket, its group's size
    // and leader; rounds, waves and lanes in storage order
    const bool live = p0 + r * 256 + threadIdx.x < np_end;
    int wrank = 0, wcount = 0, lead = lane;
    bool done = !live;
    for (;;) {
      const uint64_t pending = __ballot(!done);
      if (pending == 0ull) break;
      const int first = __ffsll((long long)pending) - 1;
      const int fb = __shfl(b[r], first, 64);
      const bool mine = !done && b[r] == fb;
      const uint64_t same = __ballot(mine);
      if (mine) {
        wrank = __popcll(same & below);
        wcount = __popcll(same);
        lead = first;
        done = true;
      }
    }
    int start = 0;
    for (int w = 0; w < 4; ++w) {
      // the group leaders of wave w take their range; buckets are distinct within a wave
      if (wave == w && live && wrank == 0) {
        start = s_h[b[r]];
        s_h[b[r]] = start + wcount;
      }
      __syncthreads();
    }
    local[r] = __shfl(start, lead, 64) + wrank;
  }
  for (int e = threadIdx.x; e < nb; e += 256) {
    const int c = s_h[e];
    if (c) s_h[e] = atomicAdd(&cnt[base + e], c);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PPT; ++r) {
    const int64_t p = p0 + r * 256 + threadIdx.x;
    if (p < np_end) rank[p] = s_h[b[r]] + local[r];
  }
}

// Exclusive scan of the bucket counts into offsets (one workgroup, 4096 buckets per
// pass); the counts are cleared for the next step as they are read.
__global__ void __launch_bounds__(1024) k_bucket_scan(int* __restrict__ cnt,
                                                      int* __restrict__ off, int nb) {
  __shared__ int s_w[16];
  __shared__ int s_carry;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 4096) {
    const int i0 = base + 4 * (int)threadIdx.x;
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = i0 + k < nb ? cnt[i0 + k] : 0;
    const int sum = v[0] + v[1] + v[2] + v[3];
    int x = sum;  // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    int e = s_carry + x - sum;
    for (int w = 0; w < wave; ++w) e += s_w[w];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (i0 + k < nb) {
        off[i0 + k] = e;
        cnt[i0 + k] = 0;
      }
      e += v[k];
    }
    __syncthreads();  // s_carry and s_w read by every thread
    if (threadIdx.x == 1023) s_carry = e;
    __syncthreads();
  }
}

// sorted[offset of the bucket + slot] = particle (the scattered side is the store).
// BATCH: the offsets run over all scenes' buckets in scene order, and the scenes'
// particle ranges are contiguous in scene order, so every scene's particles land in
// its own range.
template <bool BATCH>
__global__ void __launch_bounds__(256) k_bucket_fill(Geom g0, const float4* __restrict__ pt,
                                                     int np, int bx0, const int* __restrict__ off,
                                                     const int* __restrict__ rank,
                                                     float4* __restrict__ sorted, BatchTables tb) {
  const PartTile T = part_tile<BATCH>(g0, np, tb);
  int bx = bx0, base = 0;
  if (BATCH) {
    const BatchScene sc = tb.scene[tb.ptile[xcd_tile(blockIdx.x, gridDim.x)].x];
    bx = sc.bucket_bx;
    base = sc.bucket_base;
  }
  const int64_t p = T.p;
  if (p >= T.np_end) return;
  const float4 q = pt[p];
  const int d = off[base + particle_bucket(T.g, q, bx)] + rank[p];
  if (d >= 0 && d < np) sorted[d] = q;
}

// Plain fill used instead of hipMemsetAsync (stays valid inside captured graphs).
__global__ void __launch_bounds__(256) k_fill_f32(float* __restrict__ a, int64_t n, float val) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = val;
}

// ---------------------------------------------------------------- launchers
static inline int nblocks(int64_t n) { return (int)((n + 255) / 256); }

void launch_cell_init(hipStream_t s, const Geom& g, uint8_t* label, const uint8_t* solid,
                      float4* acc, int do_labels, int do_acc) {
  if (!do_labels && !do_acc) return;
  const CellInit ci{do_labels ? label : nullptr, solid, do_acc ? acc : nullptr};
  k_cell_init<<<nblocks((int64_t)g.nx * g.ny), 256, 0, s>>>(g, ci);
}
void launch_particle_scatter(hipStream_t s, const Geom& g, const float4* pt, int64_t np,
                             uint8_t* label, float4* acc, int do_classify, int do_p2g,
                             const BatchTables* batch, long long* acc_fx) {
  if (batch) {
    if (batch->n_ptile == 0) return;
    if (acc_fx)
      k_particle_scatter<true, true><<<batch->n_ptile, 256, 0, s>>>(g, pt, np, label, acc_fx,
                                                                    do_classify, do_p2g, *batch);
    else
      k_particle_scatter<true, false><<<batch->n_ptile, 256, 0, s>>>(g, pt, np, label, acc,
                                                                     do_classify, do_p2g, *batch);
    return;
  }
  if (np == 0) return;
  if (acc_fx)
    k_particle_scatter<false, true><<<nblocks(np), 256, 0, s>>>(g, pt, np, label, acc_fx,
                                                                do_classify, do_p2g, BatchTables{});
  else
    k_particle_scatter<false, false><<<nblocks(np), 256, 0, s>>>(g, pt, np, label, acc,
                                                                 do_classify, do_p2g, BatchTables{});
}

// Fixed-point sums -> float: out[k] = (float)(fx[k] * 2^-40) (one rounding from the
// double, exact below 2^53), and fx[k] := 0 for the next accumulation.  n values.
__global__ void __launch_bounds__(256) k_fx_to_float(int64_t n, long long* __restrict__ fx,
                                                     float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  out[k] = fx_value(fx[k]);
  fx[k] = 0;
}
void launch_fx_to_float(hipStream_t s, int64_t n, long long* fx, float* out) {
  if (n > 0) k_fx_to_float<<<nblocks(n), 256, 0, s>>>(n, fx, out);
}

int bucket_count(const Geom& g) {
  return ((g.nx + (1 << kBucketShift) - 1) >> kBucketShift) *
         ((g.ny + (1 << kBucketShift) - 1) >> kBucketShift);
}

void launch_bucket_order(hipStream_t s, const Geom& g, const float4* pt, int64_t np, int* cnt,
                         int* off, int* rank, float4* sorted, const CellInit& ci) {
  if (np <= 0) return;
  const int bx = (g.nx + (1 << kBucketShift) - 1) >> kBucketShift;
  const int nb = bucket_count(g);
  const int n = (int)np;
  constexpr int kPPT = 4;
  k_bucket_count<false, kPPT><<<nblocks((np + kPPT - 1) / kPPT), 256, (size_t)nb * sizeof(int),
                                s>>>(g, pt, n, bx, nb, cnt, rank, BatchTables{}, ci);
  k_bucket_scan<<<1, 1024, 0, s>>>(cnt, off, nb);
  k_bucket_fill<false><<<nblocks(np), 256, 0, s>>>(g, pt, n, bx, off, rank, sorted, BatchTables{});
}

void launch_bucket_order_batch(hipStream_t s, const float4* pt, int64_t np, int* cnt, int* off,
                               int* rank, float4* sorted, const BatchTables& tb, int n_buckets,
                               int max_scene_buckets) {
  if (np <= 0 || tb.n_ptile == 0) return;
  const int n = (int)np;
  k_bucket_count<true, 1><<<tb.n_ptile, 256, (size_t)max_scene_buckets * sizeof(int), s>>>(
      Geom{}, pt, n, 0, 0, cnt, rank, tb, CellInit{});
  k_bucket_scan<<<1, 1024, 0, s>>>(cnt, off, n_buckets);
  k_bucket_fill<true><<<tb.n_ptile, 256, 0, s>>>(Geom{}, pt, n, 0, off, rank, sorted, tb);
}
void launch_face_update(hipStream_t s, const Geom& g, int flags, float* u, float* v,
                        const float4* acc, float* prev_u, float* prev_v, const uint8_t* label,
                        float gx_dt, float gy_dt, float* ext_u, float* ext_v, uint8_t* mxf,
                        uint8_t* mxb, uint8_t* myf, uint8_t* myb, uint8_t* info_u,
                        uint8_t* info_v, uint8_t* dec) {
  k_face_update<<<nblocks((int64_t)g.nx * g.ny), 256, 0, s>>>(
      g, flags, u, v, acc, prev_u, prev_v, label, gx_dt, gy_dt, ext_u, ext_v, mxf, mxb, myf,
      myb, info_u, info_v, dec);
}
void launch_extend_sweep(hipStream_t s, const Geom& g, const uint8_t* label, float* ub,
                         float* vb, const uint8_t* mxf, uint8_t* mxb, const uint8_t* myf,
                         uint8_t* myb, uint8_t* info_u, uint8_t* info_v, int layer) {
  k_extend_sweep<<<nblocks((int64_t)g.nx * g.ny), 256, 0, s>>>(g, label, ub, vb, mxf, mxb, myf,
                                                                myb, info_u, info_v, layer);
}
void launch_classify_levelset(hipStream_t s, const Geom& g, const float* phi,
                              const uint8_t* solid, uint8_t* label) {
  k_classify_levelset<<<nblocks((int64_t)g.nx * g.ny), 256, 0, s>>>(g, phi, solid, label);
}
void launch_pressure_update(hipStream_t s, const Geom& g, const int* n_fluid,
                            const uint8_t* label, float* xg, double* x64,
                            const int* counts, float ppk, float* au, float* av,
                            float* bu, float* bv, float dt_over_density, int do_dirichlet,
                            int do_diff, const float* prev_u, const float* prev_v, float* diff_u,
                            float* diff_v, uint8_t* zmask, uint8_t* dec,
                            const MgFinishFuse& fin) {
  k_pressure_update<<<nblocks((int64_t)g.nx * g.ny), 256, 0, s>>>(
      g, n_fluid, label, xg, x64, counts, ppk, au, av, bu, bv, dt_over_density, do_dirichlet,
      do_diff, prev_u, prev_v, diff_u, diff_v, zmask, dec, fin);
}
void launch_count_particles(hipStream_t s, const Geom& g, const float4* pt, int64_t np,
                            int* counts) {
  launch_fill(s, reinterpret_cast<float*>(counts), (int64_t)g.nx * g.ny, 0.f);  // int 0 bits
  if (np == 0) return;
  k_count_particles<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(g, pt, np, counts);
}
void launch_emit_lattice(hipStream_t s, float4* pt, int64_t base, const float* xs, int nxp,
                         const float* ys, int nyp, float vx, float vy) {
  const int64_t n = (int64_t)nxp * nyp;
  if (n == 0) return;
  k_emit_lattice<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pt, base, xs, nxp, ys, nyp, vx, vy);
}
void launch_advect_level_set(hipStream_t s, const Geom& g, const float* phi, float* out,
                             const float* u, const float* v, float dt) {
  const int n = g.nx * g.ny;
  k_advect_level_set<<<(n + 255) / 256, 256, 0, s>>>(g, phi, out, u, v, dt);
}
void launch_diff(hipStream_t s, const Geom& g, const float* u, const float* v, const float* pu,
                 const float* pv, float* du, float* dv) {
  k_diff<<<nblocks((int64_t)g.nx * g.ny), 256, 0, s>>>(g, u, v, pu, pv, du, dv);
}
void launch_particle_gather(hipStream_t s, const Geom& g, float4* pt, int64_t np, const float* u,
                            const float* v, const float* du, const float* dv, float pic_ratio,
                            int g2p, int move, float dt, const uint8_t* label, uint8_t* reject,
                            const BatchTables* batch) {
  if (batch) {
    if (batch->n_ptile == 0) return;
    k_particle_gather<true><<<batch->n_ptile, 256, 0, s>>>(g, pt, np, u, v, du, dv, pic_ratio,
                                                           g2p, move, dt, label, reject, *batch);
    return;
  }
  if (np == 0) return;
  k_particle_gather<false><<<nblocks(np), 256, 0, s>>>(g, pt, np, u, v, du, dv, pic_ratio, g2p,
                                                       move, dt, label, reject, BatchTables{});
}
void launch_sl_scatter(hipStream_t s, const Geom& g, const uint8_t* label, const float* uf,
                       const float* vf, float* ub, float* vb, int integrator, float dt,
                       long long* fx) {
  const int64_t C = (int64_t)g.nx * g.ny;
  if (fx) {
    k_sl_scatter<true><<<nblocks(C), 256, 0, s>>>(
        g, label, uf, vf, ub, vb, reinterpret_cast<unsigned long long*>(fx), integrator, dt);
    launch_fx_to_float(s, C, fx, ub);
    launch_fx_to_float(s, C, fx + C, vb);
    return;
  }
  k_sl_scatter<false><<<nblocks(C), 256, 0, s>>>(g, label, uf, vf, ub, vb, nullptr, integrator,
                                                 dt);
}
void launch_sl_gather(hipStream_t s, const Geom& g, const uint8_t* label, const float* uf,
                      const float* vf, float* ub, float* vb, int integrator, float dt) {
  k_sl_gather<<<nblocks((int64_t)g.nx * g.ny), 256, 0, s>>>(g, label, uf, vf, ub, vb,
                                                             integrator, dt);
}
void launch_advect_grid(hipStream_t s, const Geom& g, float4* pt, int64_t np, const float* uf,
                        const float* vf, int integrator, float dt) {
  if (np == 0) return;
  k_advect_grid<<<nblocks(np), 256, 0, s>>>(g, pt, np, uf, vf, integrator, dt);
}
void launch_fill(hipStream_t s, float* a, int64_t n, float val) {
  if (n == 0) return;
  k_fill_f32<<<nblocks(n), 256, 0, s>>>(a, n, val);
}

}  // namespace fsg
