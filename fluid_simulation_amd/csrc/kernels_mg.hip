// kernels_mg.hip -- multigrid-preconditioned CG with FP64 solution and residual
// (FSG_SOLVER_MG): the fast form of the pressure solve that replaces
// Eigen::ConjugateGradient (include/FluidSolver.h:114, src/FluidSolver.cpp:81,
// 362-425).  It solves the same system as the reference -- A x = b over the
// LIQUID cells, A_ii = -n/dx^2 (n non-SOLID neighbours), A_ij = 1/dx^2 for LIQUID
// neighbours, b the cell divergence -- to a relative FP64 residual
// |b - A x| <= tol |b|, so its result is the converged pressure, not Eigen's
// iterate after a fixed iteration count.
//
// Layout: every vector is a dense cell grid (row-major i + j*nx) holding zero at
// cells that are not unknowns, so the 5-point operator needs no neighbour lists.
// The system is scaled to the integer stencil K = -dx^2 A (K_ii = n, K_ij = -1),
// right-hand side b' = -dx^2 b, kept in FP64.
//
// Precision: the solution x and the residual r = b' - K x are FP64 (the residual
// is recomputed from x at every iteration, never updated recursively); the search
// direction, the preconditioned residual and the whole V-cycle are FP32; every dot
// product is summed in FP64.
//
// Preconditioner: one V-cycle of cell-centred geometric multigrid
// (McAdams, Sifakis, Teran, "A parallel multigrid Poisson solver for fluids
// simulation on large grids", SCA 2010): a coarse cell is AIR if any child is AIR,
// else LIQUID if any child is LIQUID, else SOLID; each level is the same integer
// stencil on its own labels; prolongation is cell-centred bilinear (weights 9, 3,
// 3, 1 / 16), except that the weight of a coarse neighbour that is SOLID or
// outside the grid goes to the fine cell's own coarse cell (constant extrapolation
// across a wall: the Neumann condition; interpolating a zero there acts like a
// Dirichlet wall and costs 2-4x the iterations on closed domains), and restriction
// is its exact transpose; smoothing is two damped-Jacobi sweeps
// before and after the coarse correction.  Every component is a symmetric linear
// operator, so the V-cycle is an SPD preconditioner for CG.  In the scaled form a
// level-(l+1) right-hand side is P^T of the level-l residual (the 4^-l scale of the
// rediscretised operators cancels), so every level uses the unit stencil.
//
// Kernels per CG iteration: mg_down (one per fine level: 2 pre-smoothing sweeps,
// residual and restriction fused in one LDS tile), mg_coarse (one workgroup: the
// V-cycle of the levels of <= 1024 cells in LDS), mg_up (one per fine level:
// prolongation, correction and 2 post-smoothing sweeps fused), mg_dir (p = z +
// beta p, K p and p.Kp), mg_xr (x += alpha p in FP64 and r = b' - K x).  Every
// kernel returns at once when the solve is done, and each sums the block partials
// of the previous kernel itself in a fixed order, so the scalars are identical in
// every block and the solve is deterministic.
#include <float.h>

#include <algorithm>
#include <cmath>

#include <type_traits>

#include "fsg_common.h"
#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

namespace {

constexpr int kT = 256;         // threads of the vector kernels and the level-0 tiles
constexpr int kTileBig = 1024;  // threads of the tiles of the smaller levels
constexpr uint8_t kAct = kMgUnknown;  // info bit: the cell is an unknown
constexpr int kTileEmptyBit = (int)0x80000000u;  // k_mg_tile_map: the tile holds no unknown
constexpr uint8_t kWall = 0x40;   // info bit: SOLID (also used for cells outside the grid)

__device__ __forceinline__ bool is_act(uint8_t f) { return (f & kAct) != 0; }
__device__ __forceinline__ bool is_wall(uint8_t f) { return (f & kWall) != 0; }
// 1 / n for the Jacobi smoother (a fixed diagonal scaling: any consistent rounding
// keeps the smoother symmetric)
__device__ __forceinline__ float inv_nns(uint8_t f) {
  const int n = f & 7;
  float r = 1.0f;
  r = n == 2 ? 0.5f : r;
  r = n == 3 ? 0.333333343f : r;
  r = n == 4 ? 0.25f : r;
  return r;
}

// Bilinear prolongation of one fine cell from its coarse cell (value e0) and the
// x-, y- and diagonal neighbours (e, info f): a wall neighbour's weight is moved to
// the fine cell's own coarse cell.
__device__ __forceinline__ float prolong_value(float e0, float ex, uint8_t fx, float ey, uint8_t fy,
                                               float exy, uint8_t fxy) {
  float w0 = 0.5625f, v = 0.f;
  if (is_wall(fx)) w0 += 0.1875f; else v += 0.1875f * ex;
  if (is_wall(fy)) w0 += 0.1875f; else v += 0.1875f * ey;
  if (is_wall(fxy)) w0 += 0.0625f; else v += 0.0625f * exy;
  return v + w0 * e0;
}

// Extra restriction weight of a fine cell whose own coarse cell is the target:
// the weights its prolongation moved from wall neighbours (fx, fy, fxy).
__device__ __forceinline__ float moved_weight(uint8_t fx, uint8_t fy, uint8_t fxy) {
  return (is_wall(fx) ? 0.1875f : 0.f) + (is_wall(fy) ? 0.1875f : 0.f) +
         (is_wall(fxy) ? 0.0625f : 0.f);
}
__device__ __forceinline__ float nns_of(uint8_t f) { return (float)(f & 7); }

// Unknown flag and non-SOLID neighbour count of cell (i, j) of a label grid
// (neighbours outside the grid count as SOLID).  A LIQUID cell whose four
// neighbours are SOLID has an all-zero row and is left out of the system.
__device__ __forceinline__ uint8_t make_info(const uint8_t* __restrict__ lab, int nx, int ny,
                                             int i, int j) {
  const int c = i + j * nx;
  int n = 0;
  n += i > 0 && lab[c - 1] != kSolid;
  n += i < nx - 1 && lab[c + 1] != kSolid;
  n += j > 0 && lab[c - nx] != kSolid;
  n += j < ny - 1 && lab[c + nx] != kSolid;
  const bool act = lab[c] == kLiquid && n > 0;
  return (uint8_t)((act ? kAct : 0) | (lab[c] == kSolid ? kWall : 0) | n);
}

// (i, j) of a cell index that advances by a fixed stride: one division at the
// start, none per step (grid-stride loops over row-major cells).
struct CellWalk {
  int i, j, di, dj, nx;
  __device__ __forceinline__ CellWalk(int c, int stride, int nx_) : nx(nx_) {
    j = c / nx;
    i = c - j * nx;
    dj = stride / nx;
    di = stride - dj * nx;
  }
  __device__ __forceinline__ void next() {
    i += di;
    j += dj;
    if (i >= nx) {
      i -= nx;
      ++j;
    }
  }
};

// i = q % nx, j = q / nx for q < 2^12 via a float reciprocal: (q + 0.5) / nx is
// at least 0.5 / nx away from an integer and the product's rounding error is below
// 2^-12 / nx, so the truncation is exact.
__device__ __forceinline__ void small_divmod(int q, int nx, float inv_nx, int& i, int& j) {
  j = (int)(((float)q + 0.5f) * inv_nx);
  i = q - j * nx;
}

// Sum of n block partials, the same fixed order in every block (NT threads;
// sm: NT / 64 doubles).
template <int NT = kT>
__device__ __forceinline__ double sum_part(const double* __restrict__ part, int n, double* sm) {
  double v[1] = {0.0};
  int b = threadIdx.x;
  // four loads in flight per thread; the additions keep the order b, b + NT, ...
  for (; b + 3 * NT < n; b += 4 * NT) {
    const double a0 = part[b], a1 = part[b + NT], a2 = part[b + 2 * NT], a3 = part[b + 3 * NT];
    v[0] += a0;
    v[0] += a1;
    v[0] += a2;
    v[0] += a3;
  }
  for (; b < n; b += NT) v[0] += part[b];
  block_sum<1, NT>(v, sm);
  return v[0];
}

// Cells per thread per pass of the CG vector kernels: the loads of all of them are
// issued before the first is used, so a thread waits on memory once per pass.
constexpr int kVecU = 4;

// The 5-point stencil of cell c = (i, j) on an nx x ny grid: own, left, right, down, up.
// A neighbour outside the grid (or any, for a cell past the end, in = false) points at
// c itself and has its mask bit clear, so every index is in bounds.
struct Stencil5 {
  int q[5];
  bool m[5];
};

__device__ __forceinline__ Stencil5 stencil5(int c, int i, int j, int nx, int ny, bool in) {
  Stencil5 s;
  s.m[0] = in;
  s.m[1] = in && i > 0;
  s.m[2] = in && i < nx - 1;
  s.m[3] = in && j > 0;
  s.m[4] = in && j < ny - 1;
  s.q[0] = c;
  s.q[1] = s.m[1] ? c - 1 : c;
  s.q[2] = s.m[2] ? c + 1 : c;
  s.q[3] = s.m[3] ? c - nx : c;
  s.q[4] = s.m[4] ? c + nx : c;
  return s;
}

template <int NT = kT>
__device__ __forceinline__ void write_part(double val, double* __restrict__ part, double* sm) {
  double v[1] = {val};
  block_sum<1, NT>(v, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

}  // namespace

// ---------------------------------------------------------------- setup
// Level-0 info, b' = -dx^2 * div (the divergence exactly as the compact path forms
// it), and per-block partials (sum b', number of unknowns).
// rhs (optional): the reference's b (src/FluidSolver.cpp:413, unscaled FP32) at every
// LIQUID cell, 0 elsewhere, and part_liq the block's count of LIQUID cells (the
// reference's number of fluid indices, :331-346): what the compaction of the other
// solver forms would give, so the multigrid form needs no compaction.
__global__ void __launch_bounds__(kT) k_mg_setup0(Geom g, const uint8_t* __restrict__ label,
                                                   const float* __restrict__ uf,
                                                   const float* __restrict__ vf,
                                                   uint8_t* __restrict__ inf,
                                                   double* __restrict__ b,
                                                   double* __restrict__ part_sum,
                                                   double* __restrict__ part_cnt,
                                                   float* __restrict__ rhs,
                                                   double* __restrict__ part_liq) {
  __shared__ double sm[4];
  const int nx = g.nx, ny = g.ny, C = nx * ny;
  const double scale = -(double)g.dx * (double)g.dx;
  double s = 0.0, cnt = 0.0, liq = 0.0;
  CellWalk w(blockIdx.x * kT + threadIdx.x, gridDim.x * kT, nx);
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT, w.next()) {
    const int i = w.i, j = w.j;
    const uint8_t f = make_info(label, nx, ny, i, j);
    inf[c] = f;
    const bool fluid = label[c] == kLiquid;
    double bv = 0.0;
    float bf = 0.f;
    if (fluid) {
      const float uxp = i < nx - 1 ? uf[c + 1] : 0.f;
      const float vyp = j < ny - 1 ? vf[c + nx] : 0.f;
      const float dvx = div_x(uxp - uf[c], g);
      const float dvy = div_y(vyp - vf[c], g);
      bf = dvx + dvy;
      liq += 1.0;
    }
    if (is_act(f)) {
      bv = scale * (double)bf;
      s += bv;
      cnt += 1.0;
    }
    b[c] = bv;
    if (rhs) rhs[c] = bf;
  }
  write_part<kT>(s, part_sum, sm);
  write_part<kT>(cnt, part_cnt, sm);
  if (part_liq) write_part<kT>(liq, part_liq, sm);
}

// AIR neighbours of cell (i, j) (Dirichlet links of its row: an unknown with one
// makes the system non-singular).
__device__ __forceinline__ int air_links(const uint8_t* __restrict__ lab, int nx, int ny, int i,
                                         int j) {
  const int c = i + j * nx;
  int n = 0;
  n += i > 0 && lab[c - 1] == kAir;
  n += i < nx - 1 && lab[c + 1] == kAir;
  n += j > 0 && lab[c - nx] == kAir;
  n += j < ny - 1 && lab[c + nx] == kAir;
  return n;
}

// Adjoint solves (the step's reverse pass): the right-hand side is a given cell
// field y (the adjoint of the pressure) instead of the divergence.  With the mean
// removed the forward solution is x = S A^+ P b (P: remove the mean of b, S: the
// D-weighted constant of k_mg_finish, applied only when the system is singular, i.e.
// no unknown has an AIR neighbour), so the adjoint is P A^+ S^T y with
// S^T y = y - n (sum y) / (sum n) over the unknowns (n: non-SOLID neighbours), or
// P A^-1 y for a non-singular system.
// Pass 1: info and the partials of sum y, sum n and the AIR links (part_y / part_n /
// part_d).
__global__ void __launch_bounds__(kT) k_mg_adj_rhs0(Geom g, const uint8_t* __restrict__ label,
                                                    const double* __restrict__ y,
                                                    uint8_t* __restrict__ inf,
                                                    double* __restrict__ part_y,
                                                    double* __restrict__ part_n,
                                                    double* __restrict__ part_d) {
  __shared__ double sm[4];
  const int nx = g.nx, ny = g.ny, C = nx * ny;
  double sy = 0.0, sn = 0.0, sd = 0.0;
  CellWalk w(blockIdx.x * kT + threadIdx.x, gridDim.x * kT, nx);
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT, w.next()) {
    const uint8_t f = make_info(label, nx, ny, w.i, w.j);
    inf[c] = f;
    if (is_act(f)) {
      sy += y[c];
      sn += (double)(f & 7);
      sd += (double)air_links(label, nx, ny, w.i, w.j);
    }
  }
  write_part<kT>(sy, part_y, sm);
  write_part<kT>(sn, part_n, sm);
  write_part<kT>(sd, part_d, sm);
}

// Pass 2: b' = -dx^2 S^T y at the unknowns (S^T = I without the mean removal), and
// the partials of sum b' and of the unknown count that k_mg_setup1 reads.
__global__ void __launch_bounds__(kT) k_mg_adj_rhs1(int C, double scale, int project_mean,
                                                    const uint8_t* __restrict__ inf,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ part_y,
                                                    const double* __restrict__ part_n,
                                                    const double* __restrict__ part_d, int n_part,
                                                    double* __restrict__ b,
                                                    double* __restrict__ part_sum,
                                                    double* __restrict__ part_cnt) {
  __shared__ double sm[4];
  double shift = 0.0;
  if (project_mean && sum_part<kT>(part_d, n_part, sm) == 0.0) {
    const double a = sum_part<kT>(part_y, n_part, sm);
    const double n = sum_part<kT>(part_n, n_part, sm);
    shift = n > 0.0 ? a / n : 0.0;
  }
  double s = 0.0, cnt = 0.0;
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT) {
    const uint8_t f = inf[c];
    double bv = 0.0;
    if (is_act(f)) {
      bv = scale * (y[c] - (double)(f & 7) * shift);
      s += bv;
      cnt += 1.0;
    }
    b[c] = bv;
  }
  write_part<kT>(s, part_sum, sm);
  write_part<kT>(cnt, part_cnt, sm);
}

// P: the plain mean of x over the unknowns removed (after k_mg_finish).
__global__ void __launch_bounds__(kT) k_mg_mean_part(int C, const uint8_t* __restrict__ inf,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ part_sum,
                                                     double* __restrict__ part_cnt) {
  __shared__ double sm[4];
  double s = 0.0, cnt = 0.0;
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT)
    if (is_act(inf[c])) {
      s += x[c];
      cnt += 1.0;
    }
  write_part<kT>(s, part_sum, sm);
  write_part<kT>(cnt, part_cnt, sm);
}

__global__ void __launch_bounds__(kT) k_mg_mean_sub(int C, const uint8_t* __restrict__ inf,
                                                    double* __restrict__ x,
                                                    const double* __restrict__ part_sum,
                                                    const double* __restrict__ part_cnt,
                                                    int n_part) {
  __shared__ double sm[4];
  const double a = sum_part<kT>(part_sum, n_part, sm);
  const double n = sum_part<kT>(part_cnt, n_part, sm);
  const double mean = n > 0.0 ? a / n : 0.0;
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT)
    if (is_act(inf[c])) x[c] -= mean;
}

// Optional mean removal (singular all-Neumann systems), x0 and r0 = b' - K x0,
// |b'|^2 and |r0|^2.  Cold start: x0 = 0 (xw == null).  Warm start: x0 = xw, the
// previous solution already zeroed off this solve's unknowns (k_mg_warm), read with
// its neighbours from xw and copied to x0.
__global__ void __launch_bounds__(kT) k_mg_setup1(int nx, int ny, int project_mean,
                                                   const uint8_t* __restrict__ inf,
                                                   double* __restrict__ b,
                                                   const double* __restrict__ xw,
                                                   double* __restrict__ x0,
                                                   float* __restrict__ r,
                                                   const double* __restrict__ part_sum,
                                                   const double* __restrict__ part_cnt, int n_part,
                                                   double* __restrict__ part_bb,
                                                   double* __restrict__ part_rr, CgScalars* sc,
                                                   const double* __restrict__ part_liq,
                                                   int* __restrict__ n_fluid,
                                                   const uint8_t* __restrict__ tflag,
                                                   int n_tiles) {
  __shared__ double sm[4];
  const int C = nx * ny;
  // the live level-0 tiles, where k_mg_setup0t set their flags (else k_mg_tile_flags
  // counts them after this kernel)
  if (tflag && blockIdx.x == 0) {
    double n = 0.0;
    for (int t = threadIdx.x; t < n_tiles; t += kT) n += tflag[t] ? 1.0 : 0.0;
    double v[1] = {n};
    block_sum<1, kT>(v, sm);
    if (threadIdx.x == 0) sc->live_tiles = (int)v[0];
  }
  // the number of LIQUID cells (k_mg_setup0's partials), when the compaction did not run
  if (part_liq && blockIdx.x == 0) {
    const double n = sum_part<kT>(part_liq, n_part, sm);
    if (threadIdx.x == 0) *n_fluid = (int)n;
  }
  double mean = 0.0;
  if (project_mean) {
    const double s = sum_part<kT>(part_sum, n_part, sm);
    const double n = sum_part<kT>(part_cnt, n_part, sm);
    mean = n > 0.0 ? s / n : 0.0;
  }
  double acc_b = 0.0, acc_r = 0.0;
  CellWalk w(blockIdx.x * kT + threadIdx.x, gridDim.x * kT, nx);
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT, w.next()) {
    const uint8_t f = inf[c];
    double bv = b[c];
    if (project_mean && is_act(f)) {
      bv -= mean;
      b[c] = bv;
    }
    double rv = bv, xv = 0.0;
    if (xw) {
      xv = xw[c];
      if (is_act(f)) {
        const int i = w.i, j = w.j;
        double sn = 0.0;
        if (i > 0) sn += xw[c - 1];
        if (i < nx - 1) sn += xw[c + 1];
        if (j > 0) sn += xw[c - nx];
        if (j < ny - 1) sn += xw[c + nx];
        rv = bv - ((double)(f & 7) * xv - sn);
      }
    }
    x0[c] = xv;
    r[c] = (float)rv;
    acc_b += bv * bv;
    acc_r += rv * rv;
  }
  write_part<kT>(acc_b, part_bb, sm);
  write_part<kT>(acc_r, part_rr, sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->done = 0;
    sc->iter = 0;
    sc->spmv = 0;
    sc->err = 0.0;
    sc->rho[0] = sc->rho[1] = 0.0;
    if (!tflag) sc->live_tiles = 0;
  }
}

// Packed batch: |b'_s| over each scene's rectangle, one workgroup per scene (a
// fixed summation order), and its reciprocal.
__global__ void __launch_bounds__(1024) k_mg_scene_norm(const double* __restrict__ b, int pitch,
                                                        const int4* __restrict__ rect,
                                                        double* __restrict__ norm,
                                                        double* __restrict__ inv) {
  __shared__ double sm[16];
  const int4 r = rect[blockIdx.x];
  double v[1] = {0.0};
  for (int q = threadIdx.x; q < r.z * r.w; q += 1024) {
    const int i = q % r.z, j = q / r.z;
    const double x = b[(r.x + i) + (int64_t)(r.y + j) * pitch];
    v[0] += x * x;
  }
  block_sum<1, 1024>(v, sm);
  if (threadIdx.x == 0) {
    const double n = sqrt(v[0]);
    norm[blockIdx.x] = n;
    inv[blockIdx.x] = n > 0.0 ? 1.0 / n : 0.0;
  }
}

// b'[c] *= factor[scene of c]
__global__ void __launch_bounds__(kT) k_mg_scene_scale(int C, const int* __restrict__ sid,
                                                        const double* __restrict__ factor,
                                                        double* __restrict__ b) {
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT) {
    const int s = sid[c];
    if (s >= 0) b[c] *= factor[s];
  }
}

// Warm start: the previous solution at this solve's unknowns, 0 elsewhere.
__global__ void __launch_bounds__(kT) k_mg_warm(int C, const uint8_t* __restrict__ inf,
                                                 const double* __restrict__ prev,
                                                 double* __restrict__ xw) {
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT)
    xw[c] = is_act(inf[c]) ? prev[c] : 0.0;
}

// Coarse labels of level l+1 from level l, and (finf != null) the info of level l.
__global__ void __launch_bounds__(kT) k_mg_coarsen(int fnx, int fny, const uint8_t* __restrict__ flab,
                                                    uint8_t* __restrict__ finf, int cnx, int cny,
                                                    uint8_t* __restrict__ clab) {
  const int nf = fnx * fny, nc = cnx * cny;
  const int n = nf > nc ? nf : nc;
  for (int t = blockIdx.x * kT + threadIdx.x; t < n; t += gridDim.x * kT) {
    if (t < nc) {
      const int I = t % cnx, J = t / cnx;
      bool air = false, liq = false;
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const int i = 2 * I + a, j = 2 * J + b;
          if (i < fnx && j < fny) {
            const uint8_t L = flab[i + j * fnx];
            air |= L == kAir;
            liq |= L == kLiquid;
          }
        }
      clab[t] = air ? kAir : (liq ? kLiquid : kSolid);
    }
    if (finf && t < nf) finf[t] = make_info(flab, fnx, fny, t % fnx, t / fnx);
  }
}

// The last levels of the coarsening (level l0-1 of at most kTailCells cells and
// below) in one workgroup: level l0-1's labels staged in LDS, then per level the info
// bytes of the finer level and the labels of the coarser one formed there and written
// out, and finally the coarsest level's info -- what k_mg_coarsen does one launch per
// level.
constexpr int kTailCells = 4096;
struct MgTailArgs {
  int l0, nlev;
  int nx[kMgMaxLevels], ny[kMgMaxLevels];
  uint8_t* lab[kMgMaxLevels];
  uint8_t* inf[kMgMaxLevels];
};

__device__ __forceinline__ void coarsen_tail(const MgTailArgs& a) {
  __shared__ uint8_t s_lab[2][kTailCells];
  const int lf = a.l0 - 1;
  {
    const int n = a.nx[lf] * a.ny[lf];
    for (int t = threadIdx.x; t < n; t += 1024) s_lab[0][t] = a.lab[lf][t];
  }
  __syncthreads();
  int cur = 0;
  for (int l = a.l0; l <= a.nlev; ++l) {
    const int fnx = a.nx[l - 1], fny = a.ny[l - 1], nf = fnx * fny;
    const uint8_t* flab = s_lab[cur];
    for (int t = threadIdx.x; t < nf; t += 1024)
      a.inf[l - 1][t] = make_info(flab, fnx, fny, t % fnx, t / fnx);
    if (l == a.nlev) break;
    const int cnx = a.nx[l], cny = a.ny[l], nc = cnx * cny;
    for (int t = threadIdx.x; t < nc; t += 1024) {
      const int I = t % cnx, J = t / cnx;
      bool air = false, liq = false;
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int i = 2 * I + e, j = 2 * J + b;
          if (i < fnx && j < fny) {
            const uint8_t L = flab[i + j * fnx];
            air |= L == kAir;
            liq |= L == kLiquid;
          }
        }
      const uint8_t L = air ? kAir : (liq ? kLiquid : kSolid);
      s_lab[cur ^ 1][t] = L;
      a.lab[l][t] = L;
    }
    __syncthreads();
    cur ^= 1;
  }
}

__global__ void __launch_bounds__(1024) k_mg_coarsen_tail(MgTailArgs a) { coarsen_tail(a); }

// One domain, the divergence right-hand side: level 0's set-up by 32 x 32 tiles
// (k_mg_setup0's outputs, with one partial per tile), the tile's flag (a LIQUID cell in
// it: every unknown of level 0 or 1 lies under one, so the level-0 tile kernels skip the
// others, as k_mg_tile_flags' exact flags let them), and the labels of the levels
// 1 .. nl whose cells lie inside one tile (2^nl <= 32), coarsened in LDS by the rule of
// k_mg_coarsen (children outside the grid take no part).
struct MgTileLabels {
  int nl;
  int nx[6], ny[6];
  uint8_t* lab[6];
};
// Set: k_mg_setup0t also does k_mg_setup1's work for a cold start without the mean
// projection (x0 = 0, the FP32 residual r0 = b', one |b'|^2 partial per tile in both
// part_bb and part_rr, zero partials from n_tiles to nb, the solve's scalars reset).
struct MgSetupFold {
  double* x0 = nullptr;
  float* r0 = nullptr;
  double* part_bb = nullptr;
  double* part_rr = nullptr;
  int nb = 0;
  CgScalars* sc = nullptr;
};
__global__ void __launch_bounds__(kT) k_mg_setup0t(Geom g, const uint8_t* __restrict__ label,
                                                    const float* __restrict__ uf,
                                                    const float* __restrict__ vf,
                                                    uint8_t* __restrict__ inf,
                                                    double* __restrict__ b,
                                                    double* __restrict__ part_sum,
                                                    double* __restrict__ part_cnt,
                                                    float* __restrict__ rhs,
                                                    double* __restrict__ part_liq,
                                                    uint8_t* __restrict__ tflag, MgTileLabels L,
                                                    MgSetupFold fold) {
  constexpr uint8_t kNone = 0xFF;  // a cell outside the grid
  __shared__ double sm[4 * (kT / 64)];
  __shared__ uint8_t s_l[2][32 * 32];
  const int nx = g.nx, ny = g.ny;
  const int tiles_x = (nx + 31) >> 5;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i0 = (tile % tiles_x) * 32, j0 = (tile / tiles_x) * 32;
  const double scale = -(double)g.dx * (double)g.dx;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // sum b', unknowns, LIQUID cells, |b'|^2
  int any = 0;
#pragma unroll
  for (int k = 0; k < 32 * 32 / kT; ++k) {
    const int t = threadIdx.x + k * kT;
    const int i = i0 + (t & 31), j = j0 + (t >> 5);
    uint8_t lab = kNone;
    if (i < nx && j < ny) {
      const int c = i + j * nx;
      const uint8_t f = make_info(label, nx, ny, i, j);
      inf[c] = f;
      lab = label[c];
      const bool fluid = lab == kLiquid;
      double bv = 0.0;
      float bf = 0.f;
      if (fluid) {
        const float uxp = i < nx - 1 ? uf[c + 1] : 0.f;
        const float vyp = j < ny - 1 ? vf[c + nx] : 0.f;
        const float dvx = div_x(uxp - uf[c], g);
        const float dvy = div_y(vyp - vf[c], g);
        bf = dvx + dvy;
        acc[2] += 1.0;
        any = 1;
      }
      if (is_act(f)) {
        bv = scale * (double)bf;
        acc[0] += bv;
        acc[1] += 1.0;
      }
      b[c] = bv;
      rhs[c] = bf;
      if (fold.x0) {
        fold.x0[c] = 0.0;
        fold.r0[c] = (float)bv;
        acc[3] += bv * bv;
      }
    }
    s_l[0][t] = lab;
  }
  // the partials in k_mg_setup0's order of magnitude (one per tile, fixed order)
  block_sum<4, kT>(acc, sm);
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) {
    part_sum[blockIdx.x] = acc[0];
    part_cnt[blockIdx.x] = acc[1];
    part_liq[blockIdx.x] = acc[2];
    tflag[tile] = any ? 1 : 0;
    if (fold.x0) {
      fold.part_bb[blockIdx.x] = acc[3];
      fold.part_rr[blockIdx.x] = acc[3];
    }
  }
  if (fold.x0 && blockIdx.x == 0) {
    for (int e = (int)gridDim.x + threadIdx.x; e < fold.nb; e += kT) {
      fold.part_bb[e] = 0.0;
      fold.part_rr[e] = 0.0;
    }
    if (threadIdx.x == 0) {
      CgScalars* sc = fold.sc;
      sc->done = 0;
      sc->iter = 0;
      sc->spmv = 0;
      sc->err = 0.0;
      sc->rho[0] = sc->rho[1] = 0.0;
    }
  }
  for (int l = 1; l <= L.nl; ++l) {
    const int w = 32 >> l, wf = 2 * w;
    const uint8_t* src = s_l[(l - 1) & 1];
    uint8_t* dst = s_l[l & 1];
    for (int t = threadIdx.x; t < w * w; t += kT) {
      const int I = t % w, J = t / w;
      bool air = false, liq = false;
#pragma unroll
      for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int aa = 0; aa < 2; ++aa) {
          const uint8_t v = src[(2 * I + aa) + (2 * J + bb) * wf];
          air |= v == kAir;
          liq |= v == kLiquid;
        }
      const int Ig = (i0 >> l) + I, Jg = (j0 >> l) + J;
      const bool in = Ig < L.nx[l] && Jg < L.ny[l];
      const uint8_t v = in ? (air ? kAir : (liq ? kLiquid : kSolid)) : kNone;
      dst[t] = v;
      if (in) L.lab[l][Ig + Jg * L.nx[l]] = v;
    }
    __syncthreads();
  }
}

// Block -> tile map of one domain's level-0 tile kernels.  Workgroup b runs on XCD
// b % 8 (workgroups are handed to the XCDs in turn).  xcd_tile gives each XCD one band
// of consecutive tiles, but a dam break's unknowns fill the tiles of a few bands only:
// 5 of the 8 XCDs held all 278 live tiles of the developed 1024^2 flow, 52 each.  Here
// the live tiles (row-major order) are dealt out in 8 contiguous runs of
// lp = ceil(live / 8), run k to the first blocks of XCD k (33 per XCD there), and the
// empty tiles take the remaining blocks in order.  tmap[b] = tile, bit 31 set where the
// tile is empty.  Where a run would not fit an XCD's blocks (almost every tile live) the
// map is xcd_tile's.  One 1024-thread workgroup.
__device__ __forceinline__ int live_blocks_before(int b, int n_live, int lp) {
  int s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int cnt = min(max(n_live - k * lp, 0), lp);
    const int have = b > k ? (b - k + 7) / 8 : 0;
    s += min(cnt, have);
  }
  return s;
}

__device__ __forceinline__ void tile_map_block(int nb, const uint8_t* __restrict__ tflag,
                                               int* __restrict__ tmap) {
  __shared__ int s_w[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int n_live = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int t = base + threadIdx.x;
    n_live += __syncthreads_count(t < nb && tflag[t] != 0);
  }
  const int per = nb / 8, lp = (n_live + 7) / 8;
  if (nb < 8 || lp > per) {
    for (int b = threadIdx.x; b < nb; b += 1024) {
      const int t = xcd_tile(b, nb);
      tmap[b] = t | (tflag[t] ? 0 : kTileEmptyBit);
    }
    return;
  }
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  int carry = 0;  // live tiles before this chunk
  for (int base = 0; base < nb; base += 1024) {
    const int t = base + threadIdx.x;
    const bool live = t < nb && tflag[t] != 0;
    const uint64_t m = __ballot(live);
    if (lane == 0) s_w[wave] = __popcll(m);
    __syncthreads();
    int q = carry + __popcll(m & below);  // live tiles before t
    int chunk = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wave) q += s_w[w];
      chunk += s_w[w];
    }
    __syncthreads();  // s_w is rewritten by the next chunk
    if (t < nb) {
      if (live) {
        tmap[(q % lp) * 8 + q / lp] = t;
      } else {
        // the (t - q)-th block that takes no live tile
        const int e = t - q;
        int lo = 0, hi = nb - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (mid + 1 - live_blocks_before(mid + 1, n_live, lp) > e)
            hi = mid;
          else
            lo = mid + 1;
        }
        tmap[lo] = t | kTileEmptyBit;
      }
    }
    carry += chunk;
  }
}
__global__ void __launch_bounds__(1024) k_mg_tile_map(int nb, const uint8_t* __restrict__ tflag,
                                                      int* __restrict__ tmap) {
  tile_map_block(nb, tflag, tmap);
}

// The info bytes of levels 1 .. nli (their labels complete), by every block but the
// last, which runs the coarsening tail (levels l0-1 and below; l0 - 1 > nli) and, where
// k_mg_setup0t did k_mg_setup1's work, the solve's LIQUID count and live level-0 tiles.
struct MgInfoArgs {
  int nli;
  int nx[kMgMaxLevels], ny[kMgMaxLevels];
  const uint8_t* lab[kMgMaxLevels];
  uint8_t* inf[kMgMaxLevels];
  // set: the next-to-last block forms the level-0 tile map from tflag (nb tiles)
  int nb_map;
  const uint8_t* tflag;
  int* tmap;
  // set (setup folded): *n_fluid = the sum of n_part part_liq partials, and
  // sc->live_tiles = the tiles flagged in tflag0 (n_tiles)
  const double* part_liq;
  int n_part;
  int* n_fluid;
  const uint8_t* tflag0;
  int n_tiles;
  CgScalars* sc;
};
__global__ void __launch_bounds__(1024) k_mg_level_info(MgInfoArgs a, MgTailArgs tail) {
  if (blockIdx.x == gridDim.x - 1) {
    coarsen_tail(tail);
    if (a.part_liq) {
      __shared__ double sm[16];
      const double n = sum_part<1024>(a.part_liq, a.n_part, sm);
      int live = 0;
      for (int base = 0; base < a.n_tiles; base += 1024) {
        const int t = base + threadIdx.x;
        live += __syncthreads_count(t < a.n_tiles && a.tflag0[t] != 0);
      }
      if (threadIdx.x == 0) {
        *a.n_fluid = (int)n;
        a.sc->live_tiles = live;
      }
    }
    return;
  }
  if (a.tmap && blockIdx.x == gridDim.x - 2) {
    tile_map_block(a.nb_map, a.tflag, a.tmap);
    return;
  }
  const int stride = (gridDim.x - (a.tmap ? 2 : 1)) * 1024;
  for (int l = 1; l <= a.nli; ++l) {
    const int nx = a.nx[l], ny = a.ny[l], n = nx * ny;
    for (int t = blockIdx.x * 1024 + threadIdx.x; t < n; t += stride)
      a.inf[l][t] = make_info(a.lab[l], nx, ny, t % nx, t / nx);
  }
}

// ---------------------------------------------------------------- V-cycle
// Down: level l (fine) -> level l+1.  A block owns 16 x 16 coarse cells (32 x 32
// fine cells).  In LDS: r and info on the fine halo [-3, 35), x1 = w r / n there,
// x2 (second sweep) on [-2, 34), the residual on [-1, 33), then the restriction
// (1D weights 1/4, 3/4, 3/4, 1/4 over fine 2I-1 .. 2I+2).  The kernel given
// check != 0 first runs the stop test of the CG iteration on the partials of |r|^2:
// level 0's, or level 1's in the fused form (where level 0 forms the residual).
struct MgCheck {
  CgScalars* sc;
  const double* part_rr;
  int n_rr;
  const double* part_bb;
  int n_bb;
  double tol2;
  int max_iter;
  int external;  // 1: the stop test ran in its own kernel (per-scene CG); read `done` only
};

// Every block takes the same decision from the same partials.  `done` is read by
// one thread and broadcast, since block 0 may set it while other blocks start.
template <int NT>
__device__ __forceinline__ bool mg_check(const MgCheck& k, int it, double* sm) {
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = ld_dev(&k.sc->done);
  __syncthreads();
  if (s_done || k.external) return s_done != 0;
  const double rr = sum_part<NT>(k.part_rr, k.n_rr, sm);
  double bb, thr;
  if (it == 0) {
    bb = sum_part<NT>(k.part_bb, k.n_bb, sm);
    thr = k.tol2 * bb;
    if (thr < DBL_MIN) thr = DBL_MIN;
  } else {
    bb = k.sc->rhs_norm2;
    thr = k.sc->threshold;
  }
  const bool stop = bb == 0.0 || rr < thr || it >= k.max_iter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it == 0) {
      k.sc->rhs_norm2 = bb;
      k.sc->threshold = thr;
    }
    k.sc->rr = rr;
    k.sc->err = bb > 0.0 ? sqrt(rr / bb) : 0.0;
    k.sc->iter = it;
    if (stop) st_dev(&k.sc->done, 1);
  }
  return stop;
}

struct MgLevel {
  int nx, ny;
  const uint8_t* inf;
  float* r;   // right-hand side of this level (level 0: the CG residual)
  float* xs;  // after pre-smoothing
  float* xo;  // after the full cycle (level 0: z)
};

// Fused form: level 0's down kernel of iteration it >= 1 first finishes iteration
// it - 1 on its tile, as k_mg_xr would: x_new = x_old + alpha p (FP64), r = b' - K
// x_new (FP64, stored as FP32) and the block's partial |r|^2.  The residual on the
// tile's halo is formed from x_new on a halo one cell wider, the same values the
// neighbouring tiles store.
struct MgXr {
  const double* x_old;
  double* x_new;
  const float* p;
  const double* b;
  const double* part_pq;  // p.Kp partials of iteration it - 1
  int n_pq;
  double* part_rr;  // out: one partial |r|^2 per block
};

__device__ __forceinline__ void coarse_stage(const MgCoarseArgs& a, double* part_rz,
                                             bool coherent);

// Smoothing weights of the tile levels below level 0 when they take S > 2 sweeps
// before and after the coarse correction (sweep k uses w[k - 1]); with S = 2 every
// level uses the kernels' omega, omega2.
struct MgSweeps {
  float w[kMgMaxSweeps];
};

// Block span clock (span != null): thread 0 stores the block's start and end ticks
// of the device wall clock in span[2 b], span[2 b + 1]; the launch's duration is the
// latest end minus the earliest start over its blocks.
struct BlockSpan {
  unsigned long long* span;
  unsigned long long t0;
  __device__ __forceinline__ explicit BlockSpan(unsigned long long* s)
      : span(s), t0(s ? wall_clock64() : 0ull) {}
  __device__ __forceinline__ void end() const {
    if (span && threadIdx.x == 0) {
      span[2 * blockIdx.x] = t0;
      span[2 * blockIdx.x + 1] = wall_clock64();
    }
  }
};

// S sweeps (2 on level 0, 2 .. kMgMaxSweeps on the levels below): the right-hand side
// and info on [-(S + 1), 33 + S), sweep k on a ring one cell narrower than sweep k - 1,
// the last on [-2, 34); the sweeps alternate between the x1 and x2 buffers.
template <bool L0, int NT, bool XR = false, bool TAIL = false, int S = 2>
__global__ void __launch_bounds__(NT) k_mg_down(MgLevel f, MgLevel c, float omega, float omega2,
                                                 int it,
                                                 MgCheck chk, int check,
                                                 const uint8_t* __restrict__ tflag, MgXr xr,
                                                 unsigned long long* span,
                                                 MgCoarseArgs ca, unsigned* __restrict__ ticket,
                                                 MgSweeps sw, const int* __restrict__ tmap) {
  static_assert(!XR || L0, "the fused update belongs to level 0");
  static_assert(!TAIL || !L0, "the one-workgroup stage follows the last tile level");
  static_assert(S >= 2 && S <= kMgMaxSweeps && (!L0 || S <= 3),
                "level 0 takes two or three sweeps; smaller levels 2 .. kMgMaxSweeps");
  const BlockSpan clock(span);
  __shared__ double sm[NT / 64];
  __shared__ int s_last;
  // TAIL (the last tile level, 1024 threads): the restriction is stored
  // device-coherent and the last workgroup to finish runs the one-workgroup stage on
  // it, in place of a k_mg_coarse launch.  The arrival count is ordered after every
  // wave's stores, as in a grid barrier, but nobody waits.
  // inlined: as a call its LDS accesses would go through flat addresses
  auto run_tail = [&]() __attribute__((always_inline)) {
    if constexpr (TAIL) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (threadIdx.x == 0)
        s_last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
      __syncthreads();
      if (s_last) {
        if (threadIdx.x == 0) st_dev(ticket, 0u);  // ready for the next launch
        coarse_stage(ca, nullptr, true);
      }
    }
  };
  auto put_r = [&](float* dst, float v) {
    if constexpr (TAIL)
      st_dev(dst, v);
    else
      *dst = v;
  };
  // the tile (and its flag) are read with the stop flag (one memory round trip)
  const int tm = tmap ? tmap[blockIdx.x] : 0;
  const int tile = tmap ? tm & ~kTileEmptyBit : xcd_tile(blockIdx.x, gridDim.x);
  const bool empty = tmap ? tm < 0 : tflag && !tflag[tile];
  constexpr int W3 = 34 + 2 * S, W2 = W3 - 2, W1 = 34;  // S = 2: 38, 36, 34
  constexpr int W4 = W3 + 2;  // fused form: x_new on [-(S + 2), 34 + S)
  constexpr int WC = 18;  // coarse cells of the tile and their neighbours
  const int tiles_x = (c.nx + 15) >> 4;
  // each XCD takes one band of consecutive tiles, so halo rows are shared in its L2
  const int bx = tile % tiles_x, by = tile / tiles_x;
  const int fi0 = bx * 32, fj0 = by * 32;
  const bool do_xr = XR && it > 0;
  // Without the fused update, every load of a tile holding unknowns is issued before
  // the stop test's reads (their round trips overlap; a finished solve drops them) and
  // before the first LDS store (a load outside the grid reads cell 0 and is replaced
  // by the boundary value).
  constexpr int NC = (WC * WC + NT - 1) / NT, NF = (W3 * W3 + NT - 1) / NT;
  uint8_t cv[NC], fv[NF];
  float rv[NF];
  if (!empty && !do_xr) {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int t = threadIdx.x + k * NT;
      const int I = bx * 16 + t % WC - 1, J = by * 16 + t / WC - 1;
      const bool in = t < WC * WC && I >= 0 && I < c.nx && J >= 0 && J < c.ny;
      const uint8_t v = c.inf[in ? I + J * c.nx : 0];
      cv[k] = in ? v : kWall;
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      const int t = threadIdx.x + k * NT;
      const int i = fi0 + t % W3 - (S + 1), j = fj0 + t / W3 - (S + 1);
      const bool in = t < W3 * W3 && i >= 0 && i < f.nx && j >= 0 && j < f.ny;
      const int q = in ? i + j * f.nx : 0;
      const uint8_t fq = f.inf[q];
      const float rq = f.r[q];
      fv[k] = in ? fq : 0;
      rv[k] = in ? rq : 0.f;
    }
  }
  if (check) {
    if (mg_check<NT>(chk, it, sm)) {
      clock.end();
      return;
    }
  } else if (chk.sc->done) {
    clock.end();
    return;
  }
  __shared__ float s_r[W3 * W3], s_x1[W3 * W3], s_x2[W2 * W2];
  // the last sweep lands in x2 (S even) or x1 (S odd); the residual reuses the other
  // buffer, dead by then, so the smaller footprint lets more tiles share a CU
  float* const s_res = (S & 1) ? s_x2 : s_x1;
  const float* const s_xs = (S & 1) ? s_x1 : s_x2;
  float wk[kMgMaxSweeps];
#pragma unroll
  for (int k = 0; k < kMgMaxSweeps; ++k) wk[k] = sw.w[k];
  if (S == 2) {
    wk[0] = omega;
    wk[1] = omega2;
  }
  __shared__ uint8_t s_inf[W3 * W3], s_cinf[WC * WC];
  __shared__ double s_xn[XR ? W4 * W4 : 1];
  if (empty) {
    // no unknown in the tile on either level: its outputs are zero; written once per
    // solve (earlier solves may have left other values)
    if (it == 0) {
      for (int t = threadIdx.x; t < 32 * 32; t += NT) {
        const int i = fi0 + (t & 31), j = fj0 + (t >> 5);
        if (i < f.nx && j < f.ny) f.xs[i + j * f.nx] = 0.f;
      }
      for (int t = threadIdx.x; t < 16 * 16; t += NT) {
        const int I = bx * 16 + (t & 15), J = by * 16 + (t >> 4);
        if (I < c.nx && J < c.ny) put_r(c.r + I + J * c.nx, 0.f);
      }
    }
    if (do_xr) {
      // x = r = 0 here: both x buffers once (iterations 1 and 2), r once; every
      // iteration's stop test sums every block's partial
      if (it <= 2)
        for (int t = threadIdx.x; t < 32 * 32; t += NT) {
          const int i = fi0 + (t & 31), j = fj0 + (t >> 5);
          if (i < f.nx && j < f.ny) {
            xr.x_new[i + j * f.nx] = 0.0;
            if (it == 1) f.r[i + j * f.nx] = 0.f;
          }
        }
      if (threadIdx.x == 0) xr.part_rr[blockIdx.x] = 0.0;
    }
    run_tail();
    clock.end();
    return;
  }
  if (do_xr) {
    // every load of the tile first (a load outside the grid reads cell 0 and is
    // replaced by the boundary value), then alpha, then the stores to LDS
    constexpr int N4 = (W4 * W4 + NT - 1) / NT;
    double bv[NF], xo[N4];
    float pv[N4];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int t = threadIdx.x + k * NT;
      const int I = bx * 16 + t % WC - 1, J = by * 16 + t / WC - 1;
      const bool in = t < WC * WC && I >= 0 && I < c.nx && J >= 0 && J < c.ny;
      const uint8_t v = c.inf[in ? I + J * c.nx : 0];
      cv[k] = in ? v : kWall;
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      const int t = threadIdx.x + k * NT;
      const int i = fi0 + t % W3 - (S + 1), j = fj0 + t / W3 - (S + 1);
      const bool in = t < W3 * W3 && i >= 0 && i < f.nx && j >= 0 && j < f.ny;
      const int q = in ? i + j * f.nx : 0;
      const uint8_t fq = f.inf[q];
      const double bq = xr.b[q];
      fv[k] = in ? fq : 0;
      bv[k] = in ? bq : 0.0;
    }
#pragma unroll
    for (int k = 0; k < N4; ++k) {
      const int t = threadIdx.x + k * NT;
      const int i = fi0 + t % W4 - (S + 2), j = fj0 + t / W4 - (S + 2);
      const bool in = t < W4 * W4 && i >= 0 && i < f.nx && j >= 0 && j < f.ny;
      const int q = in ? i + j * f.nx : 0;
      const double xq = xr.x_old[q];
      const float pq = xr.p[q];
      xo[k] = in ? xq : 0.0;
      pv[k] = in ? pq : 0.f;
    }
    const double pq = sum_part<NT>(xr.part_pq, xr.n_pq, sm);
    const double alpha = chk.sc->rho[(it - 1) & 1] / pq;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int t = threadIdx.x + k * NT;
      if (t < WC * WC) s_cinf[t] = cv[k];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      const int t = threadIdx.x + k * NT;
      if (t < W3 * W3) s_inf[t] = fv[k];
    }
#pragma unroll
    for (int k = 0; k < N4; ++k) {
      const int t = threadIdx.x + k * NT;
      if (t < W4 * W4) s_xn[t] = xo[k] + alpha * (double)pv[k];
    }
    __syncthreads();
    // r = b' - K x_new on [-(S + 1), 33 + S) (cells outside the grid hold x = 0 and
    // add 0); the tile's own cells store x_new and r and add to |r|^2
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      const int t = threadIdx.x + k * NT;
      if (t >= W3 * W3) break;
      const int li = t % W3, lj = t / W3;
      const uint8_t fq = fv[k];
      const int q4 = (li + 1) + (lj + 1) * W4;
      const double xc = s_xn[q4];
      double r64 = 0.0;
      if (is_act(fq)) {
        double sn = 0.0;
        sn += s_xn[q4 - 1];
        sn += s_xn[q4 + 1];
        sn += s_xn[q4 - W4];
        sn += s_xn[q4 + W4];
        r64 = bv[k] - ((double)(fq & 7) * xc - sn);
        acc += (li > S && li < 33 + S && lj > S && lj < 33 + S) ? r64 * r64 : 0.0;
      }
      const float rv = (float)r64;
      s_r[t] = rv;
      s_x1[t] = is_act(fq) ? wk[0] * rv * inv_nns(fq) : 0.f;
      const int i = fi0 + li - (S + 1), j = fj0 + lj - (S + 1);
      if (li > S && li < 33 + S && lj > S && lj < 33 + S && i < f.nx && j < f.ny) {
        xr.x_new[i + j * f.nx] = is_act(fq) ? xc : 0.0;
        f.r[i + j * f.nx] = rv;
      }
    }
    write_part<NT>(acc, xr.part_rr, sm);
    if (blockIdx.x == 0 && threadIdx.x == 0) chk.sc->spmv = it;
  } else {
    // the tile's loads were issued before the stop test
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int t = threadIdx.x + k * NT;
      if (t < WC * WC) s_cinf[t] = cv[k];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      const int t = threadIdx.x + k * NT;
      if (t < W3 * W3) {
        s_inf[t] = fv[k];
        s_r[t] = rv[k];
        s_x1[t] = is_act(fv[k]) ? wk[0] * rv[k] * inv_nns(fv[k]) : 0.f;
      }
    }
  }
  __syncthreads();
  // sweeps 2 .. S: from the buffer of window win (offset d inside the right-hand
  // side's window) to one of window win - 2; the last lands in s_xs ([-2, 34))
#pragma unroll
  for (int k = 2; k <= S; ++k) {
    const int win = W3 - 2 * (k - 2), wout = win - 2, d = k - 2;
    const float* xin = (k & 1) ? s_x2 : s_x1;
    float* xout = (k & 1) ? s_x1 : s_x2;
    const float wgt = wk[k - 1];
    for (int t = threadIdx.x; t < wout * wout; t += NT) {
      const int li = t % wout + 1, lj = t / wout + 1;
      const int q = li + lj * win;
      const int qr = (li + d) + (lj + d) * W3;
      const uint8_t fv = s_inf[qr];
      float v = 0.f;
      if (is_act(fv)) {
        const float n = nns_of(fv), x = xin[q];
        const float kx = n * x - (xin[q - 1] + xin[q + 1] + xin[q - win] + xin[q + win]);
        v = x + wgt * (s_r[qr] - kx) * inv_nns(fv);
      }
      xout[t] = v;
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < W1 * W1; t += NT) {
    const int li = t % W1, lj = t / W1;  // fine offset li - 1, lj - 1
    constexpr int WS = 36;               // the last sweep's window, [-2, 34)
    const int q = (li + 1) + (lj + 1) * WS;
    const int q3 = (li + S) + (lj + S) * W3;
    const uint8_t fv = s_inf[q3];
    const float x = s_xs[q];
    float res = 0.f;
    if (is_act(fv)) {
      const float kx = nns_of(fv) * x - (s_xs[q - 1] + s_xs[q + 1] + s_xs[q - WS] + s_xs[q + WS]);
      res = s_r[q3] - kx;
    }
    s_res[t] = res;
    const int oi = li - 1, oj = lj - 1;
    if (oi >= 0 && oi < 32 && oj >= 0 && oj < 32) {
      const int i = fi0 + oi, j = fj0 + oj;
      if (i < f.nx && j < f.ny) f.xs[i + j * f.nx] = x;
    }
  }
  __syncthreads();
  if (threadIdx.x < 256) {
    const int ci = threadIdx.x & 15, cj = threadIdx.x >> 4;
    const int I = bx * 16 + ci, J = by * 16 + cj;
    if (I < c.nx && J < c.ny) {
      const int cc = I + J * c.nx;
      float v = 0.f;
      if (is_act(s_cinf[(ci + 1) + (cj + 1) * WC])) {
        // fine offsets 2ci-1 .. 2ci+2 are s_res columns 2ci .. 2ci+3
        const float w[4] = {0.25f, 0.75f, 0.75f, 0.25f};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const float* row = s_res + (2 * cj + b) * W1 + 2 * ci;
          const float s = 0.25f * row[0] + 0.75f * row[1] + 0.75f * row[2] + 0.25f * row[3];
          v += w[b] * s;
        }
        // the four children moved the weight of their wall neighbours here
        auto ci_at = [&](int di, int dj) { return s_cinf[(ci + 1 + di) + (cj + 1 + dj) * WC]; };
#pragma unroll
        for (int b = 1; b <= 2; ++b)
#pragma unroll
          for (int a = 1; a <= 2; ++a) {
            const int dx = a == 1 ? -1 : 1, dy = b == 1 ? -1 : 1;
            const float m = moved_weight(ci_at(dx, 0), ci_at(0, dy), ci_at(dx, dy));
            v += m * s_res[(2 * cj + b) * W1 + 2 * ci + a];
          }
      }
      put_r(c.r + cc, v);
    }
  }
  run_tail();
  if (span) {
    __syncthreads();
    clock.end();
  }
}

// Up: level l+1 -> level l.  A block owns 32 x 32 fine cells.  In LDS: the coarse
// solution around the tile, r and info on [-S, 32 + S), the corrected x = xs + P e
// there, then S post-sweeps, each on a ring one cell narrower than the one before (the
// last on the tile, written to xo), alternating between the xc and y buffers.  S = 2
// (level 0), or 2 .. kMgMaxSweeps on the levels below (sweep k uses w[k - 1]).
// Level 0 also forms the partial of r.z (z = xo) for the next direction.
template <bool L0, int NT, int S = 2>
__global__ void __launch_bounds__(NT) k_mg_up(MgLevel f, MgLevel c, float omega, float omega2,
                                               const CgScalars* sc, double* __restrict__ part_rz,
                                               int it, const uint8_t* __restrict__ tflag,
                                               MgSweeps sw, const int* __restrict__ tmap) {
  static_assert(S >= 2 && S <= kMgMaxSweeps && (!L0 || S <= 3),
                "level 0 takes two or three sweeps; smaller levels 2 .. kMgMaxSweeps");
  constexpr int W2 = 32 + 2 * S, W1 = W2 - 2;  // S = 2: 36, 34
  // coarse cells from fi0 / 2 - (S / 2 + 1): every coarse neighbour of the fine cells
  // [-S, 32 + S)
  constexpr int WC = 18 + S, CO = S / 2 + 1;
  __shared__ float s_r[W2 * W2], s_xc[W2 * W2], s_y1[W1 * W1], s_e[WC * WC];
  __shared__ uint8_t s_inf[W2 * W2], s_cinf[WC * WC];
  __shared__ double sm[NT / 64];
  float wk[kMgMaxSweeps];
#pragma unroll
  for (int k = 0; k < kMgMaxSweeps; ++k) wk[k] = sw.w[k];
  if (S == 2) {
    wk[0] = omega;
    wk[1] = omega2;
  }
  const int tiles_x = (f.nx + 31) >> 5;
  // each XCD takes one band of consecutive tiles, so halo rows are shared in its L2
  // (or the live tiles' share of them, tmap); the tile, its flag and the stop flag are
  // read together (one memory round trip)
  const int tm = tmap ? tmap[blockIdx.x] : 0;
  const int tile = tmap ? tm & ~kTileEmptyBit : xcd_tile(blockIdx.x, gridDim.x);
  const bool empty = tmap ? tm < 0 : tflag && !tflag[tile];
  const int bx = tile % tiles_x, by = tile / tiles_x;
  const int fi0 = bx * 32, fj0 = by * 32;
  const int ci0 = (fi0 >> 1) - CO, cj0 = (fj0 >> 1) - CO;  // coarse frame origin
  // every load of a tile holding unknowns is issued before the stop flag's value is
  // known (one memory round trip for both; the loads of a finished solve are dropped)
  // and before the first LDS store (a load outside the grid reads cell 0 and is
  // replaced by the boundary value)
  constexpr int NC = (WC * WC + NT - 1) / NT, NF = (W2 * W2 + NT - 1) / NT;
  uint8_t cv[NC], fv[NF];
  float ev[NC], rv[NF], xv[NF];
  if (!empty) {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int t = threadIdx.x + k * NT;
      const int I = ci0 + t % WC, J = cj0 + t / WC;
      const bool in = t < WC * WC && I >= 0 && I < c.nx && J >= 0 && J < c.ny;
      const int q = in ? I + J * c.nx : 0;
      const float e = c.xo[q];
      const uint8_t ci = c.inf[q];
      ev[k] = in ? e : 0.f;
      cv[k] = in ? ci : kWall;
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      const int t = threadIdx.x + k * NT;
      const int i = fi0 + t % W2 - S, j = fj0 + t / W2 - S;
      const bool in = t < W2 * W2 && i >= 0 && i < f.nx && j >= 0 && j < f.ny;
      const int q = in ? i + j * f.nx : 0;
      const uint8_t fq = f.inf[q];
      const float rq = f.r[q], xq = f.xs[q];
      fv[k] = in ? fq : 0;
      rv[k] = in ? rq : 0.f;
      xv[k] = in ? xq : 0.f;
    }
  }
  if (sc->done) return;
  if (empty) {
    // no unknown in the tile: z = 0 there (written once per solve) and r.z adds 0
    if (it == 0)
      for (int t = threadIdx.x; t < 32 * 32; t += NT) {
        const int i = fi0 + (t & 31), j = fj0 + (t >> 5);
        if (i < f.nx && j < f.ny) f.xo[i + j * f.nx] = 0.f;
      }
    if (L0 && threadIdx.x == 0) part_rz[blockIdx.x] = 0.0;
    return;
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int t = threadIdx.x + k * NT;
    if (t < WC * WC) {
      s_e[t] = ev[k];
      s_cinf[t] = cv[k];
    }
  }
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const int t = threadIdx.x + k * NT;
    if (t < W2 * W2) {
      s_inf[t] = fv[k];
      s_r[t] = rv[k];
      s_xc[t] = xv[k];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < W2 * W2; t += NT) {
    if (!is_act(s_inf[t])) {
      s_xc[t] = 0.f;
      continue;
    }
    const int i = fi0 + t % W2 - S, j = fj0 + t / W2 - S;
    const int I = i >> 1, J = j >> 1;  // floor for the halo left of / below 0 too
    const int In = (i & 1) ? I + 1 : I - 1;
    const int Jn = (j & 1) ? J + 1 : J - 1;
    const int q00 = (I - ci0) + (J - cj0) * WC, q10 = (In - ci0) + (J - cj0) * WC;
    const int q01 = (I - ci0) + (Jn - cj0) * WC, q11 = (In - ci0) + (Jn - cj0) * WC;
    s_xc[t] += prolong_value(s_e[q00], s_e[q10], s_cinf[q10], s_e[q01], s_cinf[q01], s_e[q11],
                             s_cinf[q11]);
  }
  __syncthreads();
  // post-sweeps 1 .. S-1: from the buffer of window win (offset d inside the
  // right-hand side's window) to one of window win - 2
#pragma unroll
  for (int k = 1; k < S; ++k) {
    const int win = W2 - 2 * (k - 1), wout = win - 2, d = k - 1;
    const float* xin = (k & 1) ? s_xc : s_y1;
    float* xout = (k & 1) ? s_y1 : s_xc;
    const float wgt = wk[k - 1];
    for (int t = threadIdx.x; t < wout * wout; t += NT) {
      const int li = t % wout + 1, lj = t / wout + 1;
      const int q = li + lj * win;
      const int qr = (li + d) + (lj + d) * W2;
      const uint8_t fv = s_inf[qr];
      float v = 0.f;
      if (is_act(fv)) {
        const float n = nns_of(fv), x = xin[q];
        const float kx = n * x - (xin[q - 1] + xin[q + 1] + xin[q - win] + xin[q + win]);
        v = x + wgt * (s_r[qr] - kx) * inv_nns(fv);
      }
      xout[t] = v;
    }
    __syncthreads();
  }
  // the last post-sweep, on the tile, from the window [-1, 33) (in y1 for S even,
  // in xc for S odd)
  constexpr int WL = 34;
  const float* const s_yl = (S & 1) ? s_xc : s_y1;
  double rz = 0.0;
  const float wl = wk[S - 1];
  for (int t = threadIdx.x; t < 32 * 32; t += NT) {
    const int oi = t & 31, oj = t >> 5;
    const int i = fi0 + oi, j = fj0 + oj;
    if (i >= f.nx || j >= f.ny) continue;
    const int q = (oi + 1) + (oj + 1) * WL;
    const int q2 = (oi + S) + (oj + S) * W2;
    const uint8_t fv = s_inf[q2];
    float v = 0.f;
    if (is_act(fv)) {
      const float n = nns_of(fv), x = s_yl[q];
      const float kx = n * x - (s_yl[q - 1] + s_yl[q + 1] + s_yl[q - WL] + s_yl[q + WL]);
      v = x + wl * (s_r[q2] - kx) * inv_nns(fv);
    }
    f.xo[i + j * f.nx] = v;
    if (L0) rz += (double)s_r[q2] * (double)v;
  }
  if (L0) write_part<NT>(rz, part_rz, sm);
}

// The levels lc .. nlev-1 (each <= 1024 cells) in one workgroup, all in LDS: the
// same V-cycle as the tile kernels.  The info bytes of all these levels are one
// contiguous array, loaded in one pass.  The coarsest level (<= 4 x 4 cells) is
// swept by one wavefront, a cell per lane, neighbours by lane shuffles, with no
// barrier between sweeps.
// Neighbour sum of cell c = (i, j) of an nx x ny level in LDS; a neighbour outside the
// level adds 0.  Branch-free: the four reads are issued together.
__device__ __forceinline__ float nbr_sum(const float* a, int c, int i, int j, int nx, int ny) {
  const bool ml = i > 0, mr = i < nx - 1, md = j > 0, mu = j < ny - 1;
  const float vl = a[ml ? c - 1 : c], vr = a[mr ? c + 1 : c];
  const float vd = a[md ? c - nx : c], vu = a[mu ? c + nx : c];
  float s = 0.f;
  s += ml ? vl : 0.f;
  s += mr ? vr : 0.f;
  s += md ? vd : 0.f;
  s += mu ? vu : 0.f;
  return s;
}

constexpr int kCoarseThreads = 1024;
constexpr float kCoarsestOmega = 0.8f;  // damped Jacobi on the coarsest level (<= 64 cells)

// The one-workgroup V-cycle over the levels 0 .. nl-1 of a small hierarchy held in
// LDS: R, X, T and INF, level l at offset off[l] (off[0] = 0).  R's level 0 holds
// the right-hand side; the coarser right-hand sides are formed here.  Level 0's
// result is left in X when nl == 1, else in T.
// Inlined into its kernels so that R, X, T and INF are addressed as LDS (a called
// function sees generic pointers and goes through the slower flat path).
__device__ __forceinline__ void coarse_cycle(int nl, const int* __restrict__ lnx,
                                             const int* __restrict__ lny,
                                             const int* __restrict__ loff, float w, float w2,
                                             int sweeps, float* R, float* X, float* T,
                                             const uint8_t* INF, int trunc) {
  if (trunc > 0) {
    // truncated: `trunc` damped-Jacobi sweeps from zero on the first level only (a
    // polynomial in D^-1 K, so the preconditioner stays symmetric positive definite),
    // alternating between X and T so that the last one lands in T
    const int nx = lnx[0], ny = lny[0], n = nx * ny;
    const float inv_nx = 1.0f / (float)nx;
    for (int sw = 0; sw < trunc; ++sw) {
      const float* src = ((trunc - sw) & 1) ? X : T;
      float* dst = ((trunc - sw) & 1) ? T : X;
      for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
        float v = 0.f;
        if (is_act(INF[q])) {
          const float rn = inv_nns(INF[q]);
          if (sw == 0) {
            v = kCoarsestOmega * R[q] * rn;
          } else {
            int i, j;
            small_divmod(q, nx, inv_nx, i, j);
            const float x = src[q];
            v = x + kCoarsestOmega * (R[q] - (nns_of(INF[q]) * x - nbr_sum(src, q, i, j, nx, ny))) * rn;
          }
        }
        dst[q] = v;
      }
      __syncthreads();
    }
    return;
  }
  for (int l = 0; l < nl - 1; ++l) {
    const int nx = lnx[l], ny = lny[l], n = nx * ny, o = loff[l];
    float *r = R + o, *x = X + o, *t = T + o;
    const uint8_t* inf = INF + o;
    const float inv_nx = 1.0f / (float)nx;
    // two sweeps from zero; the first sweep's neighbour values w r / n formed on the fly
    for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
      float v = 0.f;
      if (is_act(inf[q])) {
        int i, j;
        small_divmod(q, nx, inv_nx, i, j);
        const bool mk[4] = {i > 0, i < nx - 1, j > 0, j < ny - 1};
        const int qk[4] = {mk[0] ? q - 1 : q, mk[1] ? q + 1 : q, mk[2] ? q - nx : q,
                           mk[3] ? q + nx : q};
        uint8_t fk[4];
        float rk[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          fk[k] = inf[qk[k]];
          rk[k] = r[qk[k]];
        }
        float sn = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          sn += (mk[k] && is_act(fk[k])) ? w * rk[k] * inv_nns(fk[k]) : 0.f;
        const float nn = nns_of(inf[q]), xc = w * r[q] * inv_nns(inf[q]);
        v = xc + w2 * (r[q] - (nn * xc - sn)) * inv_nns(inf[q]);
      }
      x[q] = v;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
      float v = 0.f;
      if (is_act(inf[q])) {
        int i, j;
        small_divmod(q, nx, inv_nx, i, j);
        v = r[q] - (nns_of(inf[q]) * x[q] - nbr_sum(x, q, i, j, nx, ny));
      }
      t[q] = v;
    }
    __syncthreads();
    const int cnx = lnx[l + 1], cny = lny[l + 1], co = loff[l + 1];
    const float inv_cnx = 1.0f / (float)cnx;
    for (int q = threadIdx.x; q < cnx * cny; q += kCoarseThreads) {
      int I, J;
      small_divmod(q, cnx, inv_cnx, I, J);
      float v = 0.f;
      if (is_act(INF[co + q])) {
        // fine rows / columns 2J-1 .. 2J+2 and 2I-1 .. 2I+2, clamped into the level
        // with weight 0 outside it; every read is issued before the sums
        const float wt[4] = {0.25f, 0.75f, 0.75f, 0.25f};
        bool rin[4], cin[4];
        int row[4], col[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int j = 2 * J - 1 + b, i = 2 * I - 1 + b;
          rin[b] = j >= 0 && j < ny;
          cin[b] = i >= 0 && i < nx;
          row[b] = (rin[b] ? j : 0) * nx;
          col[b] = cin[b] ? i : 0;
        }
        float tv[4][4];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int e = 0; e < 4; ++e) tv[b][e] = t[col[e] + row[b]];
        // coarse neighbours (I -/+ 1, J), (I, J -/+ 1), (I -/+ 1, J -/+ 1); SOLID outside
        uint8_t cn[3][3];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
          for (int dx = 0; dx < 3; ++dx) {
            const int ii = I + dx - 1, jj = J + dy - 1;
            const bool in = ii >= 0 && ii < cnx && jj >= 0 && jj < cny;
            const uint8_t fv = INF[co + (in ? ii + jj * cnx : q)];
            cn[dy][dx] = in ? fv : kWall;
          }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          float srow = 0.f;
#pragma unroll
          for (int e = 0; e < 4; ++e) srow += cin[e] ? wt[e] * tv[b][e] : 0.f;
          v += rin[b] ? wt[b] * srow : 0.f;
        }
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            // child (2I + e, 2J + b) is tv[b + 1][e + 1]
            const int dx = e == 0 ? 0 : 2, dy = b == 0 ? 0 : 2;
            const float m = moved_weight(cn[1][dx], cn[dy][1], cn[dy][dx]);
            v += (rin[b + 1] && cin[e + 1]) ? m * tv[b + 1][e + 1] : 0.f;
          }
      }
      R[co + q] = v;
    }
    __syncthreads();
  }
  {
    // coarsest level: <= 64 cells (checked on the host), one wavefront
    const int l = nl - 1;
    const int nx = lnx[l], ny = lny[l], n = nx * ny, o = loff[l];
    if (threadIdx.x < 64) {
      const int q = threadIdx.x;
      const bool in = q < n;
      const uint8_t f = in ? INF[o + q] : 0;
      const float r = in ? R[o + q] : 0.f;
      const int i = in ? q % nx : 0, j = in ? q / nx : 0;
      const float nn = nns_of(f), rn = inv_nns(f);
      float x = 0.f;
      for (int sw = 0; sw < sweeps; ++sw) {
        const float xl = __shfl(x, q - 1 < 0 ? 0 : q - 1, 64);
        const float xr = __shfl(x, q + 1 > 63 ? 63 : q + 1, 64);
        const float xd = __shfl(x, q - nx < 0 ? 0 : q - nx, 64);
        const float xu = __shfl(x, q + nx > 63 ? 63 : q + nx, 64);
        float s = 0.f;
        if (i > 0) s += xl;
        if (i < nx - 1) s += xr;
        if (j > 0) s += xd;
        if (j < ny - 1) s += xu;
        x = is_act(f) ? x + kCoarsestOmega * (r - (nn * x - s)) * rn : 0.f;
      }
      if (in) X[o + q] = x;
    }
    __syncthreads();
  }
  // up: t = xs + P e, x = sweep(t), t = sweep(x); a level's result is in T (the
  // coarsest level's in X)
  for (int l = nl - 2; l >= 0; --l) {
    const int nx = lnx[l], ny = lny[l], n = nx * ny, o = loff[l];
    const int cnx = lnx[l + 1], cny = lny[l + 1], co = loff[l + 1];
    float *r = R + o, *x = X + o, *t = T + o;
    const uint8_t* inf = INF + o;
    const float* e = (l + 1 == nl - 1 ? X : T) + co;
    const uint8_t* cinf = INF + co;
    const float inv_nx = 1.0f / (float)nx;
    for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
      float v = 0.f;
      if (is_act(inf[q])) {
        int i, j;
        small_divmod(q, nx, inv_nx, i, j);
        const int I = i >> 1, J = j >> 1;
        const int In = (i & 1) ? I + 1 : I - 1;
        const int Jn = (j & 1) ? J + 1 : J - 1;
        const bool iok = In >= 0 && In < cnx, jok = Jn >= 0 && Jn < cny;
        // neighbours outside the coarse level read (I, J) and are replaced below
        const int Ic = iok ? In : I, Jc = jok ? Jn : J;
        const float e00 = e[I + J * cnx], e10r = e[Ic + J * cnx];
        const float e01r = e[I + Jc * cnx], e11r = e[Ic + Jc * cnx];
        const uint8_t f10r = cinf[Ic + J * cnx], f01r = cinf[I + Jc * cnx];
        const uint8_t f11r = cinf[Ic + Jc * cnx];
        const float e10 = iok ? e10r : 0.f, e01 = jok ? e01r : 0.f;
        const float e11 = iok && jok ? e11r : 0.f;
        const uint8_t f10 = iok ? f10r : kWall, f01 = jok ? f01r : kWall;
        const uint8_t f11 = iok && jok ? f11r : kWall;
        v = x[q] + prolong_value(e00, e10, f10, e01, f01, e11, f11);
      }
      t[q] = v;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
      float v = 0.f;
      if (is_act(inf[q])) {
        const float nn = nns_of(inf[q]);
        int i, j;
        small_divmod(q, nx, inv_nx, i, j);
        v = t[q] + w * (r[q] - (nn * t[q] - nbr_sum(t, q, i, j, nx, ny))) * inv_nns(inf[q]);
      }
      x[q] = v;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
      float v = 0.f;
      if (is_act(inf[q])) {
        const float nn = nns_of(inf[q]);
        int i, j;
        small_divmod(q, nx, inv_nx, i, j);
        v = x[q] + w2 * (r[q] - (nn * x[q] - nbr_sum(x, q, i, j, nx, ny))) * inv_nns(inf[q]);
      }
      t[q] = v;
    }
    __syncthreads();
  }
}

// The kernel's own duration from the device wall clock (clk != null): clk[0] += end -
// start (ticks), clk[1] += 1, by thread 0 after the workgroup's last barrier.
__device__ __forceinline__ void kclock_end(unsigned long long* clk, unsigned long long t0) {
  if (clk && threadIdx.x == 0) {
    atomicAdd(&clk[0], wall_clock64() - t0);
    atomicAdd(&clk[1], 1ull);
  }
}

// With check != 0 (the hierarchy starts in this kernel) it runs level 0's
// convergence test first and forms the r.z partial at the end.
// The one-workgroup stage: load its levels' info and level lc's right-hand side into
// LDS, cycle, store level lc's result (and, when the hierarchy starts here, the r.z
// partial).  `coherent`: the right-hand side was written by other workgroups of the
// same launch with device-coherent stores, so it is read the same way.
__device__ __forceinline__ void coarse_stage(const MgCoarseArgs& a, double* part_rz,
                                             bool coherent) {
  __shared__ double sm[2 * 16];
  __shared__ int s_lv[3 * kMgMaxLevels];
  extern __shared__ float lds[];
  const int tot = a.total;  // a multiple of 4
  float* R = lds;
  float* X = R + tot;
  float* T = X + tot;
  uint8_t* INF = reinterpret_cast<uint8_t*>(T + tot);
  const int nl = a.nlev - a.lc;
  if (threadIdx.x < nl) {
    const int l = a.lc + threadIdx.x;
    s_lv[threadIdx.x] = a.nx[l];
    s_lv[kMgMaxLevels + threadIdx.x] = a.ny[l];
    s_lv[2 * kMgMaxLevels + threadIdx.x] = a.off[l] - a.off[a.lc];
  }
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.inf_all);
    uint32_t* dst = reinterpret_cast<uint32_t*>(INF);
    for (int q = threadIdx.x; q < tot / 4; q += kCoarseThreads) dst[q] = src[q];
    const int n0 = a.nx[a.lc] * a.ny[a.lc];
    for (int q = threadIdx.x; q < n0; q += kCoarseThreads)
      R[q] = coherent ? ld_dev(a.r_in + q) : a.r_in[q];
  }
  __syncthreads();
  coarse_cycle(nl, s_lv, s_lv + kMgMaxLevels, s_lv + 2 * kMgMaxLevels, a.omega, a.omega2,
               a.sweeps, R, X, T, INF, a.trunc);
  const int n = a.nx[a.lc] * a.ny[a.lc];
  const float* res = nl == 1 && a.trunc == 0 ? X : T;
  double rz = 0.0;
  for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
    a.x_out[q] = res[q];
    rz += (double)R[q] * (double)res[q];
  }
  if (part_rz) {
    double v[1] = {rz};
    block_sum<1, kCoarseThreads>(v, sm);
    if (threadIdx.x == 0) part_rz[0] = v[0];
  }
}

__global__ void __launch_bounds__(kCoarseThreads) k_mg_coarse(MgCoarseArgs a, MgCheck chk, int it,
                                                             int check,
                                                             double* __restrict__ part_rz,
                                                             unsigned long long* clk) {
  const unsigned long long t0 = clk ? wall_clock64() : 0ull;
  __shared__ double sm[2 * 16];
  if (check) {
    if (mg_check<kCoarseThreads>(chk, it, sm)) {
      kclock_end(clk, t0);
      return;
    }
  } else if (chk.sc->done) {
    kclock_end(clk, t0);
    return;
  }
  coarse_stage(a, part_rz, false);
  if (clk) {
    __syncthreads();
    kclock_end(clk, t0);
  }
}

// ---------------------------------------------------------------- per-scene coarse
// Packed batch of scenes: from level `la` of the packed hierarchy on, every scene
// has a hierarchy of its own (its rectangle of level la, cells outside it count as
// SOLID), solved by one workgroup per scene.  Scene rectangles are separated by at
// least one all-SOLID cell of level la, so levels 0 .. la never couple two scenes
// either, and the preconditioner is block-diagonal over the scenes.

// Labels and info of every scene's own levels (one workgroup per scene; dynamic LDS:
// the largest scene's cells over all its levels).
__global__ void __launch_bounds__(kCoarseThreads) k_mg_scene_setup(SceneCoarseArgs a) {
  __shared__ SceneCoarse t;
  extern __shared__ uint8_t slab[];
  if (threadIdx.x == 0) t = a.tab[blockIdx.x];
  __syncthreads();
  {
    const int nx = t.nx[0], n = nx * t.ny[0];
    for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
      const int i = q % nx, j = q / nx;
      slab[q] = a.lab[(size_t)(t.cx + i) + (size_t)(t.cy + j) * a.pitch];
    }
  }
  __syncthreads();
  for (int l = 1; l < t.nl; ++l) {
    const int fnx = t.nx[l - 1], fny = t.ny[l - 1], cnx = t.nx[l], n = cnx * t.ny[l];
    const uint8_t* fl = slab + t.off[l - 1];
    for (int q = threadIdx.x; q < n; q += kCoarseThreads) {
      const int I = q % cnx, J = q / cnx;
      bool air = false, liq = false;
      for (int b = 0; b < 2; ++b)
        for (int e = 0; e < 2; ++e) {
          const int i = 2 * I + e, j = 2 * J + b;
          if (i < fnx && j < fny) {
            const uint8_t L = fl[i + j * fnx];
            air |= L == kAir;
            liq |= L == kLiquid;
          }
        }
      slab[t.off[l] + q] = air ? kAir : (liq ? kLiquid : kSolid);
    }
    __syncthreads();
  }
  uint8_t* out = a.inf_all + (size_t)blockIdx.x * a.stride;
  for (int l = 0; l < t.nl; ++l) {
    const int nx = t.nx[l], n = nx * t.ny[l];
    for (int q = threadIdx.x; q < n; q += kCoarseThreads)
      out[t.off[l] + q] = make_info(slab + t.off[l], nx, t.ny[l], q % nx, q / nx);
  }
}

// One V-cycle of every scene's own levels: right-hand side from the packed level la,
// result back into it (cells outside every rectangle keep their zero).
__global__ void __launch_bounds__(kCoarseThreads) k_mg_scene_coarse(SceneCoarseArgs a,
                                                                   const CgScalars* sc) {
  __shared__ SceneCoarse t;
  if (sc->done) return;
  if (threadIdx.x == 0) t = a.tab[blockIdx.x];
  __syncthreads();
  extern __shared__ float lds[];
  const int tot = a.stride;  // a multiple of 4, >= t.total
  float* R = lds;
  float* X = R + tot;
  float* T = X + tot;
  uint8_t* INF = reinterpret_cast<uint8_t*>(T + tot);
  const int nx = t.nx[0], n0 = nx * t.ny[0];
  {
    const uint32_t* src =
        reinterpret_cast<const uint32_t*>(a.inf_all + (size_t)blockIdx.x * a.stride);
    uint32_t* dst = reinterpret_cast<uint32_t*>(INF);
    for (int q = threadIdx.x; q < t.total / 4; q += kCoarseThreads) dst[q] = src[q];
    for (int q = threadIdx.x; q < n0; q += kCoarseThreads)
      R[q] = a.r_in[(size_t)(t.cx + q % nx) + (size_t)(t.cy + q / nx) * a.pitch];
  }
  __syncthreads();
  const int cl = t.nl - 1;
  const int sweeps = 2 * (t.nx[cl] > t.ny[cl] ? t.nx[cl] : t.ny[cl]) + 8;
  coarse_cycle(t.nl, t.nx, t.ny, t.off, a.omega, a.omega2, sweeps, R, X, T, INF, a.trunc);
  const float* res = t.nl == 1 && a.trunc == 0 ? X : T;
  for (int q = threadIdx.x; q < n0; q += kCoarseThreads)
    a.x_out[(size_t)(t.cx + q % nx) + (size_t)(t.cy + q / nx) * a.pitch] = res[q];
}

// ---------------------------------------------------------------- CG vectors
// The cells of one thread's pass of k_mg_dir (FIRST: p = z).
template <bool FIRST>
__device__ __forceinline__ double mg_dir_cells(int nx, int ny, const uint8_t* __restrict__ inf,
                                               const float* __restrict__ z,
                                               const float* __restrict__ p_old,
                                               float* __restrict__ p_new, float bf) {
  const int C = nx * ny, stride = gridDim.x * kT;
  double acc = 0.0;
  CellWalk w(blockIdx.x * kT + threadIdx.x, stride, nx);
  for (int c0 = blockIdx.x * kT + threadIdx.x; c0 < C; c0 += kVecU * stride) {
    Stencil5 st[kVecU];
    uint8_t f[kVecU];
    float zv[kVecU][5], pv[kVecU][5];
#pragma unroll
    for (int u = 0; u < kVecU; ++u) {
      const int c = c0 + u * stride;
      const bool in = c < C;
      st[u] = stencil5(in ? c : C - 1, w.i, w.j, nx, ny, in);
      w.next();
      f[u] = inf[st[u].q[0]];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        zv[u][k] = z[st[u].q[k]];
        if (!FIRST) pv[u][k] = p_old[st[u].q[k]];
      }
    }
#pragma unroll
    for (int u = 0; u < kVecU; ++u) {
      if (!st[u].m[0]) break;
      float pn[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const float v = FIRST ? zv[u][k] : zv[u][k] + bf * pv[u][k];
        pn[k] = st[u].m[k] ? v : 0.f;
      }
      float pc = 0.f;
      if (is_act(f[u])) {
        pc = pn[0];
        float sn = 0.f;
        sn += pn[1];
        sn += pn[2];
        sn += pn[3];
        sn += pn[4];
        const float kp = nns_of(f[u]) * pc - sn;
        acc += (double)pc * (double)kp;
      }
      p_new[st[u].q[0]] = pc;
    }
  }
  return acc;
}

// p = z + beta p_old (beta = 0 at the first iteration), K p, partial p.Kp.
// rho = r.z is summed from the V-cycle's partials; block 0 records it.
__global__ void __launch_bounds__(kT) k_mg_dir(int nx, int ny, const uint8_t* __restrict__ inf,
                                                const float* __restrict__ z,
                                                const float* __restrict__ p_old,
                                                float* __restrict__ p_new, int it,
                                                const double* __restrict__ part_rz, int n_rz,
                                                double* __restrict__ part_pq, CgScalars* sc) {
  __shared__ double sm[4];
  if (sc->done) return;
  const double rz = sum_part<kT>(part_rz, n_rz, sm);
  const double beta = it == 0 ? 0.0 : rz / sc->rho[(it - 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->rho[it & 1] = rz;
  const double acc = it == 0 ? mg_dir_cells<true>(nx, ny, inf, z, p_old, p_new, 0.f)
                             : mg_dir_cells<false>(nx, ny, inf, z, p_old, p_new, (float)beta);
  write_part<kT>(acc, part_pq, sm);
}

// The same on level 0's 32 x 32 tiles (one block per tile, 4 rows of a column per
// thread), skipping the tiles that hold no unknown (tflag, as the level-0 tile
// kernels): their p is zero, written in both ping-pong buffers once per solve
// (iterations 0 and 1), and their p.Kp partial is 0.
__global__ void __launch_bounds__(kT) k_mg_dir_tiles(int nx, int ny,
                                                      const uint8_t* __restrict__ inf,
                                                      const float* __restrict__ z,
                                                      const float* __restrict__ p_old,
                                                      float* __restrict__ p_new, int it,
                                                      const double* __restrict__ part_rz,
                                                      int n_rz, double* __restrict__ part_pq,
                                                      CgScalars* sc,
                                                      const uint8_t* __restrict__ tflag,
                                                      const int* __restrict__ tmap) {
  constexpr int R = 32 * 32 / kT;  // cells per thread
  __shared__ double sm[4];
  const int tiles_x = (nx + 31) >> 5;
  // the tile, its flag and the stop flag are read together (one memory round trip)
  const int tm = tmap ? tmap[blockIdx.x] : 0;
  const int tile = tmap ? tm & ~kTileEmptyBit : xcd_tile(blockIdx.x, gridDim.x);
  const bool empty = tmap ? tm < 0 : !tflag[tile];
  const int bx = tile % tiles_x, by = tile / tiles_x;
  const int i = bx * 32 + (threadIdx.x & 31), j0 = by * 32 + (threadIdx.x >> 5);
  if (sc->done) return;
  if (empty) {
    if (it <= 1)
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int j = j0 + k * (kT / 32);
        if (i < nx && j < ny) p_new[i + j * nx] = 0.f;
      }
    if (threadIdx.x == 0) part_pq[blockIdx.x] = 0.0;
    if (blockIdx.x != 0) return;
  }
  // block 0 records rho = r.z, whether its tile is empty or not
  const double rz = sum_part<kT>(part_rz, n_rz, sm);
  const double beta = it == 0 ? 0.0 : rz / sc->rho[(it - 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->rho[it & 1] = rz;
  if (empty) return;
  const float bf = (float)beta;
  // every load of the thread's cells first (cells outside the grid read their own
  // clamped index and are masked), then the stencil
  Stencil5 st[R];
  uint8_t f[R];
  float zv[R][5], pv[R][5];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int j = j0 + k * (kT / 32);
    const bool in = i < nx && j < ny;
    const int c = in ? i + j * nx : 0;
    st[k] = stencil5(c, in ? i : 0, in ? j : 0, nx, ny, in);
    f[k] = inf[st[k].q[0]];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      zv[k][q] = z[st[k].q[q]];
      pv[k][q] = it == 0 ? 0.f : p_old[st[k].q[q]];
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (!st[k].m[0]) continue;
    float pn[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const float v = it == 0 ? zv[k][q] : zv[k][q] + bf * pv[k][q];
      pn[q] = st[k].m[q] ? v : 0.f;
    }
    float pc = 0.f;
    if (is_act(f[k])) {
      pc = pn[0];
      float sn = 0.f;
      sn += pn[1];
      sn += pn[2];
      sn += pn[3];
      sn += pn[4];
      const float kp = nns_of(f[k]) * pc - sn;
      acc += (double)pc * (double)kp;
    }
    p_new[st[k].q[0]] = pc;
  }
  write_part<kT>(acc, part_pq, sm);
}

// x_new = x_old + alpha p (FP64), r = b' - K x_new (FP64, stored as FP32),
// partial |r|^2 (FP64).
__global__ void __launch_bounds__(kT) k_mg_xr(int nx, int ny, const uint8_t* __restrict__ inf,
                                               const float* __restrict__ p,
                                               const double* __restrict__ x_old,
                                               double* __restrict__ x_new,
                                               const double* __restrict__ b, float* __restrict__ r,
                                               int it, const double* __restrict__ part_pq, int n_pq,
                                               double* __restrict__ part_rr, CgScalars* sc) {
  __shared__ double sm[4];
  if (sc->done) return;
  const double pq = sum_part<kT>(part_pq, n_pq, sm);
  const double alpha = sc->rho[it & 1] / pq;
  const int C = nx * ny, stride = gridDim.x * kT;
  double acc = 0.0;
  CellWalk w(blockIdx.x * kT + threadIdx.x, stride, nx);
  for (int c0 = blockIdx.x * kT + threadIdx.x; c0 < C; c0 += kVecU * stride) {
    Stencil5 st[kVecU];
    uint8_t f[kVecU];
    double xv[kVecU][5], bv[kVecU];
    float pv[kVecU][5];
#pragma unroll
    for (int u = 0; u < kVecU; ++u) {
      const int c = c0 + u * stride;
      const bool in = c < C;
      st[u] = stencil5(in ? c : C - 1, w.i, w.j, nx, ny, in);
      w.next();
      f[u] = inf[st[u].q[0]];
      bv[u] = b[st[u].q[0]];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        xv[u][k] = x_old[st[u].q[k]];
        pv[u][k] = p[st[u].q[k]];
      }
    }
#pragma unroll
    for (int u = 0; u < kVecU; ++u) {
      if (!st[u].m[0]) break;
      double xn = 0.0;
      float rv = 0.f;
      if (is_act(f[u])) {
        double xs[5];
#pragma unroll
        for (int k = 0; k < 5; ++k)
          xs[k] = st[u].m[k] ? xv[u][k] + alpha * (double)pv[u][k] : 0.0;
        xn = xs[0];
        double sn = 0.0;
        sn += xs[1];
        sn += xs[2];
        sn += xs[3];
        sn += xs[4];
        const double r64 = bv[u] - ((double)(f[u] & 7) * xn - sn);
        rv = (float)r64;
        acc += r64 * r64;
      }
      x_new[st[u].q[0]] = xn;
      r[st[u].q[0]] = rv;
    }
  }
  write_part<kT>(acc, part_rr, sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->spmv = it + 1;
}

// Partials of sum(n_i x_i), sum(n_i) (n_i = K_ii) and of the AIR links over the
// unknowns: n_i minus the unknown neighbours (a non-SOLID neighbour of an unknown is
// an unknown or AIR).
__global__ void __launch_bounds__(kT) k_mg_wsum(int nx, int ny, const uint8_t* __restrict__ inf,
                                                 const double* __restrict__ x0,
                                                 const double* __restrict__ x1,
                                                 const CgScalars* sc, double* __restrict__ part_nx,
                                                 double* __restrict__ part_n,
                                                 double* __restrict__ part_d) {
  __shared__ double sm[4];
  const int C = nx * ny;
  const double* x = (sc->iter & 1) ? x1 : x0;
  double a = 0.0, n = 0.0, d = 0.0;
  CellWalk w(blockIdx.x * kT + threadIdx.x, gridDim.x * kT, nx);
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT, w.next()) {
    const uint8_t f = inf[c];
    if (is_act(f)) {
      a += (double)(f & 7) * x[c];
      n += (double)(f & 7);
      int na = 0;
      na += w.i > 0 && is_act(inf[c - 1]);
      na += w.i < nx - 1 && is_act(inf[c + 1]);
      na += w.j > 0 && is_act(inf[c - nx]);
      na += w.j < ny - 1 && is_act(inf[c + nx]);
      d += (double)((int)(f & 7) - na);
    }
  }
  write_part<kT>(a, part_nx, sm);
  write_part<kT>(n, part_n, sm);
  write_part<kT>(d, part_d, sm);
}

// Pressure grid for the velocity update: x at unknowns, 0 elsewhere.  With the
// mean of b removed and no unknown next to AIR (an all-Neumann system, defined up to
// a constant) the constant is fixed as Jacobi-preconditioned CG from x0 = 0 fixes
// it: every Krylov vector of D^-1 K then has zero D-weighted sum, so that solve ends
// with sum(n_i x_i) = 0.  The constant matters: the update treats SOLID neighbours
// as p = 0.  With an AIR neighbour the system is non-singular and x is kept as
// solved (part_d: the AIR-link partials).
__global__ void __launch_bounds__(kT) k_mg_finish(int C, const uint8_t* __restrict__ inf,
                                                   const double* __restrict__ x0,
                                                   const double* __restrict__ x1,
                                                   const CgScalars* sc,
                                                   const double* __restrict__ part_nx,
                                                   const double* __restrict__ part_n,
                                                   const double* __restrict__ part_d, int n_part,
                                                   double* __restrict__ xout,
                                                   float* __restrict__ xg,
                                                   const int* __restrict__ sid,
                                                   const double* __restrict__ scene_norm) {
  __shared__ double sm[4];
  double shift = 0.0;
  if (part_nx && sum_part<kT>(part_d, n_part, sm) == 0.0) {
    const double a = sum_part<kT>(part_nx, n_part, sm);
    const double n = sum_part<kT>(part_n, n_part, sm);
    shift = n > 0.0 ? a / n : 0.0;
  }
  const double* x = (sc->iter & 1) ? x1 : x0;
  for (int c = blockIdx.x * kT + threadIdx.x; c < C; c += gridDim.x * kT) {
    double v = is_act(inf[c]) ? x[c] - shift : 0.0;
    if (sid) v = sid[c] >= 0 ? v * scene_norm[sid[c]] : 0.0;
    xout[c] = v;  // may be x[c] itself: each thread reads and writes only its cells
    xg[c] = (float)v;
  }
}

// Level-0 tiles (32 x 32 fine cells and their 16 x 16 coarse cells) with no unknown on
// either level: the level-0 tile kernels skip them.  One block per tile.
// Level 0 (count != null) also counts its live tiles (cleared by k_mg_setup1).
__global__ void __launch_bounds__(kT) k_mg_tile_flags(int nx, int ny,
                                                      const uint8_t* __restrict__ inf0, int cnx,
                                                      int cny, const uint8_t* __restrict__ inf1,
                                                      uint8_t* __restrict__ flag,
                                                      int* __restrict__ count) {
  const int tiles_x = (nx + 31) >> 5;
  const int bx = blockIdx.x % tiles_x, by = blockIdx.x / tiles_x;
  int any = 0;
  for (int t = threadIdx.x; t < 32 * 32; t += kT) {
    const int i = bx * 32 + (t & 31), j = by * 32 + (t >> 5);
    if (i < nx && j < ny && is_act(inf0[i + j * nx])) any = 1;
  }
  for (int t = threadIdx.x; t < 16 * 16; t += kT) {
    const int I = bx * 16 + (t & 15), J = by * 16 + (t >> 4);
    if (I < cnx && J < cny && is_act(inf1[I + J * cnx])) any = 1;
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) {
    flag[blockIdx.x] = any ? 1 : 0;
    if (count && any) atomicAdd(count, 1);
  }
}

// ---------------------------------------------------------------- batch: per-scene CG
// A packed batch runs one CG per scene in lock step: every scene has its own alpha,
// beta and stopping test (its own dot products), so each converges as its own solve
// would, and the iteration ends once every scene meets |r_s| < tol |b'_s|.  The
// vector kernels walk 64 x 16 tiles that each lie in one scene (a block per tile, a
// column of 4 rows per thread) and add their block sums to that scene's
// accumulator, one atomic per block.  Cells outside every scene are never unknowns
// and keep their zeros.
constexpr int kMgbRows = 4;  // rows per thread; tiles are 64 x (4 * kMgbRows)

struct MgbTile {
  int i, j0, s, jend;  // column, first row, scene, end row (exclusive) of the scene
  bool col;            // the column lies in the scene
  bool live;           // the tile holds an unknown (tiles without one skip their work)
};

// The tile's three table entries are independent loads (one memory round trip); the
// kernels read them before their stop-flag test so that the two waits overlap.
__device__ __forceinline__ MgbTile mgb_tile(const MgbArgs& a) {
  const int4 t = a.tile[blockIdx.x];
  const int2 e = a.tend[blockIdx.x];
  const uint8_t vf = a.vflag ? a.vflag[blockIdx.x] : 1;
  MgbTile m;
  m.s = t.z;
  m.i = t.x + (threadIdx.x & 63);
  m.j0 = t.y + (threadIdx.x >> 6);
  m.jend = e.y;
  m.col = m.i < e.x;
  m.live = vf != 0;
  return m;
}

// Vector tiles with no unknown (flags for the batch kernels).
__global__ void __launch_bounds__(kT) k_mgb_flags(MgbArgs a, const uint8_t* __restrict__ inf,
                                                  uint8_t* __restrict__ vflag) {
  const int4 t4 = a.tile[blockIdx.x];
  const int4 r = a.rect[t4.z];
  const int i = t4.x + (threadIdx.x & 63);
  int any = 0;
  if (i < r.x + r.z)
    for (int k = 0, j = t4.y + (threadIdx.x >> 6); k < kMgbRows && j < r.y + r.w; ++k, j += 4)
      if (is_act(inf[i + j * a.W])) any = 1;
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) vflag[blockIdx.x] = any ? 1 : 0;
}

__device__ __forceinline__ void mgb_add(double v, double* dst, double* sm) {
  double t[1] = {v};
  block_sum<1, kT>(t, sm);
  if (threadIdx.x == 0 && t[0] != 0.0) unsafeAtomicAdd(dst, t[0]);
}

// |b'_s|^2 (1 after the per-scene normalisation, 0 for a scene without unknowns) and
// |r_s|^2 of x0 = 0.
__global__ void __launch_bounds__(kT) k_mgb_init(MgbArgs a, const double* __restrict__ norm) {
  for (int s = threadIdx.x; s < a.n_scenes; s += kT) {
    const double b = norm[s] > 0.0 ? 1.0 : 0.0;
    a.bb[s] = b;
    a.rr[s] = b;
  }
}

// Stop test of iteration `it` on the scenes' |r_s|^2, then the accumulators of this
// iteration cleared.  One workgroup.
__global__ void __launch_bounds__(kT) k_mgb_check(MgbArgs a, int it, int max_iter, double tol2,
                                                  CgScalars* sc) {
  __shared__ int s_done;
  __shared__ double s_w[kT / 64];
  if (threadIdx.x == 0) s_done = ld_dev(&sc->done);
  __syncthreads();
  if (s_done) return;
  const int S = a.n_scenes;
  int ok = 1;
  double worst = 0.0;
  for (int s = threadIdx.x; s < S; s += kT) {
    const double bb = a.bb[s], rr = a.rr[s];
    if (bb > 0.0) {
      double thr = tol2 * bb;
      if (thr < DBL_MIN) thr = DBL_MIN;
      if (!(rr < thr)) ok = 0;
      worst = fmax(worst, rr / bb);
    }
    a.rr[s] = 0.0;
    a.pq[s] = 0.0;
    a.rz[(it & 1) * S + s] = 0.0;
  }
  const int all_ok = __syncthreads_and(ok);
  for (int off = 32; off > 0; off >>= 1) worst = fmax(worst, __shfl_down(worst, off, 64));
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = worst;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kT / 64; ++w) worst = fmax(worst, s_w[w]);
    sc->iter = it;
    sc->err = sqrt(worst);
    if (all_ok || it >= max_iter) st_dev(&sc->done, 1);
  }
}

// The rows k = 0 .. kMgbRows-1 of a thread's column in its tile: cell index (0 for a
// row outside the scene, in[k] = false) and info byte.  Loads first, then use.
struct MgbCells {
  int c[kMgbRows];
  bool in[kMgbRows];
  uint8_t f[kMgbRows];
};

__device__ __forceinline__ MgbCells mgb_cells(const MgbTile& t, int W,
                                              const uint8_t* __restrict__ inf) {
  MgbCells m;
#pragma unroll
  for (int k = 0; k < kMgbRows; ++k) {
    const int j = t.j0 + 4 * k;
    m.in[k] = t.col && j < t.jend;
    m.c[k] = m.in[k] ? t.i + j * W : 0;
  }
#pragma unroll
  for (int k = 0; k < kMgbRows; ++k) {
    const uint8_t f = inf[m.c[k]];
    m.f[k] = m.in[k] ? f : 0;
  }
  return m;
}

// r_s . z_s.  A block takes kRzTiles consecutive tiles (the list runs scene by scene,
// so they mostly share one): their r and z loads are all in flight before the first
// use, one block sum covers them, and each run of tiles of one scene adds one atomic.
constexpr int kRzTiles = 4;

__global__ void __launch_bounds__(kT) k_mgb_rz(MgbArgs a, const float* __restrict__ r,
                                               const float* __restrict__ z, int it,
                                               const CgScalars* sc) {
  __shared__ double sm[kRzTiles * (kT / 64)];
  const int tb = blockIdx.x * kRzTiles;
  // flags, table entries and the stop flag in one memory round trip (the batch always
  // has its tile flags)
  const int done = sc->done;
  int4 t4[kRzTiles];
  int2 e[kRzTiles];
  int vf[kRzTiles];
#pragma unroll
  for (int k = 0; k < kRzTiles; ++k) {
    const int ti = tb + k < a.n_tiles ? tb + k : a.n_tiles - 1;
    t4[k] = a.tile[ti];
    e[k] = a.tend[ti];
    vf[k] = a.vflag[ti];
  }
  // r = z = 0 on a tile without unknowns.  (x0 >= 0 always holds; the test makes the
  // early return below wait for the table entries, so that they are loaded before the
  // branch together with the flags rather than after it.)
  bool live[kRzTiles];
  int any = 0;
#pragma unroll
  for (int k = 0; k < kRzTiles; ++k) {
    live[k] = (tb + k < a.n_tiles) & (vf[k] != 0) & (t4[k].x >= 0);
    any |= (int)live[k];
  }
  if (done | !any) return;
  // r = z = 0 at every cell that is not an unknown, so the info bytes are not needed
  float rv[kRzTiles][kMgbRows], zv[kRzTiles][kMgbRows];
  bool in[kRzTiles][kMgbRows];
#pragma unroll
  for (int k = 0; k < kRzTiles; ++k) {
    const int i = t4[k].x + (int)(threadIdx.x & 63);
#pragma unroll
    for (int q = 0; q < kMgbRows; ++q) {
      const int j = t4[k].y + (int)(threadIdx.x >> 6) + 4 * q;
      in[k][q] = live[k] && i < e[k].x && j < e[k].y;
      const int c = in[k][q] ? i + j * a.W : 0;
      rv[k][q] = r[c];
      zv[k][q] = z[c];
    }
  }
  double v[kRzTiles];
#pragma unroll
  for (int k = 0; k < kRzTiles; ++k) {
    v[k] = 0.0;
#pragma unroll
    for (int q = 0; q < kMgbRows; ++q)
      if (in[k][q]) v[k] += (double)rv[k][q] * (double)zv[k][q];
  }
  block_sum<kRzTiles, kT>(v, sm);
  if (threadIdx.x != 0) return;
  double* rz = a.rz + (it & 1) * a.n_scenes;
  double acc = 0.0;
  int s = t4[0].z;
#pragma unroll
  for (int k = 0; k < kRzTiles; ++k) {
    if (tb + k >= a.n_tiles) break;
    if (t4[k].z != s) {
      if (acc != 0.0) unsafeAtomicAdd(&rz[s], acc);
      s = t4[k].z;
      acc = 0.0;
    }
    acc += v[k];
  }
  if (acc != 0.0) unsafeAtomicAdd(&rz[s], acc);
}

// p_s = z + beta_s p_old (beta_s = r.z / previous r.z of the scene), K p, p_s . K p_s
__global__ void __launch_bounds__(kT) k_mgb_dir(MgbArgs a, const uint8_t* __restrict__ inf,
                                                const float* __restrict__ z,
                                                const float* __restrict__ p_old,
                                                float* __restrict__ p_new, int it,
                                                const CgScalars* sc) {
  // p on the tile and a 1-cell halo, staged in LDS: every z and p_old value is read from
  // memory once per tile, the stencil reads LDS
  constexpr int LW = 64 + 2, LH = 4 * kMgbRows + 2, NL = (LW * LH + kT - 1) / kT;
  __shared__ double sm[kT / 64];
  __shared__ float s_p[LW * LH];
  const MgbTile t = mgb_tile(a);
  if (sc->done) return;
  // a tile without unknowns writes p = 0 into both ping-pong buffers (iterations 0, 1)
  if (!t.live && it >= 2) return;
  const int S = a.n_scenes, W = a.W;
  const int x0 = t.i - (int)(threadIdx.x & 63), y0 = t.j0 - (int)(threadIdx.x >> 6);
  const bool first = it == 0;
  const MgbCells m = mgb_cells(t, W, inf);
  float zv[NL], pv[NL];
#pragma unroll
  for (int e = 0; e < NL; ++e) {
    // halo cells outside the packed grid read a clamped cell; no unknown uses them
    const int q = threadIdx.x + e * kT;
    const int g = (x0 - 1 + q % LW) + (y0 - 1 + q / LW) * W;
    const int gc = g < 0 ? 0 : (g >= a.C ? a.C - 1 : g);
    zv[e] = z[gc];
    pv[e] = first ? 0.f : p_old[gc];
  }
  const double rzp = first ? 0.0 : a.rz[((it - 1) & 1) * S + t.s];
  const double rzc = a.rz[(it & 1) * S + t.s];
  const float bf = rzp != 0.0 ? (float)(rzc / rzp) : 0.f;
#pragma unroll
  for (int e = 0; e < NL; ++e) {
    const int q = threadIdx.x + e * kT;
    if (q < LW * LH) s_p[q] = first ? zv[e] : zv[e] + bf * pv[e];
  }
  __syncthreads();
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kMgbRows; ++k) {
    if (!m.in[k]) break;
    float pc = 0.f;
    if (is_act(m.f[k])) {
      // neighbours of an unknown lie in its scene (its border is SOLID: z = p = 0 there)
      const int l = (int)(threadIdx.x & 63) + 1 + ((int)(threadIdx.x >> 6) + 4 * k + 1) * LW;
      pc = s_p[l];
      const float sn = s_p[l - 1] + s_p[l + 1] + s_p[l - LW] + s_p[l + LW];
      const float kp = nns_of(m.f[k]) * pc - sn;
      acc += (double)pc * (double)kp;
    }
    p_new[m.c[k]] = pc;
  }
  mgb_add(acc, &a.pq[t.s], sm);
}

// x_new = x_old + alpha_s p, r = b' - K x_new (FP64), |r_s|^2
__global__ void __launch_bounds__(kT) k_mgb_xr(MgbArgs a, const uint8_t* __restrict__ inf,
                                               const float* __restrict__ p,
                                               const double* __restrict__ x_old,
                                               double* __restrict__ x_new,
                                               const double* __restrict__ b,
                                               float* __restrict__ r, int it, CgScalars* sc) {
  // x_new on the tile and a 1-cell halo, staged in LDS: every x_old and p value is read
  // from memory once per tile, the stencil reads LDS
  constexpr int LW = 64 + 2, LH = 4 * kMgbRows + 2, NL = (LW * LH + kT - 1) / kT;
  __shared__ double sm[kT / 64];
  __shared__ double s_x[LW * LH];
  const MgbTile t = mgb_tile(a);
  if (sc->done) return;
  // a tile without unknowns: x[0] = r = 0 from the setup, x[1] = 0 from iteration 0
  if (!t.live && it >= 1) return;
  const int W = a.W;
  const int x0 = t.i - (int)(threadIdx.x & 63), y0 = t.j0 - (int)(threadIdx.x >> 6);
  const MgbCells m = mgb_cells(t, W, inf);
  double bv[kMgbRows], xo[NL];
  float pv[NL];
#pragma unroll
  for (int k = 0; k < kMgbRows; ++k) bv[k] = b[m.c[k]];
#pragma unroll
  for (int e = 0; e < NL; ++e) {
    // halo cells outside the packed grid read a clamped cell; no unknown uses them
    const int q = threadIdx.x + e * kT;
    const int g = (x0 - 1 + q % LW) + (y0 - 1 + q / LW) * W;
    const int gc = g < 0 ? 0 : (g >= a.C ? a.C - 1 : g);
    xo[e] = x_old[gc];
    pv[e] = p[gc];
  }
  const double pq = a.pq[t.s];
  const double rz = a.rz[(it & 1) * a.n_scenes + t.s];
  const double alpha = pq != 0.0 ? rz / pq : 0.0;
#pragma unroll
  for (int e = 0; e < NL; ++e) {
    const int q = threadIdx.x + e * kT;
    if (q < LW * LH) s_x[q] = xo[e] + alpha * (double)pv[e];
  }
  __syncthreads();
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kMgbRows; ++k) {
    if (!m.in[k]) break;
    const int c = m.c[k];
    const uint8_t f = m.f[k];
    double xn = 0.0;
    float rv = 0.f;
    if (is_act(f)) {
      const int l = (int)(threadIdx.x & 63) + 1 + ((int)(threadIdx.x >> 6) + 4 * k + 1) * LW;
      xn = s_x[l];
      const double sn = s_x[l - 1] + s_x[l + 1] + s_x[l - LW] + s_x[l + LW];
      const double r64 = bv[k] - ((double)(f & 7) * xn - sn);
      rv = (float)r64;
      acc += r64 * r64;
    }
    x_new[c] = xn;
    r[c] = rv;
  }
  mgb_add(acc, &a.rr[t.s], sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->spmv = it + 1;
}

void launch_mgb_init(hipStream_t s, const MgbArgs& a, const double* norm, const MgHost& m) {
  k_mgb_init<<<1, kT, 0, s>>>(a, norm);
  if (a.vflag && a.n_tiles > 0)
    k_mgb_flags<<<a.n_tiles, kT, 0, s>>>(a, m.inf[0], const_cast<uint8_t*>(a.vflag));
}

void launch_mgb_check(hipStream_t s, const MgbArgs& a, int it, int max_iter, double tol,
                      CgScalars* sc) {
  k_mgb_check<<<1, kT, 0, s>>>(a, it, max_iter, tol * tol, sc);
}

// a.vflag is set with the tile table when a batch is packed (k_mgb_rz reads it as given)
void launch_mgb_vectors(hipStream_t s, const MgbArgs& a, const MgHost& m, int it, CgScalars* sc,
                        int which) {
  if (a.n_tiles <= 0) return;
  if (which == 0)
    k_mgb_rz<<<(a.n_tiles + kRzTiles - 1) / kRzTiles, kT, 0, s>>>(a, m.rl[0], m.xo[0], it,
                                                                   sc);
  else if (which == 1)
    k_mgb_dir<<<a.n_tiles, kT, 0, s>>>(a, m.inf[0], m.xo[0], m.p[it & 1], m.p[(it + 1) & 1], it,
                                        sc);
  else
    k_mgb_xr<<<a.n_tiles, kT, 0, s>>>(a, m.inf[0], m.p[(it + 1) & 1], m.x[it & 1],
                                       m.x[(it + 1) & 1], m.b, m.rl[0], it, sc);
}

// ---------------------------------------------------------------- launchers
static inline int vec_blocks(int C) {
  int nb = (C + 4 * kT - 1) / (4 * kT);
  if (nb > kMgMaxVecBlocks) nb = kMgMaxVecBlocks;
  return nb < 1 ? 1 : nb;
}

int mg_vec_blocks(int C) { return vec_blocks(C); }
int mg_up_blocks(int nx, int ny) { return ((nx + 31) / 32) * ((ny + 31) / 32); }

hipError_t mg_set_lds_limit(size_t bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_coarse),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  // the last tile level's down kernel hosts the same stage (its own static LDS aside)
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_down<false, kTileBig, false, true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  for (const void* f : {reinterpret_cast<const void*>(&k_mg_down<false, kTileBig, false, true, 3>),
                        reinterpret_cast<const void*>(&k_mg_down<false, kTileBig, false, true, 4>),
                        reinterpret_cast<const void*>(&k_mg_down<false, kTileBig, false, true, 5>),
                        reinterpret_cast<const void*>(&k_mg_down<false, kTileBig, false, true, 6>)}) {
    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
  }
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_scene_coarse),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

void launch_mg_scene_setup(hipStream_t s, const SceneCoarseArgs& a) {
  k_mg_scene_setup<<<a.n_scenes, kCoarseThreads, a.stride, s>>>(a);
}

void launch_mg_scene_coarse(hipStream_t s, const SceneCoarseArgs& a, CgScalars* sc) {
  k_mg_scene_coarse<<<a.n_scenes, kCoarseThreads, (size_t)a.stride * (3 * sizeof(float) + 1), s>>>(
      a, sc);
}

void launch_mg_setup(hipStream_t s, const Geom& g, const MgHost& m, const uint8_t* label,
                     const float* uf, const float* vf, int project_mean, int warm,
                     CgScalars* sc, const MgBatchScale* batch, const double* adj_rhs,
                     float* rhs, int* n_fluid) {
  const int C = g.nx * g.ny, nb = vec_blocks(C);
  // levels coarsened one launch each, down to the first level l0 whose finer level
  // (>= 1: level 0's info comes from k_mg_setup0) fits the one-workgroup tail
  int l0 = 2;
  while (l0 < m.nlev && m.nx[l0 - 1] * m.ny[l0 - 1] > kTailCells) ++l0;
  if (m.nlev < 2) l0 = m.nlev;
  // one domain's forward solve: level 0's set-up by tiles, which also forms the tile
  // flags and the labels of levels 1 .. lk (k_mg_setup0t), then the info of the levels
  // and the coarsening tail in one launch (k_mg_level_info)
  const bool tiled = !adj_rhs && !batch && rhs && m.tflag && m.nlev >= 3 && l0 < m.nlev;
  const int n_tiles = mg_up_blocks(g.nx, g.ny);
  // a cold start without the mean projection: k_mg_setup0t does k_mg_setup1's per-cell
  // work (its tile partials in the vector kernels' nb slots) and k_mg_level_info the
  // counts; k_mg_setup1 does not run
  const bool folded = tiled && !project_mean && !warm && n_tiles <= nb;
  const int lk = tiled ? std::min(std::min(5, l0 - 1), m.nlev - 1) : 0;
  int n_part = nb;
  if (adj_rhs) {
    // part_bb / part_rr are free until k_mg_setup1 writes them
    k_mg_adj_rhs0<<<nb, kT, 0, s>>>(g, label, adj_rhs, m.inf[0], m.part_bb, m.part_rr,
                                    m.part_pq);
    k_mg_adj_rhs1<<<nb, kT, 0, s>>>(C, -(double)g.dx * (double)g.dx, project_mean, m.inf[0],
                                    adj_rhs, m.part_bb, m.part_rr, m.part_pq, nb, m.b, m.part_sum,
                                    m.part_cnt);
  } else if (tiled) {
    MgTileLabels L{};
    L.nl = lk;
    for (int l = 1; l <= lk; ++l) {
      L.nx[l] = m.nx[l];
      L.ny[l] = m.ny[l];
      L.lab[l] = m.lab[l];
    }
    MgSetupFold fold;
    if (folded) fold = MgSetupFold{m.x[0], m.rl[0], m.part_bb, m.part_rr, nb, sc};
    k_mg_setup0t<<<n_tiles, kT, 0, s>>>(g, label, uf, vf, m.inf[0], m.b, m.part_sum, m.part_cnt,
                                         rhs, m.part_liq, m.tflag, L, fold);
    n_part = n_tiles;
  } else {
    k_mg_setup0<<<nb, kT, 0, s>>>(g, label, uf, vf, m.inf[0], m.b, m.part_sum, m.part_cnt, rhs,
                                  rhs ? m.part_liq : nullptr);
  }
  if (batch) {
    k_mg_scene_norm<<<batch->n_scenes, 1024, 0, s>>>(m.b, g.nx, batch->rect, batch->norm,
                                                     batch->inv);
    k_mg_scene_scale<<<nb, kT, 0, s>>>(C, batch->sid, batch->inv, m.b);
  }
  // x[0] holds the previous solve's solution (k_mg_finish); x[1] is free until the
  // first update of this solve
  if (warm) k_mg_warm<<<nb, kT, 0, s>>>(C, m.inf[0], m.x[0], m.x[1]);
  // an adjoint right-hand side is not mean-projected: S^T y already has zero sum for a
  // singular system, and a non-singular one solves A y = x-bar itself (P follows the
  // solve)
  if (!folded)
    k_mg_setup1<<<nb, kT, 0, s>>>(g.nx, g.ny, adj_rhs ? 0 : project_mean, m.inf[0], m.b,
                                  warm ? m.x[1] : nullptr,
                                  m.x[0], m.rl[0], m.part_sum, m.part_cnt, n_part, m.part_bb,
                                  m.part_rr, sc, rhs && !adj_rhs ? m.part_liq : nullptr, n_fluid,
                                  tiled ? m.tflag : nullptr, n_tiles);
  // the levels not formed by k_mg_setup0t, one launch each, down to level l0 - 1
  for (int l = std::max(1, lk + 1); l < std::min(l0, m.nlev); ++l) {
    const int nf = m.nx[l - 1] * m.ny[l - 1], ncl = m.nx[l] * m.ny[l];
    const int n = nf > ncl ? nf : ncl;
    int blocks = (n + kT - 1) / kT;
    if (blocks > 2048) blocks = 2048;
    k_mg_coarsen<<<blocks, kT, 0, s>>>(m.nx[l - 1], m.ny[l - 1], m.lab[l - 1],
                                       l - 1 >= 1 ? m.inf[l - 1] : nullptr, m.nx[l], m.ny[l],
                                       m.lab[l]);
  }
  if (tiled) {
    // the info of levels 1 .. lk - 1 (lk .. l0 - 2 came with the launches above) and the
    // tail from l0 in one launch
    MgInfoArgs ia{};
    ia.nli = lk - 1;
    int cells = 0;
    for (int l = 1; l <= ia.nli; ++l) cells += m.nx[l] * m.ny[l];
    for (int l = 0; l < m.nlev; ++l) {
      ia.nx[l] = m.nx[l];
      ia.ny[l] = m.ny[l];
      ia.lab[l] = m.lab[l];
      ia.inf[l] = m.inf[l];
    }
    MgTailArgs a{};
    a.l0 = l0;
    a.nlev = m.nlev;
    for (int l = 0; l < m.nlev; ++l) {
      a.nx[l] = m.nx[l];
      a.ny[l] = m.ny[l];
      a.lab[l] = m.lab[l];
      a.inf[l] = m.inf[l];
    }
    // the level-0 tile flags came with k_mg_setup0t: the map is formed here too
    if (m.tmap && !batch) {
      ia.nb_map = mg_up_blocks(m.nx[0], m.ny[0]);
      ia.tflag = m.tflag;
      ia.tmap = m.tmap;
    }
    if (folded) {
      ia.part_liq = m.part_liq;
      ia.n_part = n_tiles;
      ia.n_fluid = n_fluid;
      ia.tflag0 = m.tflag;
      ia.n_tiles = n_tiles;
      ia.sc = sc;
    }
    const int blocks = std::min(std::max((cells + 1023) / 1024, 1), 1024) + (ia.tmap ? 2 : 1);
    k_mg_level_info<<<blocks, 1024, 0, s>>>(ia, a);
  } else if (m.nlev > 1 && l0 < m.nlev) {
    MgTailArgs a{};
    a.l0 = l0;
    a.nlev = m.nlev;
    for (int l = 0; l < m.nlev; ++l) {
      a.nx[l] = m.nx[l];
      a.ny[l] = m.ny[l];
      a.lab[l] = m.lab[l];
      a.inf[l] = m.inf[l];
    }
    k_mg_coarsen_tail<<<1, 1024, 0, s>>>(a);
  } else if (m.nlev > 1) {
    const int l = m.nlev - 1, n = m.nx[l] * m.ny[l];
    // info of the last level: a coarsening call with no coarse output
    k_mg_coarsen<<<(n + kT - 1) / kT, kT, 0, s>>>(m.nx[l], m.ny[l], m.lab[l], m.inf[l], 0, 0,
                                                  nullptr);
  }
  for (int l = tiled ? 1 : 0; l < m.lc; ++l) {
    uint8_t* flag = l == 0 ? m.tflag : m.tflag_l[l];
    if (flag)
      k_mg_tile_flags<<<mg_up_blocks(m.nx[l], m.ny[l]), kT, 0, s>>>(
          m.nx[l], m.ny[l], m.inf[l], m.nx[l + 1], m.ny[l + 1], m.inf[l + 1], flag,
          l == 0 ? &sc->live_tiles : nullptr);
  }
  if (m.tmap && m.tflag && !batch && m.lc > 0 && !tiled)
    k_mg_tile_map<<<1, 1024, 0, s>>>(mg_up_blocks(m.nx[0], m.ny[0]), m.tflag, m.tmap);
}

static int mg_down_blocks(const MgHost& m, int l) {
  return ((m.nx[l + 1] + 15) / 16) * ((m.ny[l + 1] + 15) / 16);
}

MgCheck mg_check_args(const MgHost& m, int it, int max_iter, double tol, CgScalars* sc) {
  const int nb = vec_blocks(m.nx[0] * m.ny[0]);
  // |r|^2: k_mg_setup1's partials (it == 0), else k_mg_xr's or, fused, level 0's down
  // kernel's (one per tile); |b'|^2 from k_mg_setup1
  const int n_rr = (m.fused && it > 0) ? mg_down_blocks(m, 0) : nb;
  return MgCheck{sc, m.part_rr, n_rr, m.part_bb, nb, tol * tol, max_iter, m.per_scene ? 1 : 0};
}

int mg_part_rr_size(const MgHost& m) {
  return std::max(kMgMaxVecBlocks, m.nlev > 1 ? mg_down_blocks(m, 0) : 1);
}

static MgLevel mg_level(const MgHost& m, int l) {
  return MgLevel{m.nx[l], m.ny[l], m.inf[l], m.rl[l], m.xs[l], m.xo[l]};
}

// The tile levels' weights for S > 2 sweeps: the reciprocal roots of the degree-S
// Chebyshev polynomial on [lo_S, 2] (the high end of the spectrum of D^-1 K), smallest
// first (every partial product of the factors stays within [-1, 1] on [0, 2]).  lo_S
// per degree from a model of the solve over developed dam-break domains (fewest
// iterations): 0.2, 0.2, 0.03, 0.05 for S = 3 .. 6 (S = 5: 48 iterations over six
// developed 1024^2 dam-break solves against 50 with 0.05).
static MgSweeps cheb_sweeps(int S) {
  MgSweeps w{};
  if (S <= 2 || S > kMgMaxSweeps) return w;
  static const double kLo[kMgMaxSweeps + 1] = {0, 0, 0.4, 0.2, 0.2, 0.03, 0.05};
  const double lo = kLo[S], hi = 2.0, pi = 3.14159265358979323846;
  double r[kMgMaxSweeps];
  for (int k = 0; k < S; ++k)
    r[k] = 1.0 / ((hi + lo) / 2 + (hi - lo) / 2 * std::cos(pi * (2 * k + 1) / (2.0 * S)));
  std::sort(r, r + S);
  for (int k = 0; k < S; ++k) w.w[k] = (float)r[k];
  return w;
}
static MgSweeps mg_sweeps(const MgHost& m) { return cheb_sweeps(m.level_sweeps); }
// level 0: three sweeps take the degree-3 weights on [0.2, 2]
static MgSweeps mg_sweeps0(const MgHost& m) { return cheb_sweeps(m.level0_sweeps); }

// One CG iteration = V-cycle (z = B r) + direction + update.  Level 0's down kernel
// (or the one-workgroup kernel when the hierarchy starts there) runs the stop test
// of this iteration first.
int mg_down_blocks0(const MgHost& m) { return m.nlev > 1 ? mg_down_blocks(m, 0) : 0; }

void launch_mg_down(hipStream_t s, MgHost& m, int l, int it, int max_iter, double tol,
                    CgScalars* sc) {
  const int blocks = mg_down_blocks(m, l);
  unsigned long long* span = nullptr;
  if (l == 0 && m.clk_span_on && m.clk_span && m.clk_span_n < m.clk_span_cap &&
      m.clk_span_nb == blocks)
    span = m.clk_span + 2 * (size_t)blocks * m.clk_span_n++;
  const MgCheck chk = mg_check_args(m, it, max_iter, tol, sc);
  // the stop test: level 0's kernel, or level 1's in the fused form (level 0 forms
  // the residual it tests)
  const int check = l == (m.fused ? 1 : 0) ? 1 : 0;
  // level 0: m.l0_threads per tile (256 by default: many tiles); smaller levels: 1024
  // threads, so that each LDS phase of the few tiles takes one or two passes
  if (l == 0) {
    // fused form: iteration it - 1's update, the buffers k_mg_xr would use then
    const MgXr xr = m.fused ? MgXr{m.x[(it + 1) & 1], m.x[it & 1], m.p[it & 1], m.b, m.part_pq,
                                   mg_dir_blocks(m), m.part_rr}
                            : MgXr{};
    const int ck = m.fused ? 0 : check;
    const MgSweeps sw0 = mg_sweeps0(m);
    auto go = [&](auto nt, auto ns) {
      constexpr int NT = decltype(nt)::value, S = decltype(ns)::value;
      if (m.fused)
        k_mg_down<true, NT, true, false, S><<<blocks, NT, 0, s>>>(
            mg_level(m, 0), mg_level(m, 1), m.coarse.omega, m.coarse.omega2, it, chk, ck, m.tflag,
            xr, span, MgCoarseArgs{}, nullptr, sw0, m.tmap);
      else
        k_mg_down<true, NT, false, false, S><<<blocks, NT, 0, s>>>(
            mg_level(m, 0), mg_level(m, 1), m.coarse.omega, m.coarse.omega2, it, chk, ck, m.tflag,
            xr, span, MgCoarseArgs{}, nullptr, sw0, m.tmap);
    };
    auto go_s = [&](auto nt) {
      if (m.level0_sweeps == 3)
        go(nt, std::integral_constant<int, 3>{});
      else
        go(nt, std::integral_constant<int, 2>{});
    };
    if (m.l0_threads == 1024)
      go_s(std::integral_constant<int, 1024>{});
    else if (m.l0_threads == 512)
      go_s(std::integral_constant<int, 512>{});
    else
      go_s(std::integral_constant<int, kT>{});
  } else {
    // the last tile level also runs the one-workgroup stage (mg_tail_coarse): its own
    // instance, so that the other levels (and every level of a batch) carry none of it;
    // m.level_sweeps (2 or 4) sweeps before the coarse correction
    const MgSweeps sw = mg_sweeps(m);
    auto go = [&](auto ns) {
      constexpr int S = decltype(ns)::value;
      if (l == m.lc - 1 && mg_tail_coarse(m))
        k_mg_down<false, kTileBig, false, true, S><<<blocks, kTileBig, m.coarse_lds, s>>>(
            mg_level(m, l), mg_level(m, l + 1), m.coarse.omega, m.coarse.omega2, it, chk, check,
            m.tflag_l[l], MgXr{}, nullptr, m.coarse, m.ticket, sw, nullptr);
      else
        k_mg_down<false, kTileBig, false, false, S><<<blocks, kTileBig, 0, s>>>(
            mg_level(m, l), mg_level(m, l + 1), m.coarse.omega, m.coarse.omega2, it, chk, check,
            m.tflag_l[l], MgXr{}, nullptr, MgCoarseArgs{}, nullptr, sw, nullptr);
    };
    switch (m.level_sweeps) {
      case 3: go(std::integral_constant<int, 3>{}); break;
      case 4: go(std::integral_constant<int, 4>{}); break;
      case 5: go(std::integral_constant<int, 5>{}); break;
      case 6: go(std::integral_constant<int, 6>{}); break;
      default: go(std::integral_constant<int, 2>{}); break;
    }
  }
}

bool mg_tail_coarse(const MgHost& m) {
  return !m.per_scene && m.lc >= 2 && m.ticket != nullptr && m.coarse_tail;
}

void launch_mg_coarse(hipStream_t s, const MgHost& m, int it, int max_iter, double tol,
                      CgScalars* sc) {
  const MgCheck chk = mg_check_args(m, it, max_iter, tol, sc);
  k_mg_coarse<<<1, kCoarseThreads, m.coarse_lds, s>>>(m.coarse, chk, it, m.lc == 0 ? 1 : 0,
                                                      m.lc == 0 ? m.part_rz : nullptr,
                                                      m.clk_on ? m.clk : nullptr);
}

void launch_mg_up(hipStream_t s, const MgHost& m, int l, int it, CgScalars* sc) {
  const int blocks = mg_up_blocks(m.nx[l], m.ny[l]);
  auto go0 = [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    if (m.level0_sweeps == 3)
      k_mg_up<true, NT, 3><<<blocks, NT, 0, s>>>(mg_level(m, 0), mg_level(m, 1), m.coarse.omega,
                                                 m.coarse.omega2, sc, m.part_rz, it, m.tflag,
                                                 mg_sweeps0(m), m.tmap);
    else
      k_mg_up<true, NT><<<blocks, NT, 0, s>>>(mg_level(m, 0), mg_level(m, 1), m.coarse.omega,
                                              m.coarse.omega2, sc, m.part_rz, it, m.tflag,
                                              MgSweeps{}, m.tmap);
  };
  if (l == 0 && m.l0_threads == 1024)
    go0(std::integral_constant<int, 1024>{});
  else if (l == 0 && m.l0_threads == 512)
    go0(std::integral_constant<int, 512>{});
  else if (l == 0)
    go0(std::integral_constant<int, kT>{});
  else {
    auto go = [&](auto ns) {
      constexpr int S = decltype(ns)::value;
      k_mg_up<false, kTileBig, S><<<blocks, kTileBig, 0, s>>>(mg_level(m, l), mg_level(m, l + 1),
                                                              m.coarse.omega, m.coarse.omega2, sc,
                                                              nullptr, it, m.tflag_l[l],
                                                              mg_sweeps(m), nullptr);
    };
    switch (m.level_sweeps) {
      case 3: go(std::integral_constant<int, 3>{}); break;
      case 4: go(std::integral_constant<int, 4>{}); break;
      case 5: go(std::integral_constant<int, 5>{}); break;
      case 6: go(std::integral_constant<int, 6>{}); break;
      default: go(std::integral_constant<int, 2>{}); break;
    }
  }
}

int mg_dir_blocks(const MgHost& m) {
  return m.tflag ? mg_up_blocks(m.nx[0], m.ny[0]) : vec_blocks(m.nx[0] * m.ny[0]);
}

void launch_mg_dir(hipStream_t s, const MgHost& m, int it, CgScalars* sc) {
  const int n_rz = m.lc == 0 ? 1 : mg_up_blocks(m.nx[0], m.ny[0]);
  if (m.tflag) {
    k_mg_dir_tiles<<<mg_dir_blocks(m), kT, 0, s>>>(m.nx[0], m.ny[0], m.inf[0], m.xo[0],
                                                   m.p[it & 1], m.p[(it + 1) & 1], it, m.part_rz,
                                                   n_rz, m.part_pq, sc, m.tflag, m.tmap);
    return;
  }
  const int nb = vec_blocks(m.nx[0] * m.ny[0]);
  k_mg_dir<<<nb, kT, 0, s>>>(m.nx[0], m.ny[0], m.inf[0], m.xo[0], m.p[it & 1], m.p[(it + 1) & 1],
                             it, m.part_rz, n_rz, m.part_pq, sc);
}

void launch_mg_xr(hipStream_t s, const MgHost& m, int it, CgScalars* sc) {
  const int nb = vec_blocks(m.nx[0] * m.ny[0]);
  k_mg_xr<<<nb, kT, 0, s>>>(m.nx[0], m.ny[0], m.inf[0], m.p[(it + 1) & 1], m.x[it & 1],
                            m.x[(it + 1) & 1], m.b, m.rl[0], it, m.part_pq, mg_dir_blocks(m),
                            m.part_rr, sc);
}

void launch_mg_remove_mean(hipStream_t s, const MgHost& m) {
  const int C = m.nx[0] * m.ny[0], nb = vec_blocks(C);
  k_mg_mean_part<<<nb, kT, 0, s>>>(C, m.inf[0], m.x[0], m.part_sum, m.part_cnt);
  k_mg_mean_sub<<<nb, kT, 0, s>>>(C, m.inf[0], m.x[0], m.part_sum, m.part_cnt, nb);
}

__global__ void __launch_bounds__(64) k_read_scalars(const CgScalars* __restrict__ sc,
                                                    const int* __restrict__ n_fluid,
                                                    HostScalars* __restrict__ out) {
  constexpr int kW = (int)(sizeof(CgScalars) / 4);
  static_assert(sizeof(CgScalars) % 4 == 0 && kW < 64, "one word per lane");
  const int t = threadIdx.x;
  if (t < kW) reinterpret_cast<int*>(&out->sc)[t] = reinterpret_cast<const int*>(sc)[t];
  if (t == 63) out->n_fluid = n_fluid ? *n_fluid : -1;
  __threadfence_system();
}
void launch_read_scalars(hipStream_t s, const CgScalars* sc, const int* n_fluid,
                         HostScalars* out) {
  k_read_scalars<<<1, 64, 0, s>>>(sc, n_fluid, out);
}

void launch_mg_finish(hipStream_t s, const MgHost& m, int project_mean, CgScalars* sc,
                      float* xg, const MgBatchScale* batch) {
  const int C = m.nx[0] * m.ny[0], nb = vec_blocks(C);
  const int* sid = batch ? batch->sid : nullptr;
  const double* norm = batch ? batch->norm : nullptr;
  if (project_mean) {
    // part_sum / part_cnt are free again after the setup, part_bb after iteration 0
    k_mg_wsum<<<nb, kT, 0, s>>>(m.nx[0], m.ny[0], m.inf[0], m.x[0], m.x[1], sc, m.part_sum,
                                m.part_cnt, m.part_bb);
    k_mg_finish<<<nb, kT, 0, s>>>(C, m.inf[0], m.x[0], m.x[1], sc, m.part_sum, m.part_cnt,
                                  m.part_bb, nb, m.x[0], xg, sid, norm);
  } else {
    k_mg_finish<<<nb, kT, 0, s>>>(C, m.inf[0], m.x[0], m.x[1], sc, nullptr, nullptr, nullptr, 0,
                                  m.x[0], xg, sid, norm);
  }
}

}  // namespace fsg
