// kernels_cg.hip -- pressure Poisson system setup and the Jacobi-preconditioned CG
// solvers that replace Eigen::ConjugateGradient (include/FluidSolver.h:114,
// src/FluidSolver.cpp:81,362-425).
//
// Unknowns are the LIQUID cells in row-major order, exactly the reference's
// fluid_indices (src/FluidSolver.cpp:331-346), built by a three-kernel scan.  The
// operator is applied matrix-free: one info byte per unknown holds the four
// LIQUID-neighbour bits and the count of non-SOLID neighbours, so A's row is
// diag_tab[count] at the centre and off = 1/dx^2 at each LIQUID neighbour.
// Vectors that neighbours read (the search direction p and z = M^-1 r) live in
// grid layout (nx*ny floats, indexed by cell); everything else is compact.
//
// Reductions are deterministic: per-thread partial sums, a fixed wave/block tree,
// per-block partials in FP64, and a fixed-order final sum.
#include <float.h>

#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

// ------------------------------------------------------------------ helpers
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sm) {
  // sm: NV * 4 doubles (256 threads = 4 waves of 64)
#pragma unroll
  for (int q = 0; q < NV; ++q)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) sm[q * 4 + wave] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = ((sm[q * 4 + 0] + sm[q * 4 + 1]) + sm[q * 4 + 2]) + sm[q * 4 + 3];
  __syncthreads();
}

// Block partials + last-block-done total.  Returns true in the last block, where
// thread 0 holds the totals in v (summed over blocks in block order).
template <int NV>
__device__ __forceinline__ bool last_block_total(double (&v)[NV], double* partials,
                                                 unsigned* counter) {
  __shared__ double sm[NV * 4];
  __shared__ int is_last;
  block_sum<NV>(v, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[blockIdx.x * NV + q] = v[q];
    __threadfence();
    const unsigned t = atomicAdd(counter, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < NV; ++q)
      acc[q] += __hip_atomic_load(&partials[b * NV + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(acc, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = acc[q];
    *counter = 0;
  }
  return true;
}

__device__ __forceinline__ float sel5(int n, const float (&t)[5]) {
  float r = t[0];
  r = n == 1 ? t[1] : r;
  r = n == 2 ? t[2] : r;
  r = n == 3 ? t[3] : r;
  r = n == 4 ? t[4] : r;
  return r;
}

// Row of the Laplacian at cell c: diag * p[c] first, then the LIQUID neighbours in
// the order (i-1), (i+1), (j-1), (j+1) (src/FluidSolver.cpp:379-410 insert order).
__device__ __forceinline__ float apply_row(const float* __restrict__ pg, int c, int nx,
                                           uint8_t inf, const CgParams& cp) {
  float acc = sel5((inf >> 4) & 7, cp.diag_tab) * pg[c];
  if (inf & 1) acc += cp.off * pg[c - 1];
  if (inf & 2) acc += cp.off * pg[c + 1];
  if (inf & 4) acc += cp.off * pg[c - nx];
  if (inf & 8) acc += cp.off * pg[c + nx];
  return acc;
}

// ------------------------------------------------------------ system setup
// Per 2048-cell chunk: number of LIQUID cells (8 cells per thread, 8-byte loads).
__global__ void __launch_bounds__(256) k_count_liquid(int n, const uint8_t* __restrict__ label,
                                                      int* __restrict__ counts) {
  const int base = blockIdx.x * kScanChunk + threadIdx.x * 8;
  int cnt = 0;
  if (base + 8 <= n) {
    const uint2 w = *reinterpret_cast<const uint2*>(label + base);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cnt += ((w.x >> (8 * q)) & 0xFF) == kLiquid;
      cnt += ((w.y >> (8 * q)) & 0xFF) == kLiquid;
    }
  } else {
    for (int q = 0; q < 8; ++q)
      if (base + q < n) cnt += label[base + q] == kLiquid;
  }
  double v[1] = {(double)cnt};
  __shared__ double sm[4];
  block_sum<1>(v, sm);
  if (threadIdx.x == 0) counts[blockIdx.x] = (int)v[0];
}

// Exclusive scan of the chunk counts in one block of 1024 threads.
__global__ void __launch_bounds__(1024) k_scan_counts(int nb, const int* __restrict__ counts,
                                                      int* __restrict__ offsets,
                                                      int* __restrict__ n_fluid) {
  __shared__ int sm[1024];
  const int per = (nb + 1023) / 1024;
  const int b0 = threadIdx.x * per;
  int s = 0;
  for (int q = 0; q < per; ++q)
    if (b0 + q < nb) s += counts[b0 + q];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
    __syncthreads();
    sm[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sm[threadIdx.x] - s;  // exclusive prefix of this thread's first chunk
  for (int q = 0; q < per; ++q)
    if (b0 + q < nb) {
      offsets[b0 + q] = run;
      run += counts[b0 + q];
    }
  if (threadIdx.x == 1023) *n_fluid = sm[1023];
}

// fluid_indices (src/FluidSolver.cpp:331-346), the compact cell list, the
// divergence right-hand side b = divVelX + divVelY (:413, MacGrid.h:98-111) and the
// neighbour info byte (:379-410).
__global__ void __launch_bounds__(256) k_compact(Geom g, const uint8_t* __restrict__ label,
                                                 const int* __restrict__ offsets,
                                                 int* __restrict__ fluid_index,
                                                 int* __restrict__ cell_of,
                                                 const float* __restrict__ uf,
                                                 const float* __restrict__ vf,
                                                 float* __restrict__ b,
                                                 uint8_t* __restrict__ info) {
  const int n = g.nx * g.ny, nx = g.nx;
  const int base = blockIdx.x * kScanChunk + threadIdx.x * 8;
  uint8_t lab[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) lab[q] = (base + q < n) ? label[base + q] : kAir;
  int cnt = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) cnt += lab[q] == kLiquid;
  // exclusive scan over the block's 256 threads
  __shared__ int sm[256];
  sm[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
    __syncthreads();
    sm[threadIdx.x] += v;
    __syncthreads();
  }
  int k = offsets[blockIdx.x] + sm[threadIdx.x] - cnt;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int c = base + q;
    if (c >= n) break;
    if (lab[q] != kLiquid) {
      fluid_index[c] = -1;
      continue;
    }
    fluid_index[c] = k;
    cell_of[k] = c;
    const int i = c % nx, j = c / nx;
    // out-of-range neighbours count as SOLID (a SOLID border makes this moot)
    const uint8_t lxm = i > 0 ? label[c - 1] : kSolid;
    const uint8_t lxp = i < nx - 1 ? label[c + 1] : kSolid;
    const uint8_t lym = j > 0 ? label[c - nx] : kSolid;
    const uint8_t lyp = j < g.ny - 1 ? label[c + nx] : kSolid;
    const int bits = (lxm == kLiquid) | ((lxp == kLiquid) << 1) | ((lym == kLiquid) << 2) |
                     ((lyp == kLiquid) << 3);
    const int nns = (lxm != kSolid) + (lxp != kSolid) + (lym != kSolid) + (lyp != kSolid);
    info[k] = (uint8_t)(bits | (nns << 4));
    const float uxp = i < nx - 1 ? uf[c + 1] : 0.f;
    const float vyp = j < g.ny - 1 ? vf[c + nx] : 0.f;
    const float dvx = div_x(uxp - uf[c], g);
    const float dvy = div_y(vyp - vf[c], g);
    b[k] = dvx + dvy;
    ++k;
  }
}

// ---------------------------------------------------------------- plain form
// Eigen's conjugate_gradient restated (x0 = 0; stop when |r|^2 < max(tol^2 |b|^2,
// min normal) or after max_iter iterations), one iteration = 3 kernels:
//   spmv:    Ap = A p, p.Ap                       -> alpha = r.z / p.Ap
//   update:  x += alpha p, r -= alpha Ap, |r|^2, z = r / diag, r.z -> stop / beta
//   pupdate: p = z + beta p
// Scalars are formed in float like Eigen's RealScalar; sums are reduced in FP64.

__global__ void __launch_bounds__(256) k_cg_mean(int F, const float* __restrict__ b,
                                                 double* partials, CgScalars* sc) {
  double v[1] = {0.0};
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) v[0] += b[k];
  if (last_block_total<1>(v, partials, &sc->counter) && threadIdx.x == 0)
    sc->mean = (double)((float)v[0] / (float)F);
}

__global__ void __launch_bounds__(256) k_cg_init(int F, CgParams cp, const int* __restrict__ cell_of,
                                                 const uint8_t* __restrict__ info,
                                                 const float* __restrict__ b, float* __restrict__ r,
                                                 float* __restrict__ pg, float* __restrict__ xg,
                                                 double* partials, CgScalars* sc) {
  const float mean = cp.project_mean ? (float)sc->mean : 0.f;
  double v[2] = {0.0, 0.0};
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) {
    const int c = cell_of[k];
    const float rk = cp.project_mean ? b[k] - mean : b[k];
    const float z = sel5((info[k] >> 4) & 7, cp.invd_tab) * rk;
    r[k] = rk;
    pg[c] = z;
    xg[c] = 0.f;
    v[0] += (double)(rk * rk);
    v[1] += (double)(rk * z);
  }
  if (last_block_total<2>(v, partials, &sc->counter) && threadIdx.x == 0) {
    const float rhs = (float)v[0];
    sc->rhs_norm2 = rhs;
    sc->iter = 0;
    sc->spmv = 0;
    sc->err = 0.0;
    if (rhs == 0.f) {
      sc->done = 1;
      return;
    }
    float thr = cp.tol * cp.tol * rhs;
    if (thr < FLT_MIN) thr = FLT_MIN;
    sc->threshold = thr;
    sc->rr = rhs;
    sc->abs_new = (float)v[1];
    sc->done = 0;
    if (rhs < thr) {
      sc->done = 1;
      sc->err = 1.0;
    }
  }
}

__global__ void __launch_bounds__(256) k_cg_spmv(int F, int nx, CgParams cp,
                                                 const int* __restrict__ cell_of,
                                                 const uint8_t* __restrict__ info,
                                                 const float* __restrict__ pg, float* __restrict__ ap,
                                                 double* partials, CgScalars* sc) {
  if (sc->done) return;
  double v[1] = {0.0};
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) {
    const int c = cell_of[k];
    const float a = apply_row(pg, c, nx, info[k], cp);
    ap[k] = a;
    v[0] += (double)(pg[c] * a);
  }
  if (last_block_total<1>(v, partials, &sc->counter) && threadIdx.x == 0) {
    sc->pap = (float)v[0];
    sc->alpha = (float)sc->abs_new / (float)v[0];
    sc->spmv += 1;
  }
}

__global__ void __launch_bounds__(256) k_cg_update(int F, CgParams cp, const int* __restrict__ cell_of,
                                                   const uint8_t* __restrict__ info,
                                                   float* __restrict__ r, const float* __restrict__ ap,
                                                   const float* __restrict__ pg, float* __restrict__ xg,
                                                   double* partials, CgScalars* sc) {
  if (sc->done) return;
  const float alpha = sc->alpha;
  double v[2] = {0.0, 0.0};
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) {
    const int c = cell_of[k];
    xg[c] = xg[c] + alpha * pg[c];
    const float rk = r[k] - alpha * ap[k];
    r[k] = rk;
    const float z = sel5((info[k] >> 4) & 7, cp.invd_tab) * rk;
    v[0] += (double)(rk * rk);
    v[1] += (double)(rk * z);
  }
  if (last_block_total<2>(v, partials, &sc->counter) && threadIdx.x == 0) {
    const float rr = (float)v[0];
    sc->rr = rr;
    sc->err = sqrt((double)rr / sc->rhs_norm2);
    if (rr < (float)sc->threshold) {
      sc->done = 1;
      return;
    }
    const float abs_old = (float)sc->abs_new;
    const float abs_new = (float)v[1];
    sc->abs_new = abs_new;
    sc->beta = abs_new / abs_old;
    sc->iter += 1;
    if (sc->iter >= cp.max_iter) sc->done = 1;
  }
}

__global__ void __launch_bounds__(256) k_cg_pupdate(int F, CgParams cp, const int* __restrict__ cell_of,
                                                    const uint8_t* __restrict__ info,
                                                    const float* __restrict__ r, float* __restrict__ pg,
                                                    const CgScalars* sc) {
  if (sc->done) return;
  const float beta = sc->beta;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) {
    const int c = cell_of[k];
    const float z = sel5((info[k] >> 4) & 7, cp.invd_tab) * r[k];
    pg[c] = z + beta * pg[c];
  }
}

static inline int cg_blocks(int F) {
  int nb = (F + 1023) / 1024;  // ~4 unknowns per thread
  if (nb < 1) nb = 1;
  if (nb > 2048) nb = 2048;
  return nb;
}

void launch_count_liquid(hipStream_t s, const Geom& g, const uint8_t* label, int* counts) {
  const int n = g.nx * g.ny;
  k_count_liquid<<<(n + kScanChunk - 1) / kScanChunk, 256, 0, s>>>(n, label, counts);
}
void launch_scan_counts(hipStream_t s, int nb, const int* counts, int* offsets, int* n_fluid) {
  k_scan_counts<<<1, 1024, 0, s>>>(nb, counts, offsets, n_fluid);
}
void launch_compact(hipStream_t s, const Geom& g, const uint8_t* label, const int* offsets,
                    int* fluid_index, int* cell_of, const float* uf, const float* vf, float* b,
                    uint8_t* info) {
  const int n = g.nx * g.ny;
  k_compact<<<(n + kScanChunk - 1) / kScanChunk, 256, 0, s>>>(g, label, offsets, fluid_index,
                                                               cell_of, uf, vf, b, info);
}
void launch_cg_mean(hipStream_t s, int F, const float* b, double* partials, CgScalars* sc) {
  k_cg_mean<<<cg_blocks(F), 256, 0, s>>>(F, b, partials, sc);
}
void launch_cg_init(hipStream_t s, int F, const CgParams& cp, const int* cell_of,
                    const uint8_t* info, const float* b, float* r, float* pg, float* xg,
                    double* partials, CgScalars* sc) {
  k_cg_init<<<cg_blocks(F), 256, 0, s>>>(F, cp, cell_of, info, b, r, pg, xg, partials, sc);
}
void launch_cg_iteration(hipStream_t s, int F, int nx, const CgParams& cp, const int* cell_of,
                         const uint8_t* info, float* r, float* ap, float* pg, float* xg,
                         double* partials, CgScalars* sc) {
  const int nb = cg_blocks(F);
  k_cg_spmv<<<nb, 256, 0, s>>>(F, nx, cp, cell_of, info, pg, ap, partials, sc);
  k_cg_update<<<nb, 256, 0, s>>>(F, cp, cell_of, info, r, ap, pg, xg, partials, sc);
  k_cg_pupdate<<<nb, 256, 0, s>>>(F, cp, cell_of, info, r, pg, sc);
}

// ----------------------------------------------------------- persistent form
// All iterations of one solve in one cooperative launch.  Each block owns a
// contiguous range of unknowns.  One iteration = two grid-wide reduce-barriers:
//   A: p_k = z_k + beta p_{k-1} for own cells and (recomputed) for neighbours,
//      publish p_k, Ap_k, partial p.Ap                        -> barrier, alpha
//   B: x += alpha p, r -= alpha Ap, z = r/diag published, partial |r|^2, r.z
//                                                             -> barrier, beta
// Phase A reads z (published before barrier B) and p_{k-1} (published in the
// previous phase A, before barrier A), and writes p_k into the other ping-pong
// buffer, so no third barrier is needed.
//
// The refined form wraps the FP32 inner solve in FP64 iterative refinement:
// r64 = b - A x64 (FP64 SpMV on x64 published in grid layout), inner solve
// A e = r64 in FP32 to a relative tolerance, x64 += e.

// Barrier + sum of NV doubles over all blocks.  Slots are double-buffered by the
// barrier's parity with a fixed stride of kSlotNV values per block, so the two sets
// never overlap whatever NV the neighbouring barriers use: a block cannot reuse a
// set before every block has left the barrier that read it.  Returns the totals in
// every thread.
constexpr int kSlotNV = 4;
template <int NV>
__device__ __forceinline__ void grid_sum(double (&v)[NV], double* slots, unsigned* bar,
                                         unsigned& epoch, double* sm) {
  static_assert(NV <= kSlotNV, "grid_sum: too many values");
  block_sum<NV>(v, sm);
  const unsigned G = gridDim.x;
  double* set = slots + (size_t)(epoch & 1) * G * kSlotNV;
  epoch += 1;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) set[blockIdx.x * NV + q] = v[q];
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = epoch * G;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
      __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  // every block sums the same slots in the same order
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int b = threadIdx.x; b < (int)G; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < NV; ++q)
      acc[q] += __hip_atomic_load(&set[b * NV + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(acc, sm);
  __shared__ double bc[NV];
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) bc[q] = acc[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = bc[q];
  __syncthreads();
}

// Plain barrier (no data): used where only publication order matters.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned& epoch) {
  __syncthreads();
  epoch += 1;
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = epoch * gridDim.x;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
      __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// One FP32 Jacobi-PCG solve of A e = rhs over this block's unknowns [k0, k1).
// rhs is r (compact, overwritten with the running residual); the solution is
// accumulated in xg (grid layout, zeroed here).  Returns the iteration count in
// Eigen's convention; sets *spmv_out and the final relative residual.
__device__ int pcg_inner(const PersistentArgs& a, int F, int k0, int k1, float tol, int max_iter,
                         unsigned& epoch, double* sm, int* spmv_out, double* err_out,
                         double* rhs_out) {
  const CgParams& cp = a.cp;
  const int nx = a.nx;
  double v2[2] = {0.0, 0.0};
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    const int c = a.cell_of[k];
    const float rk = a.r[k];
    const float z = sel5((a.info[k] >> 4) & 7, cp.invd_tab) * rk;
    a.zg[c] = z;
    a.xg[c] = 0.f;
    v2[0] += (double)(rk * rk);
    v2[1] += (double)(rk * z);
  }
  grid_sum<2>(v2, a.partials, a.bar, epoch, sm);
  const float rhs = (float)v2[0];
  *rhs_out = v2[0];
  *spmv_out = 0;
  if (rhs == 0.f) {
    *err_out = 0.0;
    return 0;
  }
  float thr = tol * tol * rhs;
  if (thr < FLT_MIN) thr = FLT_MIN;
  float rr = rhs;
  if (rr < thr) {
    *err_out = 1.0;
    return 0;
  }
  float abs_new = (float)v2[1];
  float beta = 0.f;
  float* pold = a.pg0;
  float* pnew = a.pg1;
  int it = 0, spmv = 0;
  while (true) {
    // phase A
    double v1[1] = {0.0};
    for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
      const int c = a.cell_of[k];
      const uint8_t inf = a.info[k];
      float pc, acc;
      if (it == 0) {
        pc = a.zg[c];
        acc = sel5((inf >> 4) & 7, cp.diag_tab) * pc;
        if (inf & 1) acc += cp.off * a.zg[c - 1];
        if (inf & 2) acc += cp.off * a.zg[c + 1];
        if (inf & 4) acc += cp.off * a.zg[c - nx];
        if (inf & 8) acc += cp.off * a.zg[c + nx];
      } else {
        pc = a.zg[c] + beta * pold[c];
        acc = sel5((inf >> 4) & 7, cp.diag_tab) * pc;
        if (inf & 1) acc += cp.off * (a.zg[c - 1] + beta * pold[c - 1]);
        if (inf & 2) acc += cp.off * (a.zg[c + 1] + beta * pold[c + 1]);
        if (inf & 4) acc += cp.off * (a.zg[c - nx] + beta * pold[c - nx]);
        if (inf & 8) acc += cp.off * (a.zg[c + nx] + beta * pold[c + nx]);
      }
      pnew[c] = pc;
      a.e32[k] = acc;  // Ap, read back by the same thread in phase B
      v1[0] += (double)(pc * acc);
    }
    grid_sum<1>(v1, a.partials, a.bar, epoch, sm);
    ++spmv;
    const float alpha = abs_new / (float)v1[0];
    // phase B
    v2[0] = 0.0;
    v2[1] = 0.0;
    for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
      const int c = a.cell_of[k];
      const float pc = pnew[c];
      a.xg[c] = a.xg[c] + alpha * pc;
      const float rk = a.r[k] - alpha * a.e32[k];
      a.r[k] = rk;
      const float z = sel5((a.info[k] >> 4) & 7, cp.invd_tab) * rk;
      a.zg[c] = z;
      v2[0] += (double)(rk * rk);
      v2[1] += (double)(rk * z);
    }
    grid_sum<2>(v2, a.partials, a.bar, epoch, sm);
    rr = (float)v2[0];
    if (rr < thr) break;
    const float abs_old = abs_new;
    abs_new = (float)v2[1];
    beta = abs_new / abs_old;
    ++it;
    if (it >= max_iter) break;
    float* t = pold;
    pold = pnew;
    pnew = t;
  }
  *spmv_out = spmv;
  *err_out = sqrt((double)rr / (double)rhs);
  return it;
}

__global__ void __launch_bounds__(256) k_cg_persistent(PersistentArgs a) {
  __shared__ double sm[8];
  const int F = *a.n_fluid;
  CgScalars* sc = a.sc;
  if (F == 0) return;  // uniform across the grid: no barrier is entered
  const int G = gridDim.x;
  const int per = (F + G - 1) / G;
  const int k0 = min(F, (int)blockIdx.x * per);
  const int k1 = min(F, k0 + per);
  unsigned epoch = 0;
  const CgParams& cp = a.cp;
  const int nx = a.nx;

  // optional mean removal of b (fully Neumann systems)
  float mean = 0.f;
  if (cp.project_mean) {
    double v[1] = {0.0};
    for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) v[0] += a.b[k];
    grid_sum<1>(v, a.partials, a.bar, epoch, sm);
    mean = (float)v[0] / (float)F;
  }

  if (!a.refined) {
    for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x)
      a.r[k] = cp.project_mean ? a.b[k] - mean : a.b[k];
    int spmv = 0;
    double err = 0.0, rhs = 0.0;
    const int it = pcg_inner(a, F, k0, k1, cp.tol, cp.max_iter, epoch, sm, &spmv, &err, &rhs);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->iter = it;
      sc->spmv = spmv;
      sc->err = err;
      sc->rhs_norm2 = rhs;
      sc->refine_sweeps = 0;
      sc->total_unknown_iters += (unsigned long long)F * (unsigned long long)spmv;
      sc->total_solves += 1;
    }
    return;
  }

  // ---- FP64 iterative refinement
  double v[1] = {0.0};
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    const double bk = (double)(cp.project_mean ? a.b[k] - mean : a.b[k]);
    a.x64[k] = 0.0;
    a.r[k] = (float)bk;
    v[0] += bk * bk;
  }
  grid_sum<1>(v, a.partials, a.bar, epoch, sm);
  const double b2 = v[0];
  double res2 = b2;
  int total_it = 0, total_spmv = 0, sweeps = 0;
  if (b2 > 0.0) {
    const double thr2 = a.outer_tol * a.outer_tol * b2;
    while (res2 >= thr2 && sweeps < a.max_outer) {
      int spmv = 0;
      double err = 0.0, rhs = 0.0;
      total_it += pcg_inner(a, F, k0, k1, a.inner_tol, cp.max_iter, epoch, sm, &spmv, &err, &rhs);
      total_spmv += spmv;
      ++sweeps;
      // x64 += e ; publish x64 in grid layout via zg/pg0 pair (hi/lo split)
      for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        const int c = a.cell_of[k];
        const double x = a.x64[k] + (double)a.xg[c];
        a.x64[k] = x;
        const float hi = (float)x;
        a.zg[c] = hi;
        a.pg0[c] = (float)(x - (double)hi);
      }
      grid_barrier(a.bar, epoch);
      // r64 = b - A x64 in FP64
      double w[1] = {0.0};
      const double off = (double)cp.off;
      for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        const int c = a.cell_of[k];
        const uint8_t inf = a.info[k];
        auto X = [&](int cc) { return (double)a.zg[cc] + (double)a.pg0[cc]; };
        double ax = (double)sel5((inf >> 4) & 7, cp.diag_tab) * X(c);
        if (inf & 1) ax += off * X(c - 1);
        if (inf & 2) ax += off * X(c + 1);
        if (inf & 4) ax += off * X(c - nx);
        if (inf & 8) ax += off * X(c + nx);
        const double bk = (double)(cp.project_mean ? a.b[k] - mean : a.b[k]);
        const double rk = bk - ax;
        a.r[k] = (float)rk;
        w[0] += rk * rk;
      }
      grid_sum<1>(w, a.partials, a.bar, epoch, sm);
      res2 = w[0];
      total_spmv += 1;
    }
  }
  // final solution into xg (FP32) for the gradient update
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) a.xg[a.cell_of[k]] = (float)a.x64[k];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->iter = total_it;
    sc->spmv = total_spmv;
    sc->err = b2 > 0.0 ? sqrt(res2 / b2) : 0.0;
    sc->rhs_norm2 = b2;
    sc->b64_norm2 = b2;
    sc->res64_norm2 = res2;
    sc->refine_sweeps = sweeps;
    sc->total_unknown_iters += (unsigned long long)F * (unsigned long long)total_spmv;
    sc->total_solves += 1;
  }
}

int persistent_grid_size(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg_persistent, 256, 0) !=
      hipSuccess)
    return 0;
  if (per_cu < 1) return 0;
  return prop.multiProcessorCount * (per_cu > 1 ? 1 : per_cu);
}

hipError_t launch_cg_persistent(hipStream_t s, const PersistentArgs& a, int grid) {
  void* args[] = {(void*)&a};
  return hipLaunchCooperativeKernel((const void*)k_cg_persistent, dim3(grid), dim3(256), args, 0, s);
}

}  // namespace fsg
