// kernels_cg.hip -- pressure Poisson system setup and the Jacobi-preconditioned CG
// solvers that replace Eigen::ConjugateGradient (include/FluidSolver.h:114,
// src/FluidSolver.cpp:81,362-425).
//
// Unknowns are the LIQUID cells in row-major order, exactly the reference's
// fluid_indices (src/FluidSolver.cpp:331-346), built by a three-kernel scan.  The
// operator is applied matrix-free: one info byte per unknown holds the four
// LIQUID-neighbour bits and the count of non-SOLID neighbours, so A's row is
// diag_tab[count] at the centre and off = 1/dx^2 at each LIQUID neighbour.  Left
// and right neighbours of unknown k are k-1 and k+1 (row-major order); the up and
// down neighbours' indices are stored per unknown.
//
// Synchronisation on gfx950 (measured): a contended returning atomic costs ~90 ns
// per block and agent-scope release/acquire fences write back and invalidate L2.
// So no kernel here uses a last-block ticket or an agent-scope fence:
//  - the plain form lets every block re-sum the previous kernel's per-block
//    partials in a fixed order (kernel boundaries order the data);
//  - the persistent forms use a fence-free grid barrier: data that crosses blocks
//    (halo values, partial sums) is written with agent-scope atomic stores and read
//    with agent-scope atomic loads, both of which go to the device coherence point,
//    and the arrival counter takes one non-returning atomic add per block.
// Every sum is deterministic (fixed per-thread order, fixed wave/block tree, fixed
// block order), so every block of a persistent launch derives identical scalars.
#include <float.h>

#include "fsg_common.h"
#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

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
  block_sum<1, 256>(v, sm);
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

// Compact indices of the (j-1) and (j+1) neighbours of every unknown (-1 where the
// neighbour is not LIQUID).  Runs after k_compact has written fluid_index.
__global__ void __launch_bounds__(256) k_updown(int nx, const int* __restrict__ n_fluid,
                                                const int* __restrict__ cell_of,
                                                const uint8_t* __restrict__ info,
                                                const int* __restrict__ fluid_index,
                                                int2* __restrict__ updn) {
  const int F = *n_fluid;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) {
    const int c = cell_of[k];
    const uint8_t inf = info[k];
    updn[k] = make_int2((inf & 4) ? fluid_index[c - nx] : -1, (inf & 8) ? fluid_index[c + nx] : -1);
  }
}

// ---------------------------------------------------------------- plain form
// Eigen's conjugate_gradient restated (x0 = 0; stop when |r|^2 < max(tol^2 |b|^2,
// min normal) or after max_iter iterations), one iteration = 3 kernels:
//   spmv:    Ap = A p, partial p.Ap
//   update:  alpha = r.z / p.Ap; x += alpha p, r -= alpha Ap; partial |r|^2, r.z
//   pupdate: stop test, beta = r.z_new / r.z; p = z + beta p
// Each kernel's blocks re-sum the previous kernel's partials in block order;
// block 0 records the scalars for the host.  Scalars are formed in float like
// Eigen's RealScalar; sums are reduced in FP64.

constexpr int kPlainThreads = 256;

template <int NV>
__device__ __forceinline__ void sum_partials(const double* part, int nb, double (&v)[NV],
                                             double* sm) {
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kPlainThreads)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += part[b * NV + q];
  block_sum<NV, kPlainThreads>(v, sm);
}

template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* part, double* sm) {
  block_sum<NV, kPlainThreads>(v, sm);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) part[blockIdx.x * NV + q] = v[q];
}

__global__ void __launch_bounds__(256) k_cg_sum_b(int F, const float* __restrict__ b,
                                                  double* __restrict__ part) {
  __shared__ double sm[4];
  double v[1] = {0.0};
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) v[0] += b[k];
  write_partials<1>(v, part, sm);
}

__global__ void __launch_bounds__(256) k_cg_init(int F, CgParams cp, const int* __restrict__ cell_of,
                                                 const uint8_t* __restrict__ info,
                                                 const float* __restrict__ b, float* __restrict__ r,
                                                 float* __restrict__ pg, float* __restrict__ xg,
                                                 const double* __restrict__ mean_part,
                                                 double* __restrict__ part, CgScalars* sc) {
  __shared__ double sm[2 * 4];
  float mean = 0.f;
  if (cp.project_mean) {
    double m[1];
    sum_partials<1>(mean_part, gridDim.x, m, sm);
    mean = (float)m[0] / (float)F;
  }
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
  write_partials<2>(v, part, sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->done = 0;
    sc->iter = 0;
    sc->spmv = 0;
    sc->err = 0.0;
  }
}

// part0: (|r|^2, r.z) of the previous update (or init, it == 0); part1: p.Ap out
__global__ void __launch_bounds__(256) k_cg_spmv(int F, int nx, int it, CgParams cp,
                                                 const int* __restrict__ cell_of,
                                                 const uint8_t* __restrict__ info,
                                                 const float* __restrict__ pg, float* __restrict__ ap,
                                                 const double* __restrict__ part_rz,
                                                 double* __restrict__ part_pap, CgScalars* sc) {
  __shared__ double sm[2 * 4];
  if (it == 0) {
    // Eigen's entry checks: |b|^2 == 0 -> x = 0; |r0|^2 < threshold -> done
    double v[2];
    sum_partials<2>(part_rz, gridDim.x, v, sm);
    const float rhs = (float)v[0];
    float thr = cp.tol * cp.tol * rhs;
    if (thr < FLT_MIN) thr = FLT_MIN;
    const bool stop = rhs == 0.f || rhs < thr;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->rhs_norm2 = rhs;
      sc->threshold = thr;
      sc->rz[0] = (float)v[1];
      sc->rr = rhs;
      sc->err = rhs == 0.f ? 0.0 : (stop ? 1.0 : 0.0);
      if (stop) sc->done = 1;
    }
    if (stop) return;
  } else if (sc->done) {
    return;
  }
  double v[1] = {0.0};
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) {
    const int c = cell_of[k];
    const uint8_t inf = info[k];
    const float pc = pg[c];
    float acc = sel5((inf >> 4) & 7, cp.diag_tab) * pc;
    if (inf & 1) acc += cp.off * pg[c - 1];
    if (inf & 2) acc += cp.off * pg[c + 1];
    if (inf & 4) acc += cp.off * pg[c - nx];
    if (inf & 8) acc += cp.off * pg[c + nx];
    ap[k] = acc;
    v[0] += (double)(pc * acc);
  }
  write_partials<1>(v, part_pap, sm);
}

__global__ void __launch_bounds__(256) k_cg_update(int F, int it, CgParams cp,
                                                   const int* __restrict__ cell_of,
                                                   const uint8_t* __restrict__ info,
                                                   float* __restrict__ r, const float* __restrict__ ap,
                                                   const float* __restrict__ pg, float* __restrict__ xg,
                                                   const double* __restrict__ part_pap,
                                                   double* __restrict__ part_rz, CgScalars* sc) {
  __shared__ double sm[2 * 4];
  if (sc->done) return;
  double pap[1];
  sum_partials<1>(part_pap, gridDim.x, pap, sm);
  const float alpha = sc->rz[it & 1] / (float)pap[0];
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
  write_partials<2>(v, part_rz, sm);
}

__global__ void __launch_bounds__(256) k_cg_pupdate(int F, int it, CgParams cp,
                                                    const int* __restrict__ cell_of,
                                                    const uint8_t* __restrict__ info,
                                                    const float* __restrict__ r, float* __restrict__ pg,
                                                    const double* __restrict__ part_rz, CgScalars* sc) {
  __shared__ double sm[2 * 4];
  if (sc->done) return;
  double v[2];
  sum_partials<2>(part_rz, gridDim.x, v, sm);
  const float rr = (float)v[0];
  const bool conv = rr < (float)sc->threshold;
  const float abs_old = sc->rz[it & 1];
  const float abs_new = (float)v[1];
  const float beta = abs_new / abs_old;
  const bool last = !conv && it + 1 >= cp.max_iter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rr = rr;
    sc->err = sqrt((double)rr / sc->rhs_norm2);
    sc->spmv = it + 1;
    if (conv) {
      sc->iter = it;
      sc->done = 1;
    } else {
      sc->iter = it + 1;
      sc->rz[(it + 1) & 1] = abs_new;
      sc->beta = beta;
      if (last) sc->done = 1;
    }
  }
  if (conv || last) return;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < F; k += gridDim.x * 256) {
    const int c = cell_of[k];
    const float z = sel5((info[k] >> 4) & 7, cp.invd_tab) * r[k];
    pg[c] = z + beta * pg[c];
  }
}

static inline int cg_blocks(int F) {
  int nb = (F + 1023) / 1024;  // ~4 unknowns per thread
  if (nb < 1) nb = 1;
  if (nb > 1024) nb = 1024;
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
void launch_updown(hipStream_t s, int nx, int F_bound, const int* n_fluid, const int* cell_of,
                   const uint8_t* info, const int* fluid_index, int2* updn) {
  int nb = (F_bound + 255) / 256;
  if (nb > 4096) nb = 4096;
  if (nb < 1) nb = 1;
  k_updown<<<nb, 256, 0, s>>>(nx, n_fluid, cell_of, info, fluid_index, updn);
}
void launch_cg_init(hipStream_t s, int F, const CgParams& cp, const int* cell_of,
                    const uint8_t* info, const float* b, float* r, float* pg, float* xg,
                    double* partials, CgScalars* sc) {
  const int nb = cg_blocks(F);
  // partials layout: [0, 4096) mean, [4096, 8192) r.z sets (2 x 2 values), [12288,...) p.Ap
  double* mean_part = partials;
  if (cp.project_mean) k_cg_sum_b<<<nb, 256, 0, s>>>(F, b, mean_part);
  k_cg_init<<<nb, 256, 0, s>>>(F, cp, cell_of, info, b, r, pg, xg, mean_part, partials + 4096, sc);
}
void launch_cg_iteration(hipStream_t s, int F, int nx, int it, const CgParams& cp,
                         const int* cell_of, const uint8_t* info, float* r, float* ap, float* pg,
                         float* xg, double* partials, CgScalars* sc) {
  const int nb = cg_blocks(F);
  double* rz_prev = partials + 4096 + (it & 1) * 2048;
  double* rz_next = partials + 4096 + ((it + 1) & 1) * 2048;
  double* pap = partials + 8192;
  k_cg_spmv<<<nb, 256, 0, s>>>(F, nx, it, cp, cell_of, info, pg, ap, rz_prev, pap, sc);
  k_cg_update<<<nb, 256, 0, s>>>(F, it, cp, cell_of, info, r, ap, pg, xg, pap, rz_next, sc);
  k_cg_pupdate<<<nb, 256, 0, s>>>(F, it, cp, cell_of, info, r, pg, rz_next, sc);
}

}  // namespace fsg
