// kernels_pcg.hip -- persistent (one launch per solve) Jacobi-preconditioned CG,
// optionally inside FP64 iterative refinement.  Replaces Eigen::ConjugateGradient
// (include/FluidSolver.h:114, src/FluidSolver.cpp:81,424-425).
//
// Launch: G co-resident blocks of 1024 threads (cooperative launch).  Block b owns
// the contiguous range [k0, k1) = [b*per, (b+1)*per) of unknowns (LIQUID cells in
// row-major order); a thread owns up to KPT of them (k = k0 + tid + j*1024) and
// keeps their x, r and Ap in registers.
//
// Neighbour values.  The SpMV of unknown k reads k-1, k+1 and its up/down
// neighbours, all within hw = nx of k in the compact order.  Each block keeps an
// LDS window over [k0 - hw, k1 + hw) of z = M^-1 r and of the search direction p
// (two buffers, by iteration parity).  Own entries are written by the owning
// threads; the two halo stripes are staged from device-coherent exchange arrays
// into which every unknown within hw of its block's ends is published.
//
// Iteration (Eigen's PCG, two grid barriers, each also a reduction):
//   A: stage the halo of z and p_{k-1}; p_k = z_k + beta p_{k-1} for own unknowns
//      and, in the stencil, for neighbours; publish p_k; Ap; partial p.Ap
//      -- barrier --  alpha = r.z / p.Ap
//   B: x += alpha p, r -= alpha Ap, z = M^-1 r (published), partial |r|^2, r.z
//      -- barrier --  stop test, beta
// Phase A reads z (published before barrier B) and p_{k-1} (published in the
// previous phase A, before barrier A) and writes p_k into the other buffer of the
// pair, so no third barrier is needed.
//
// Grid barrier: agent-scope atomic stores for everything that crosses blocks, one
// non-returning atomic add per block on a monotonic counter, relaxed polling; no
// agent-scope fences (they write back and invalidate L2 on gfx950).  The wait is
// bounded: a block that waits too long raises an abort flag, every block leaves at
// its next barrier, and the host reports the failure.
//
// Refined form: FP64 iterative refinement around the FP32 solve -- r64 = b - A x64
// with the SpMV and its reduction in FP64, inner solve of A e = r64 to a relative
// tolerance, x64 += e -- until |b - A x64| < tol |b| in FP64.
#include <float.h>

#include "fsg_common.h"
#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

constexpr int kPcgThreads = 1024;
constexpr unsigned kSpinLimit = 1u << 22;  // seconds of waiting: treated as a fault

struct BarrierState {
  unsigned* count;  // monotonic arrival counter
  unsigned* abort;  // set when a block gives up waiting
  unsigned epoch;
  double* slots;    // 2 parity sets x G blocks x 4 values
};

// Grid barrier carrying up to 4 partial sums per block; returns the grid totals
// (identical in every block).  Returns false if the wait was abandoned, in which
// case the caller leaves the kernel without entering another barrier.
template <int NV>
__device__ __forceinline__ bool grid_reduce(BarrierState& bs, double (&v)[NV], double* sm) {
  static_assert(NV <= 4, "grid_reduce: at most 4 values");
  block_sum<NV, kPcgThreads>(v, sm);
  const unsigned G = gridDim.x;
  double* set = bs.slots + (size_t)(bs.epoch & 1) * G * 4;
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) st_dev(&set[blockIdx.x * 4 + q], v[q]);
  // every wave's device-coherent stores complete before the block arrives
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  bs.epoch += 1;
  __shared__ int ok;
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(bs.count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = bs.epoch * G;
    unsigned spins = 0;
    int good = 1;
    while (ld_dev(bs.count) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kSpinLimit) {
        st_dev(bs.abort, 1u);
        good = 0;
        break;
      }
      if ((spins & 255u) == 0 && ld_dev(bs.abort)) {
        good = 0;
        break;
      }
    }
    ok = good;
  }
  __syncthreads();
  if (!ok) return false;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int b = threadIdx.x; b < (int)G; b += kPcgThreads)
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += ld_dev(&set[b * 4 + q]);
  block_sum<NV, kPcgThreads>(acc, sm);
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = acc[q];
  return true;
}

template <int KPT, bool REFINED>
__global__ void __launch_bounds__(kPcgThreads) k_pcg(PersistentArgs a) {
  extern __shared__ float lds[];
  __shared__ double sm[4 * (kPcgThreads / 64)];
  const int F = *a.n_fluid;
  if (F == 0) return;  // uniform across the grid: no barrier is entered
  const int G = gridDim.x;
  const int per = (F + G - 1) / G;
  const int hw = a.nx;                 // halo width in compact indices
  const int W = per + 2 * hw;          // LDS window length
  const int k0 = min(F, (int)blockIdx.x * per);
  const int k1 = min(F, k0 + per);
  const int wb = k0 - hw;              // window base: slot = k - wb
  const int tid = threadIdx.x;
  float* lz = lds;
  float* lp0 = lds + W;
  float* lp1 = lds + 2 * W;
  float* ez = a.exch;
  float* ep0 = a.exch + a.F_max;
  float* ep1 = a.exch + 2 * (size_t)a.F_max;
  const CgParams& cp = a.cp;

  // unknown j of this thread: k = k0 + tid + j*1024, LDS slot hw + tid + j*1024
  int su[KPT], sd[KPT];  // LDS slots of the up / down neighbours
  unsigned valid = 0, infw = 0, nnsw = 0, exported = 0;
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int k = k0 + tid + j * kPcgThreads;
    su[j] = sd[j] = 0;
    if (k >= k1) continue;
    valid |= 1u << j;
    const uint8_t f = a.info[k];
    const int2 ud = a.updn[k];
    su[j] = ud.x - wb;
    sd[j] = ud.y - wb;
    infw |= (unsigned)(f & 15) << (4 * j);
    nnsw |= (unsigned)((f >> 4) & 7) << (3 * j);
    if (k - k0 < hw || k1 - 1 - k < hw) exported |= 1u << j;
  }
#define FSG_VALID(j) (valid & (1u << (j)))
#define FSG_K(j) (k0 + tid + (j) * kPcgThreads)
#define FSG_S(j) (hw + tid + (j) * kPcgThreads)
  BarrierState bs{a.bar, a.bar + 1, 0u, a.partials};
  CgScalars* sc = a.sc;

  // optional mean removal of b (fully Neumann systems)
  float mean = 0.f;
  if (cp.project_mean) {
    double v[1] = {0.0};
#pragma unroll
    for (int j = 0; j < KPT; ++j)
      if (FSG_VALID(j)) v[0] += (double)a.b[FSG_K(j)];
    if (!grid_reduce<1>(bs, v, sm)) return;
    mean = (float)v[0] / (float)F;
  }

  float x[KPT], r[KPT], ap[KPT];
  int total_it = 0, total_spmv = 0, sweeps = 0;
  double err = 0.0, b2 = 0.0, res2 = 0.0, thr2 = 0.0;
  // Start: x0 = 0 (Eigen's solve()), or with warm_start the previous solution kept
  // in xg (zero at cells that were not LIQUID).  r0 = b - A x0 is formed in FP64
  // from xg, which nothing writes during the solve.
  {
    const double off = (double)cp.off;
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      r[j] = 0.f;
      if (!FSG_VALID(j)) continue;
      const int k = FSG_K(j);
      const double bk = (double)(cp.project_mean ? a.b[k] - mean : a.b[k]);
      double rk = bk;
      double x0 = 0.0;
      if (a.warm_start) {
        const int c = a.cell_of[k];
        const unsigned f = (infw >> (4 * j)) & 15u;
        x0 = (double)a.xg[c];
        double ax = (double)sel5((nnsw >> (3 * j)) & 7, cp.diag_tab) * x0;
        if (f & 1u) ax += off * (double)a.xg[c - 1];
        if (f & 2u) ax += off * (double)a.xg[c + 1];
        if (f & 4u) ax += off * (double)a.xg[c - a.nx];
        if (f & 8u) ax += off * (double)a.xg[c + a.nx];
        rk = bk - ax;
      }
      if (REFINED) st_dev(a.x64 + k, x0);
      r[j] = (float)rk;
      v[0] += bk * bk;
      v[1] += rk * rk;
    }
    if (!grid_reduce<2>(bs, v, sm)) return;
    b2 = v[0];
    res2 = v[1];
    thr2 = a.outer_tol * a.outer_tol * b2;
  }

  const float tol = REFINED ? a.inner_tol : cp.tol;
  while (true) {
    if (REFINED && !(b2 > 0.0 && res2 >= thr2 && sweeps < a.max_outer)) break;
    // ------------- one FP32 Jacobi-PCG solve of A e = r (x0 = 0) -------------
    double v2[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      x[j] = 0.f;
      if (!FSG_VALID(j)) continue;
      const float z = sel5((nnsw >> (3 * j)) & 7, cp.invd_tab) * r[j];
      lz[FSG_S(j)] = z;
      if (exported & (1u << j)) st_dev(ez + FSG_K(j), z);
      v2[0] += (double)(r[j] * r[j]);
      v2[1] += (double)(r[j] * z);
    }
    if (!grid_reduce<2>(bs, v2, sm)) return;
    const float rhs = (float)v2[0];
    int it = 0, spmv = 0;
    float rr = rhs;
    // Eigen's test |r| < tol |b|; the refined form's inner solves are relative to
    // their own right-hand side (the current residual)
    const float ref = REFINED ? rhs : (float)b2;
    float thr = tol * tol * ref;
    if (thr < FLT_MIN) thr = FLT_MIN;
    if (ref != 0.f && rhs != 0.f && !(rr < thr)) {
      float abs_new = (float)v2[1];
      float beta = 0.f;
      int par = 0;  // p_k goes to buffer par, p_{k-1} is read from par ^ 1
      while (true) {
        const bool first = it == 0;
        float* lp_new = par ? lp1 : lp0;
        const float* lp_old = par ? lp0 : lp1;
        float* ep_new = par ? ep1 : ep0;
        const float* ep_old = par ? ep0 : ep1;
        // stage the halo stripes [k0 - hw, k0) and [k1, k1 + hw)
        for (int t = tid; t < 2 * hw; t += kPcgThreads) {
          const int q = t < hw ? k0 - hw + t : k1 + (t - hw);
          if (q < 0 || q >= F) continue;
          lz[q - wb] = ld_dev(ez + q);
          if (!first) const_cast<float*>(lp_old)[q - wb] = ld_dev(ep_old + q);
        }
        __syncthreads();
        // phase A: p_k (own, published) and Ap with neighbours' p_k recomputed
        double v1[1] = {0.0};
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
          ap[j] = 0.f;
          if (!FSG_VALID(j)) continue;
          const int s = FSG_S(j);
          const float pk = first ? lz[s] : lz[s] + beta * lp_old[s];
          const unsigned f = (infw >> (4 * j)) & 15u;
          float acc = sel5((nnsw >> (3 * j)) & 7, cp.diag_tab) * pk;
          if (f & 1u) acc += cp.off * (first ? lz[s - 1] : lz[s - 1] + beta * lp_old[s - 1]);
          if (f & 2u) acc += cp.off * (first ? lz[s + 1] : lz[s + 1] + beta * lp_old[s + 1]);
          if (f & 4u) {
            const int t2 = su[j];
            acc += cp.off * (first ? lz[t2] : lz[t2] + beta * lp_old[t2]);
          }
          if (f & 8u) {
            const int t2 = sd[j];
            acc += cp.off * (first ? lz[t2] : lz[t2] + beta * lp_old[t2]);
          }
          ap[j] = acc;
          v1[0] += (double)(pk * acc);
          lp_new[s] = pk;
          if (exported & (1u << j)) st_dev(ep_new + FSG_K(j), pk);
        }
        if (!grid_reduce<1>(bs, v1, sm)) return;
        ++spmv;
        const float alpha = abs_new / (float)v1[0];
        // phase B
        v2[0] = 0.0;
        v2[1] = 0.0;
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
          if (!FSG_VALID(j)) continue;
          const int s = FSG_S(j);
          x[j] = x[j] + alpha * lp_new[s];
          r[j] = r[j] - alpha * ap[j];
          const float z = sel5((nnsw >> (3 * j)) & 7, cp.invd_tab) * r[j];
          lz[s] = z;
          if (exported & (1u << j)) st_dev(ez + FSG_K(j), z);
          v2[0] += (double)(r[j] * r[j]);
          v2[1] += (double)(r[j] * z);
        }
        if (!grid_reduce<2>(bs, v2, sm)) return;
        rr = (float)v2[0];
        if (rr < thr) break;
        const float abs_old = abs_new;
        abs_new = (float)v2[1];
        beta = abs_new / abs_old;
        ++it;
        if (it >= cp.max_iter) break;
        par ^= 1;
      }
    }
    total_it += it;
    total_spmv += spmv;
    err = ref != 0.f ? sqrt((double)rr / (double)ref) : 0.0;
    if (!REFINED) break;
    // ------------- FP64 refinement: x64 += e; r = b - A x64 -------------
    ++sweeps;
#pragma unroll
    for (int j = 0; j < KPT; ++j)
      if (FSG_VALID(j)) {
        const int k = FSG_K(j);
        st_dev(a.x64 + k, ld_dev(a.x64 + k) + (double)x[j]);
      }
    double w[1] = {0.0};
    if (!grid_reduce<1>(bs, w, sm)) return;
    const double off = (double)cp.off;
    w[0] = 0.0;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      if (!FSG_VALID(j)) {
        r[j] = 0.f;
        continue;
      }
      const int k = FSG_K(j);
      const unsigned f = (infw >> (4 * j)) & 15u;
      double ax = (double)sel5((nnsw >> (3 * j)) & 7, cp.diag_tab) * ld_dev(a.x64 + k);
      if (f & 1u) ax += off * ld_dev(a.x64 + k - 1);
      if (f & 2u) ax += off * ld_dev(a.x64 + k + 1);
      if (f & 4u) ax += off * ld_dev(a.x64 + su[j] + wb);
      if (f & 8u) ax += off * ld_dev(a.x64 + sd[j] + wb);
      const double bk = (double)(cp.project_mean ? a.b[k] - mean : a.b[k]);
      const double rk = bk - ax;
      r[j] = (float)rk;
      w[0] += rk * rk;
    }
    if (!grid_reduce<1>(bs, w, sm)) return;
    res2 = w[0];
    total_spmv += 1;
  }
  // solution into grid layout for the gradient update
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    if (!FSG_VALID(j)) continue;
    const int k = FSG_K(j);
    const int c = a.cell_of[k];
    float xk;
    if (b2 == 0.0) xk = 0.f;  // Eigen: |b| = 0 gives x = 0
    else if (REFINED) xk = (float)ld_dev(a.x64 + k);
    else xk = a.warm_start ? a.xg[c] + x[j] : x[j];
    a.xg[c] = xk;
  }
#undef FSG_VALID
#undef FSG_K
#undef FSG_S
  if (REFINED) err = b2 > 0.0 ? sqrt(res2 / b2) : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->iter = total_it;
    sc->spmv = total_spmv;
    sc->err = err;
    sc->b64_norm2 = b2;
    sc->res64_norm2 = res2;
    sc->refine_sweeps = sweeps;
    sc->total_unknown_iters += (unsigned long long)F * (unsigned long long)total_spmv;
    sc->total_solves += 1;
  }
}

int pcg_max_kpt() { return 4; }

size_t pcg_lds_bytes(int F, int grid, int nx) {
  const int per = (F + grid - 1) / grid;
  return (size_t)3 * (per + 2 * nx) * sizeof(float);
}

hipError_t launch_pcg(hipStream_t s, const PersistentArgs& a, int grid, int kpt) {
  void* args[] = {(void*)&a};
  const size_t lds = pcg_lds_bytes(a.F_actual, grid, a.nx);
  const void* f = nullptr;
  const bool ref = a.refined != 0;
  switch (kpt) {
    case 1: f = ref ? (const void*)k_pcg<1, true> : (const void*)k_pcg<1, false>; break;
    case 2: f = ref ? (const void*)k_pcg<2, true> : (const void*)k_pcg<2, false>; break;
    // 4 unknowns per thread only in the FP32 form: the refined form's FP64 state
    // exceeds the 128 VGPRs of a 1024-thread block there (it would spill)
    case 4:
      if (ref) return hipErrorInvalidValue;
      f = (const void*)k_pcg<4, false>;
      break;
    default: return hipErrorInvalidValue;
  }
  return hipLaunchCooperativeKernel(f, dim3(grid), dim3(kPcgThreads), args, lds, s);
}

hipError_t pcg_set_lds_limit(size_t max_bytes) {
  const void* fs[] = {(const void*)k_pcg<1, false>, (const void*)k_pcg<2, false>,
                      (const void*)k_pcg<4, false>, (const void*)k_pcg<1, true>,
                      (const void*)k_pcg<2, true>};
  for (const void* f : fs) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)max_bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace fsg
