// fsg_common.h -- reduction and device-coherent memory helpers shared by the
// pressure-solver kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace fsg {

// Sum NV doubles over a block of NT threads; the result is returned in every
// thread.  The order is fixed (wave shuffle tree, then waves in order), so equal
// inputs give bit-identical sums in every block and every launch.
// sm: NV * (NT / 64) doubles of LDS.
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sm) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int q = 0; q < NV; ++q)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) sm[q * NW + wave] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += sm[q * NW + w];
    v[q] = s;
  }
}

// t[n] for n in 0..4 without dynamic indexing (keeps the table in registers).
__device__ __forceinline__ float sel5(int n, const float (&t)[5]) {
  float r = t[0];
  r = n == 1 ? t[1] : r;
  r = n == 2 ? t[2] : r;
  r = n == 3 ? t[3] : r;
  r = n == 4 ? t[4] : r;
  return r;
}

// Agent-scope relaxed atomic load / store: served at the device coherence point,
// so values written by blocks on other XCDs are seen without cache maintenance.
template <typename T>
__device__ __forceinline__ T ld_dev(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_dev(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace fsg
