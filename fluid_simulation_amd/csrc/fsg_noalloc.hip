// fsg_noalloc.hip -- fsg_p2g_noalloc: the P2G transfer of fsg_run_phase(FSG_PH_P2G)
// on caller-owned device buffers, for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops, harnesses that forbid allocation).
//
// Its only scratch, the four fixed-point sums per cell of the deterministic splat
// (fsg_params.deterministic), is the file-scope device array `scratch` below: each
// GPU the library runs on holds one copy of it until the process ends.  Its device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) and kept; host code passes that address, never the array's name, to the
// kernels.  Every call on a GPU shares the array: calls that overlap on one GPU race.
// A call clears the sums it uses, splats, and normalises; it calls no hipMalloc,
// hipFree or allocating library routine.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "fsg.h"
#include "fsg_kernels.h"

namespace {

constexpr int64_t kMaxCells = FSG_NOALLOC_MAX_CELLS;
constexpr int kMaxDevices = 64;

// 4 x int64 per cell: (u sum, u weight, v sum, v weight) in 2^-40 units
static __device__ __align__(256) long long scratch[4 * kMaxCells];

std::mutex g_mu;
std::atomic<long long*> g_ptr[kMaxDevices];

// scratch's address on the current device (looked up on the first call there)
long long* scratch_on_device(int dev) {
  long long* p = g_ptr[dev].load(std::memory_order_acquire);
  if (p) return p;
  std::lock_guard<std::mutex> lk(g_mu);
  p = g_ptr[dev].load(std::memory_order_relaxed);
  if (p) return p;
  void* addr = nullptr;
  if (hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)) != hipSuccess || !addr) return nullptr;
  p = static_cast<long long*>(addr);
  g_ptr[dev].store(p, std::memory_order_release);
  return p;
}

__global__ void __launch_bounds__(256) k_clear(int64_t n, long long* __restrict__ fx) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < n) fx[k] = 0;
}

// transferVelocityToGridSpread's normalise (src/FluidSolver.cpp:898-917) on the
// fixed-point sums: back := sum / weight where weight > 1e-6, as k_face_update does
// with FU_P2G after the deterministic splat (the same conversion, the same division).
__global__ void __launch_bounds__(256) k_normalise(int64_t C, const long long* __restrict__ fx,
                                                   float* __restrict__ u, float* __restrict__ v) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= C) return;
  const float su = fsg::fx_value(fx[4 * c + 0]), wu = fsg::fx_value(fx[4 * c + 1]);
  const float sv = fsg::fx_value(fx[4 * c + 2]), wv = fsg::fx_value(fx[4 * c + 3]);
  if (wu > 0.000001f) u[c] = su / wu;
  if (wv > 0.000001f) v[c] = sv / wv;
}

inline unsigned blocks_of(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int fsg_p2g_noalloc(const float* particles_xyuv, int64_t n, int32_t nx, int32_t ny,
                               float length_x, float length_y, float* back_u, float* back_v,
                               void* stream) {
  // every check before the first HIP call
  if (nx < 1 || ny < 1 || n < 0 || n > (int64_t)INT32_MAX * 256) return FSG_E_INVALID;
  if ((int64_t)nx * ny > kMaxCells) return FSG_E_INVALID;
  if (!(length_x > 0.f) || !(length_y > 0.f)) return FSG_E_INVALID;
  if (!back_u || !back_v || (n > 0 && !particles_xyuv)) return FSG_E_INVALID;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FSG_E_HIP;
  long long* fx = scratch_on_device(dev);
  if (!fx) return FSG_E_HIP;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const fsg::Geom g = fsg::make_geom(nx, ny, length_x, length_y);
  const int64_t C = (int64_t)nx * ny;
  k_clear<<<blocks_of(4 * C), 256, 0, s>>>(4 * C, fx);
  if (n > 0)
    fsg::launch_particle_scatter(s, g, reinterpret_cast<const float4*>(particles_xyuv), n,
                                 nullptr, nullptr, /*do_classify=*/0, /*do_p2g=*/1, nullptr, fx);
  k_normalise<<<blocks_of(C), 256, 0, s>>>(C, fx, back_u, back_v);
  return hipGetLastError() == hipSuccess ? FSG_OK : FSG_E_HIP;
}
