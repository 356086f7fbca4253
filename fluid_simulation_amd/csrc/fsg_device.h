// fsg_device.h -- device helpers shared by the gfx950 kernels.
//
// Every helper rounds exactly like the reference's float build (no contraction:
// the library is compiled with -ffp-contract=off and correctly rounded fp32
// division), so per-element results can be compared bit-for-bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsg {

constexpr uint8_t kLiquid = 0, kAir = 1, kSolid = 2;  // include/MacGrid.h:14-17

// Grid geometry passed by value to kernels.
struct Geom {
  int nx, ny;
  float dx, dy;          // MacGrid spacing (length / size in float, src/MacGrid.cpp:8)
  float inv_dx, inv_dy;  // exact reciprocals when the spacing is a power of two
  int pow2_x, pow2_y;    // 1: x / dx == x * inv_dx exactly
  float len_x, len_y;    // GridInterface::lengthX() = size * delta (include/Grid.h:54-55)
};

// MathDefinitions.h:16-19: CLAMP is a MyFloat function, integer clamps go through
// float.  Exact for |v| < 2^24, which covers every index here.
__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  float d = (float)v;
  const float t = d < (float)lo ? (float)lo : d;
  return (int)(t > (float)hi ? (float)hi : t);
}

__device__ __forceinline__ float div_x(float a, const Geom& g) {
  return g.pow2_x ? a * g.inv_dx : a / g.dx;
}
__device__ __forceinline__ float div_y(float a, const Geom& g) {
  return g.pow2_y ? a * g.inv_dy : a / g.dy;
}

// Grid<T>::valueInterpolated (include/Grid.h:117-144) on a row-major grid:
// truncating casts, fraction before clamping, i+1 clamped after i, x then y lerp.
// xs, ys are the coordinates already divided by the spacing; nx, ny bound the
// indices and `pitch` is the row stride (nx for a grid of its own, larger for one
// scene inside a packed grid).
__device__ __forceinline__ float interp_pitched(const float* __restrict__ g, int nx, int ny,
                                                int pitch, float xs, float ys) {
  int i = (int)xs;
  int j = (int)ys;
  const float fi = xs - (float)i;
  const float fj = ys - (float)j;
  i = clampi(i, 0, nx - 1);
  j = clampi(j, 0, ny - 1);
  const int i1 = clampi(i + 1, 0, nx - 1);
  const int j1 = clampi(j + 1, 0, ny - 1);
  const float v00 = g[i + j * pitch];
  const float v10 = g[i1 + j * pitch];
  const float v01 = g[i + j1 * pitch];
  const float v11 = g[i1 + j1 * pitch];
  const float v0 = (1.0f - fi) * v00 + fi * v10;
  const float v1 = (1.0f - fi) * v01 + fi * v11;
  return (1.0f - fj) * v0 + fj * v1;
}

__device__ __forceinline__ float interp_scaled(const float* __restrict__ g, int nx, int ny,
                                               float xs, float ys) {
  return interp_pitched(g, nx, ny, nx, xs, ys);
}

// MacGrid::velXInterpolated: sample u at (x, y - dy*0.5); the offset is formed in
// double in the reference and rounded once to float, which equals the float
// subtraction of the exact half spacing (a single correctly rounded operation).
__device__ __forceinline__ float vel_x_at(const float* __restrict__ u, const Geom& g, float x,
                                          float y) {
  const float yy = y - g.dy * 0.5f;
  return interp_scaled(u, g.nx, g.ny, div_x(x, g), div_y(yy, g));
}
// MacGrid::velYInterpolated: sample v at (x - dx*0.5, y).
__device__ __forceinline__ float vel_y_at(const float* __restrict__ v, const Geom& g, float x,
                                          float y) {
  const float xx = x - g.dx * 0.5f;
  return interp_scaled(v, g.nx, g.ny, div_x(xx, g), div_y(y, g));
}

// LDS float add of (addr, v) per lane, summed first over the lanes of each quad that
// target the same address (neighbouring particles of one cell share their corners):
// lanes L and L^1, then L and L^2, compare addresses through DPP quad permutations and
// the lower lane of an equal pair takes both values.  Every value is added once; a
// lane left with 0 (or addr < 0) issues no atomic, so the LDS sees up to 4x fewer
// conflicting atomics.
__device__ __forceinline__ int quad_xor1(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
}
__device__ __forceinline__ int quad_xor2(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
}
__device__ __forceinline__ void lds_add_quad(float* tile, int addr, float v) {
  const int lane = threadIdx.x & 63;
  const int a1 = quad_xor1(addr);
  const float v1 = __int_as_float(quad_xor1(__float_as_int(v)));
  if (a1 == addr) v = (lane & 1) ? 0.f : v + v1;
  const int a2 = quad_xor2(addr);
  const float v2 = __int_as_float(quad_xor2(__float_as_int(v)));
  if (a2 == addr) v = (lane & 2) ? 0.f : v + v2;
  if (addr >= 0 && v != 0.f) atomicAdd(tile + addr, v);
}

// Fixed-point splat terms (deterministic accumulation): t -> rint(t * 2^40) as a
// 64-bit integer, and a sum back to float through one double.
constexpr float kFxScale = 1099511627776.0f;  // 2^40
__device__ __forceinline__ long long fx_term(float t) {
  return (long long)__builtin_rintf(t * kFxScale);
}
__device__ __forceinline__ float fx_value(long long v) {
  return (float)((double)v * (1.0 / 1099511627776.0));
}

struct SplatCorners {
  int i, j, i1, j1;
  float fi, fj;
};

// Corners and fractions of Grid<T>::addToValueInterpolated (include/Grid.h:
// 152-184) at scaled coordinates: i+1 taken before the clamp.
__device__ __forceinline__ SplatCorners splat_corners(float xs, float ys, int nx, int ny) {
  SplatCorners s;
  s.i = (int)xs;
  s.j = (int)ys;
  s.i1 = s.i + 1;
  s.j1 = s.j + 1;
  s.fi = xs - (float)s.i;
  s.fj = ys - (float)s.j;
  s.i = clampi(s.i, 0, nx - 1);
  s.j = clampi(s.j, 0, ny - 1);
  s.i1 = clampi(s.i1, 0, nx - 1);
  s.j1 = clampi(s.j1, 0, ny - 1);
  return s;
}

// Corners and fractions of Grid<T>::valueInterpolated (include/Grid.h:117-144):
// i+1 taken after the clamp (differs from the splat only below the first node).
__device__ __forceinline__ SplatCorners interp_corners(float xs, float ys, int nx, int ny) {
  SplatCorners s;
  s.i = (int)xs;
  s.j = (int)ys;
  s.fi = xs - (float)s.i;
  s.fj = ys - (float)s.j;
  s.i = clampi(s.i, 0, nx - 1);
  s.j = clampi(s.j, 0, ny - 1);
  s.i1 = clampi(s.i + 1, 0, nx - 1);
  s.j1 = clampi(s.j + 1, 0, ny - 1);
  return s;
}

// Grid<T>::addToValueInterpolated (include/Grid.h:152-184) with hardware fp32
// atomics: i+1 taken before the clamp, spread in y then x.
__device__ __forceinline__ void splat_scaled(float* g, int nx, int ny, float xs, float ys,
                                             float value) {
  int i = (int)xs;
  int j = (int)ys;
  int i1 = i + 1, j1 = j + 1;
  const float fi = xs - (float)i;
  const float fj = ys - (float)j;
  i = clampi(i, 0, nx - 1);
  j = clampi(j, 0, ny - 1);
  i1 = clampi(i1, 0, nx - 1);
  j1 = clampi(j1, 0, ny - 1);
  const float val0 = (1.0f - fj) * value;
  const float val1 = fj * value;
  unsafeAtomicAdd(&g[i + j * nx], (1.0f - fi) * val0);
  unsafeAtomicAdd(&g[i1 + j * nx], fi * val0);
  unsafeAtomicAdd(&g[i + j1 * nx], (1.0f - fi) * val1);
  unsafeAtomicAdd(&g[i1 + j1 * nx], fi * val1);
}

// Velocity field sampler used by the ODE steps (FluidSolver::getAdvectedPosition,
// src/FluidSolver.cpp:793-814).
struct VelField {
  const float* __restrict__ u;
  const float* __restrict__ v;
  Geom g;
  __device__ __forceinline__ void at(float x, float y, float& ox, float& oy) const {
    ox = vel_x_at(u, g, x, y);
    oy = vel_y_at(v, g, x, y);
  }
};

constexpr int kRK3 = 0, kEuler = 1, kRK2 = 2, kRK4 = 3;

// Displacement of one ODE step in Vec2 float ops (include/FluidSolver.h:120-141).
//  RK3 (include/OdeSolver.h:102-113): ((k1*2 + k2*3 + k3*4) * h * 1) / 9 with k2 at
//      x + ((k1*h)*1)/2 and k3 at x + ((k2*h)*3)/4.
//  Euler (include/OdeSolver.h:78-86): f(x + h) * h, the scalar added to both axes.
//  RK2 (new, explicit midpoint): k2 at x + (k1*h)*0.5, displacement k2*h.
//  RK4 (include/OdeSolver.h:117-142): the template's stages with its scalar shift of
//      both coordinates (Vec2 + MyFloat, like Euler): f1 = f(x), f2 = f3 = f(x + h/2),
//      f4 = f(x + h), displacement (f1 + f2*2 + f3*2 + f4) * h / 6.  (The template
//      cannot be instantiated with the solver's Vec2, which has no MyFloat * Vec2;
//      the Vec2 * MyFloat order is used.)
__device__ __forceinline__ void ode_disp(const VelField& f, int integrator, float x, float y,
                                         float h, float& ddx, float& ddy) {
  if (integrator == kRK4) {
    float f1x, f1y, f2x, f2y, f4x, f4y;
    f.at(x, y, f1x, f1y);
    const float hh = h / 2.0f;
    f.at(x + hh, y + hh, f2x, f2y);
    f.at(x + h, y + h, f4x, f4y);
    ddx = (f1x + f2x * 2.0f + f2x * 2.0f + f4x) * h / 6.0f;
    ddy = (f1y + f2y * 2.0f + f2y * 2.0f + f4y) * h / 6.0f;
    return;
  }
  if (integrator == kEuler) {
    float kx, ky;
    f.at(x + h, y + h, kx, ky);
    ddx = kx * h;
    ddy = ky * h;
    return;
  }
  float k1x, k1y;
  f.at(x, y, k1x, k1y);
  if (integrator == kRK2) {
    float k2x, k2y;
    f.at(x + k1x * h * 0.5f, y + k1y * h * 0.5f, k2x, k2y);
    ddx = k2x * h;
    ddy = k2y * h;
    return;
  }
  float k2x, k2y, k3x, k3y;
  f.at(x + k1x * h * 1.0f / 2.0f, y + k1y * h * 1.0f / 2.0f, k2x, k2y);
  f.at(x + k2x * h * 3.0f / 4.0f, y + k2y * h * 3.0f / 4.0f, k3x, k3y);
  ddx = (k1x * 2.0f + k2x * 3.0f + k3x * 4.0f) * h * 1.0f / 9.0f;
  ddy = (k1y * 2.0f + k2y * 3.0f + k3y * 4.0f) * h * 1.0f / 9.0f;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).  Map
// dispatch order to logical tiles so that each XCD gets one contiguous band of
// tiles (rows of the grid / runs of particles) and its L2 sees the halo rows.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int kXcd = 8;
  const int per = nb / kXcd;          // full tiles per XCD
  const int full = per * kXcd;
  if (b >= full) return b;            // the ragged tail stays in place
  return (b % kXcd) * per + b / kXcd;
}

}  // namespace fsg
