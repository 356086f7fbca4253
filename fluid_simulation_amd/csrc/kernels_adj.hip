// kernels_adj.hip -- reverse-mode (adjoint) kernels of the PIC / FLIP / PICFLIP step
// (FluidSolver::stepPIC / stepFLIP / stepPICFLIP, src/FluidSolver.cpp:136-251).
//
// The step is differentiated with its discrete decisions held fixed (labels, the
// P2G weight test, Dirichlet zeroing, extension masks, obstacle rejection): the
// adjoint of the piecewise-smooth map on the piece the forward run took.  Names:
//   X0, V0   particle positions / velocities at the step's start (taped copy)
//   G1       grid after P2G (P2G at faces with weight > 1e-6, stale back elsewhere)
//   G2, G3   after gravity, after the first Dirichlet
//   G4       the front after the extension init (u zeroed where u- or v-invalid)
//   E        the extended field (the back buffer after the step, the next step's
//            stale source)
//   H'       the front after the step (projected + Dirichlet), D = H' - G1 the diff
//   X1, V1   particles after G2P and advection
// The buffer a step starts with as its front is overwritten before it is read (the
// extension init writes every back face), so only the back buffer and the particles
// carry adjoints between steps.
#include "fsg_kernels.h"

namespace fsg {

namespace {

constexpr int kAT = 256;

__device__ __forceinline__ int64_t gtid() {
  return (int64_t)blockIdx.x * kAT + threadIdx.x;
}

// d/dfi and d/dfj of Grid::valueInterpolated's bilinear form at corners s
__device__ __forceinline__ void interp_grad(const float* __restrict__ g, int nx,
                                            const SplatCorners& s, float& dfi, float& dfj) {
  const float v00 = g[s.i + s.j * nx], v10 = g[s.i1 + s.j * nx];
  const float v01 = g[s.i + s.j1 * nx], v11 = g[s.i1 + s.j1 * nx];
  dfi = (1.0f - s.fj) * (v10 - v00) + s.fj * (v11 - v01);
  dfj = ((1.0f - s.fi) * v01 + s.fi * v11) - ((1.0f - s.fi) * v00 + s.fi * v10);
}

// transpose of the bilinear sample: value * weights added to the 4 corners
__device__ __forceinline__ void splat_add(float* g, int nx, const SplatCorners& s, float val) {
  const float a0 = (1.0f - s.fj) * val, a1 = s.fj * val;
  unsafeAtomicAdd(&g[s.i + s.j * nx], (1.0f - s.fi) * a0);
  unsafeAtomicAdd(&g[s.i1 + s.j * nx], s.fi * a0);
  unsafeAtomicAdd(&g[s.i + s.j1 * nx], (1.0f - s.fi) * a1);
  unsafeAtomicAdd(&g[s.i1 + s.j1 * nx], s.fi * a1);
}

// The four bilinear terms of splat_add, handed to `add(i, j, term)`.
template <typename Add>
__device__ __forceinline__ void splat_terms2(const SplatCorners& s, float val, Add add) {
  const float a0 = (1.0f - s.fj) * val, a1 = s.fj * val;
  add(s.i, s.j, (1.0f - s.fi) * a0);
  add(s.i1, s.j, s.fi * a0);
  add(s.i, s.j1, (1.0f - s.fi) * a1);
  add(s.i1, s.j1, s.fi * a1);
}

__device__ __forceinline__ float inv_or_div(float a, float d, int pow2, float inv) {
  return pow2 ? a * inv : a / d;
}

}  // namespace

// ---------------------------------------------------------------- G2P + advect
// Advect (MarkerParticleSet::advect / advectAndEnsureOutsideObstacles,
// src/MarkerParticleSet.cpp:40-62): X1 = X0 + V1 dt, or X0 + V1 dt - V1 dt when
// rejected.  G2P (src/FluidSolver.cpp:921-963): V1 = r I(H', X0) + (1-r)(V0 + I(D, X0))
// per component (r = 1 for PIC, 0 for FLIP).  In: pbar = (X1bar, V1bar).  Out:
// pbar = (X0bar, V0bar) of these two phases; s = I^T V1bar on the u / v grids (the
// adjoint of H' is s, of D is (1-r) s).
// Cells of a block's LDS splat window (float2 each: the u and v adjoint terms).
constexpr int kAdjSplatTile = 2048;

__global__ void __launch_bounds__(kAT) k_adj_g2p(Geom g, const float4* __restrict__ x0,
                                                 int64_t np, const float* __restrict__ hu,
                                                 const float* __restrict__ hv,
                                                 const float* __restrict__ du,
                                                 const float* __restrict__ dv, float r, float dt,
                                                 const uint8_t* __restrict__ reject,
                                                 float4* __restrict__ pbar,
                                                 float* __restrict__ su, float* __restrict__ sv) {
  const int64_t p = gtid();
  const bool live = p < np;
  const float4 q = live ? x0[p] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 a = live ? pbar[p] : make_float4(0.f, 0.f, 0.f, 0.f);
  float vbx = a.z, vby = a.w;
  if (live && !(reject && reject[p])) {
    vbx += dt * a.x;
    vby += dt * a.y;
  }
  const int nx = g.nx, ny = g.ny;
  const SplatCorners cu =
      interp_corners(div_x(q.x, g), div_y(q.y - g.dy * 0.5f, g), nx, ny);
  const SplatCorners cv =
      interp_corners(div_x(q.x - g.dx * 0.5f, g), div_y(q.y, g), nx, ny);
  // the splat of vbx into su and vby into sv: summed in an LDS copy of the block's
  // footprint when it fits (neighbouring particles share cells), then one atomic per
  // touched value; otherwise straight to the grids
  int i0 = live ? min(cu.i, cv.i) : INT_MAX, j0 = live ? min(cu.j, cv.j) : INT_MAX;
  int i1 = live ? max(cu.i1, cv.i1) : INT_MIN, j1 = live ? max(cu.j1, cv.j1) : INT_MIN;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    i0 = min(i0, __shfl_xor(i0, off, 64));
    j0 = min(j0, __shfl_xor(j0, off, 64));
    i1 = max(i1, __shfl_xor(i1, off, 64));
    j1 = max(j1, __shfl_xor(j1, off, 64));
  }
  __shared__ int s_box[4][kAT / 64];
  __shared__ float2 s_tile[kAdjSplatTile];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    s_box[0][wave] = i0;
    s_box[1][wave] = j0;
    s_box[2][wave] = i1;
    s_box[3][wave] = j1;
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kAT / 64; ++w) {
    i0 = min(i0, s_box[0][w]);
    j0 = min(j0, s_box[1][w]);
    i1 = max(i1, s_box[2][w]);
    j1 = max(j1, s_box[3][w]);
  }
  const int bw = i1 - i0 + 1, bh = j1 - j0 + 1;
  const bool tiled = bw > 0 && bh > 0 && (int64_t)bw * bh <= kAdjSplatTile;
  if (!tiled) {
    if (live) {
      splat_add(su, nx, cu, vbx);
      splat_add(sv, nx, cv, vby);
    }
  } else {
    const int area = bw * bh;
    for (int c = threadIdx.x; c < area; c += kAT) s_tile[c] = make_float2(0.f, 0.f);
    __syncthreads();
    {
      // every lane takes part (the quad exchange needs the whole wave)
      float* tile = reinterpret_cast<float*>(s_tile);
      splat_terms2(cu, vbx, [&](int i, int j, float t) {
        lds_add_quad(tile, live ? 2 * ((i - i0) + (j - j0) * bw) : -1, live ? t : 0.f);
      });
      splat_terms2(cv, vby, [&](int i, int j, float t) {
        lds_add_quad(tile, live ? 2 * ((i - i0) + (j - j0) * bw) + 1 : -1, live ? t : 0.f);
      });
    }
    __syncthreads();
    for (int c = threadIdx.x; c < area; c += kAT) {
      const float2 v = s_tile[c];
      const int cell = (i0 + c % bw) + (j0 + c / bw) * nx;
      if (v.x != 0.f) unsafeAtomicAdd(su + cell, v.x);
      if (v.y != 0.f) unsafeAtomicAdd(sv + cell, v.y);
    }
  }
  if (!live) return;
  float ui, uj, vi, vj;
  interp_grad(hu, nx, cu, ui, uj);
  interp_grad(hv, nx, cv, vi, vj);
  float gux = r * ui, guy = r * uj, gvx = r * vi, gvy = r * vj;
  if (r != 1.0f) {
    float dui, duj, dvi, dvj;
    interp_grad(du, nx, cu, dui, duj);
    interp_grad(dv, nx, cv, dvi, dvj);
    gux += (1.0f - r) * dui;
    guy += (1.0f - r) * duj;
    gvx += (1.0f - r) * dvi;
    gvy += (1.0f - r) * dvj;
  }
  const float xbar = a.x + inv_or_div(vbx * gux + vby * gvx, g.dx, g.pow2_x, g.inv_dx);
  const float ybar = a.y + inv_or_div(vbx * guy + vby * gvy, g.dy, g.pow2_y, g.inv_dy);
  pbar[p] = make_float4(xbar, ybar, (1.0f - r) * vbx, (1.0f - r) * vby);
}

// ---------------------------------------------------------------- pressure update
// H = E - dt/density grad p at union faces U (L(i,j) | L(i-1,j) | L(i,j-1) LIQUID,
// clamped), G4 elsewhere; H' = Dirichlet(H) with the zeroed faces recorded in zmask
// (bit 0 u, bit 1 v) by the taped forward.  In: s (the G2P's adjoint of H', through
// H' itself and the diff D = H' - G1) and f (optional: H''s adjoint as the front
// after the step, read by a following semi-Lagrangian step).  Out: ebar = adjoint of
// E through the direct term, g4bar, and g1bar = -(1-r) s from the diff.
__global__ void __launch_bounds__(kAT) k_adj_post(Geom g, const uint8_t* __restrict__ label,
                                                  const uint8_t* __restrict__ zmask, float r,
                                                  int flip, const float* __restrict__ su,
                                                  const float* __restrict__ sv,
                                                  const float* __restrict__ fu,
                                                  const float* __restrict__ fv,
                                                  float* __restrict__ eu, float* __restrict__ ev,
                                                  float* __restrict__ g4u,
                                                  float* __restrict__ g4v,
                                                  float* __restrict__ g1u,
                                                  float* __restrict__ g1v) {
  const int64_t c = gtid();
  if (c >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(c % g.nx), j = (int)(c / g.nx);
  const float s_u = su[c], s_v = sv[c];
  g1u[c] = flip ? -(1.0f - r) * s_u : 0.f;
  g1v[c] = flip ? -(1.0f - r) * s_v : 0.f;
  const uint8_t z = zmask[c];
  const float hbu = (z & 1) ? 0.f : s_u + (fu ? fu[c] : 0.f);
  const float hbv = (z & 2) ? 0.f : s_v + (fv ? fv[c] : 0.f);
  const uint8_t L = label[c];
  const uint8_t Lim = label[clampi(i - 1, 0, g.nx - 1) + j * g.nx];
  const uint8_t Ljm = label[i + clampi(j - 1, 0, g.ny - 1) * g.nx];
  const bool U = L == kLiquid || Lim == kLiquid || Ljm == kLiquid;
  eu[c] = U ? hbu : 0.f;
  ev[c] = U ? hbv : 0.f;
  g4u[c] = U ? 0.f : hbu;
  g4v[c] = U ? 0.f : hbv;
}

// Adjoint of the pressure at LIQUID cells: p is used by the four faces of its cell
// (all in U): pbar = dt/density ((Hbar_u(i+1) - Hbar_u(i)) / dx + (Hbar_v(j+1) -
// Hbar_v(j)) / dy), in FP64 like the update.
__global__ void __launch_bounds__(kAT) k_adj_pgrad(Geom g, const uint8_t* __restrict__ label,
                                                   const float* __restrict__ eu,
                                                   const float* __restrict__ ev,
                                                   float dt_over_density,
                                                   double* __restrict__ pbar) {
  const int64_t c = gtid();
  if (c >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(c % g.nx), j = (int)(c / g.nx);
  double v = 0.0;
  if (label[c] == kLiquid) {
    const double ax = (i + 1 < g.nx ? (double)eu[c + 1] : 0.0) - (double)eu[c];
    const double ay = (j + 1 < g.ny ? (double)ev[c + g.nx] : 0.0) - (double)ev[c];
    v = (double)dt_over_density * (ax / (double)g.dx + ay / (double)g.dy);
  }
  pbar[c] = v;
}

// Adjoint of the divergence b(i,j) = (u(i+1,j) - u(i,j)) / dx + (v(i,j+1) - v(i,j)) / dy
// (MacGrid::divVelX/Y, include/MacGrid.h:98-111) at the solve's unknowns (info byte
// active), plus the next step's adjoint of E (its stale source), into ebar.
__global__ void __launch_bounds__(kAT) k_adj_div(Geom g, const uint8_t* __restrict__ inf,
                                                 const double* __restrict__ bbar,
                                                 const float* __restrict__ nu,
                                                 const float* __restrict__ nv,
                                                 float* __restrict__ eu, float* __restrict__ ev) {
  const int64_t c = gtid();
  if (c >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(c % g.nx), j = (int)(c / g.nx);
  auto act = [&](int64_t k) { return (inf[k] & 0x80) != 0; };
  double au = 0.0, av = 0.0;
  if (act(c)) {
    au -= bbar[c] / (double)g.dx;
    av -= bbar[c] / (double)g.dy;
  }
  if (i > 0 && act(c - 1)) au += bbar[c - 1] / (double)g.dx;
  if (j > 0 && act(c - g.nx)) av += bbar[c - g.nx] / (double)g.dy;
  eu[c] += (float)au + (nu ? nu[c] : 0.f);
  ev[c] += (float)av + (nv ? nv[c] : 0.f);
}

// ---------------------------------------------------------------- extension
// extendVelocityIndividual (src/FluidSolver.cpp:485-622).  Masks as bits (0: u,
// 1: v).  Init: M0 = LIQUID on either side of the face; both mask buffers = M0.
// Sweep k reads the front mask R_k and writes W_k (invalid interior faces between
// two non-SOLID cells with a valid neighbour) into the back mask; the buffers swap:
// R_{k+1} = B_k | W_k, B_{k+1} = R_k, so R_{k+1} = R_{k-1} | W_k (R_0 := R_1 = M0).
__global__ void __launch_bounds__(kAT) k_adj_mask_init(Geom g, const uint8_t* __restrict__ label,
                                                       uint8_t* __restrict__ r0,
                                                       uint8_t* __restrict__ b0) {
  const int64_t c = gtid();
  if (c >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(c % g.nx), j = (int)(c / g.nx);
  const uint8_t L = label[c];
  const uint8_t Lim = label[clampi(i - 1, 0, g.nx - 1) + j * g.nx];
  const uint8_t Ljm = label[i + clampi(j - 1, 0, g.ny - 1) * g.nx];
  const uint8_t m = (uint8_t)((L == kLiquid || Lim == kLiquid ? 1 : 0) |
                              (L == kLiquid || Ljm == kLiquid ? 2 : 0));
  r0[c] = m;
  if (b0) b0[c] = m;
}

// W_k bits of face c (and the neighbour counts n_u, n_v) from the read mask R.
__device__ __forceinline__ uint8_t ext_written(const Geom& g, const uint8_t* __restrict__ label,
                                               const uint8_t* __restrict__ R, int64_t c, int i,
                                               int j, int& nu_, int& nv_) {
  nu_ = nv_ = 0;
  const int nx = g.nx;
  if (i < 1 || j < 1 || i > nx - 2 || j > g.ny - 2) return 0;
  if (label[c] == kSolid) return 0;
  uint8_t w = 0;
  const uint8_t rl = R[c - 1], rd = R[c - nx], ru = R[c + nx], rr = R[c + 1], rc = R[c];
  if (!(rc & 1) && label[c - 1] != kSolid) {
    nu_ = (rl & 1) + (rd & 1) + (ru & 1) + (rr & 1);
    if (nu_ > 0) w |= 1;
  }
  if (!(rc & 2) && label[c - nx] != kSolid) {
    nv_ = ((rl >> 1) & 1) + ((rd >> 1) & 1) + ((ru >> 1) & 1) + ((rr >> 1) & 1);
    if (nv_ > 0) w |= 2;
  }
  return w;
}

__global__ void __launch_bounds__(kAT) k_adj_mask_step(Geom g, const uint8_t* __restrict__ label,
                                                       const uint8_t* __restrict__ rk,
                                                       const uint8_t* __restrict__ bk,
                                                       uint8_t* __restrict__ rn,
                                                       uint8_t* __restrict__ bn) {
  const int64_t c = gtid();
  if (c >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(c % g.nx), j = (int)(c / g.nx);
  int a, b;
  const uint8_t w = ext_written(g, label, rk, c, i, j, a, b);
  rn[c] = bk[c] | w;
  if (bn) bn[c] = rk[c];
}

// Adjoint of sweep k: a written face f took E(f) = sum_{valid nb} E(nb) / n_f, its
// old value discarded.  out(g) = [g not written] in(g) + [g read-valid] sum over
// written neighbours f of in(f) / n_f.
__global__ void __launch_bounds__(kAT) k_adj_ext_sweep(Geom g, const uint8_t* __restrict__ label,
                                                       const uint8_t* __restrict__ rk,
                                                       const float* __restrict__ inu,
                                                       const float* __restrict__ inv,
                                                       float* __restrict__ outu,
                                                       float* __restrict__ outv) {
  const int64_t c = gtid();
  const int nx = g.nx, ny = g.ny;
  if (c >= (int64_t)nx * ny) return;
  const int i = (int)(c % nx), j = (int)(c / nx);
  int n0u, n0v;
  const uint8_t w0 = ext_written(g, label, rk, c, i, j, n0u, n0v);
  float ou = (w0 & 1) ? 0.f : inu[c];
  float ov = (w0 & 2) ? 0.f : inv[c];
  const uint8_t rc = rk[c];
  if (rc & 3) {
    const int di[4] = {-1, 0, 0, 1}, dj[4] = {0, -1, 1, 0};
    for (int k = 0; k < 4; ++k) {
      const int ii = i + di[k], jj = j + dj[k];
      if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) continue;
      const int64_t f = (int64_t)ii + (int64_t)jj * nx;
      int nfu, nfv;
      const uint8_t wf = ext_written(g, label, rk, f, ii, jj, nfu, nfv);
      if ((rc & 1) && (wf & 1)) ou += inu[f] / (float)nfu;
      if ((rc & 2) && (wf & 2)) ov += inv[f] / (float)nfv;
    }
  }
  outu[c] = ou;
  outv[c] = ov;
}

// Adjoint of the extension init and of the first Dirichlet and gravity, and the
// P2G normalisation split.  E0 = G3 at M0 faces (0 elsewhere); G4 = G3 with u zeroed
// where u- or v-invalid; G3 = Dirichlet(G2) (decision recomputed from G2 = G1 +
// gravity exactly as the forward); G1 = sum / w where w > 1e-6, else the stale back
// value.  Out: q = G1bar / w at normalised faces (0 elsewhere), nbar = G1bar at the
// stale faces (the adjoint of the back buffer the step started with).
__global__ void __launch_bounds__(kAT) k_adj_pre(
    Geom g, const uint8_t* __restrict__ label, const uint8_t* __restrict__ m0,
    const float* __restrict__ e0u, const float* __restrict__ e0v, const float* __restrict__ g4u,
    const float* __restrict__ g4v, const float* __restrict__ g1u_in,
    const float* __restrict__ g1v_in, const float* __restrict__ prev_u,
    const float* __restrict__ prev_v, const float4* __restrict__ acc, float gx_dt, float gy_dt,
    int closed, float* __restrict__ qu, float* __restrict__ qv, float* __restrict__ nu,
    float* __restrict__ nv) {
  const int64_t c = gtid();
  if (c >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(c % g.nx), j = (int)(c / g.nx);
  const uint8_t M = m0[c];
  float g3u = ((M & 1) ? e0u[c] : 0.f) + ((M & 3) == 3 ? g4u[c] : 0.f);
  float g3v = ((M & 2) ? e0v[c] : 0.f) + g4v[c];
  const uint8_t L = label[c];
  const uint8_t Lim = label[clampi(i - 1, 0, g.nx - 1) + j * g.nx];
  const uint8_t Ljm = label[i + clampi(j - 1, 0, g.ny - 1) * g.nx];
  float uu = prev_u[c], vv = prev_v[c];
  if (L == kLiquid) {
    uu = uu + gx_dt;
    vv = vv + gy_dt;
  }
  bool zu, zv;
  if (closed) {
    zu = Lim == kSolid || L == kSolid;
    zv = Ljm == kSolid || L == kSolid;
  } else {
    zu = (Lim == kSolid && uu < 0.f) || (L == kSolid && uu > 0.f);
    zv = (Ljm == kSolid && vv < 0.f) || (L == kSolid && vv > 0.f);
  }
  const float gbu = (zu ? 0.f : g3u) + g1u_in[c];
  const float gbv = (zv ? 0.f : g3v) + g1v_in[c];
  const float4 a = acc[c];
  const bool nu_ok = a.y > 0.000001f, nv_ok = a.w > 0.000001f;
  qu[c] = nu_ok ? gbu / a.y : 0.f;
  qv[c] = nv_ok ? gbv / a.w : 0.f;
  nu[c] = nu_ok ? 0.f : gbu;
  nv[c] = nv_ok ? 0.f : gbv;
}

// ---------------------------------------------------------------- P2G
// transferVelocityToGridSpread (src/FluidSolver.cpp:873-919): G1(f) = sum_p W_pf V_p /
// sum_p W_pf with the splat weights of Grid::addToValueInterpolated on the pool's
// (dx, dx) spacing.  dG1(f)/dV_p = W_pf / w_f, dG1(f)/dX_p = dW_pf/dX_p (V_p - G1(f)) / w_f;
// with q = G1bar / w: V0bar += sum_f W q, X0bar += sum_f dW/dX (V - G1) q.
__global__ void __launch_bounds__(kAT) k_adj_p2g(Geom g, const float4* __restrict__ x0,
                                                 int64_t np, const float* __restrict__ g1u,
                                                 const float* __restrict__ g1v,
                                                 const float* __restrict__ qu,
                                                 const float* __restrict__ qv,
                                                 float4* __restrict__ pbar) {
  const int64_t p = gtid();
  if (p >= np) return;
  const float4 q = x0[p];
  const int nx = g.nx, ny = g.ny;
  const SplatCorners su = splat_corners(div_x(q.x, g), div_x(q.y - g.dy * 0.5f, g), nx, ny);
  const SplatCorners sv = splat_corners(div_x(q.x - g.dx * 0.5f, g), div_x(q.y, g), nx, ny);
  float dfi = 0.f, dfj = 0.f, vbu = 0.f, vbv = 0.f;
  auto acc = [&](const SplatCorners& s, const float* __restrict__ g1,
                 const float* __restrict__ qq, float vel, float& vb) {
    const int k[4] = {s.i + s.j * nx, s.i1 + s.j * nx, s.i + s.j1 * nx, s.i1 + s.j1 * nx};
    const float w[4] = {(1.0f - s.fi) * (1.0f - s.fj), s.fi * (1.0f - s.fj),
                        (1.0f - s.fi) * s.fj, s.fi * s.fj};
    const float wi[4] = {-(1.0f - s.fj), 1.0f - s.fj, -s.fj, s.fj};
    const float wj[4] = {-(1.0f - s.fi), -s.fi, 1.0f - s.fi, s.fi};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float qk = qq[k[t]];
      vb += w[t] * qk;
      const float d = (vel - g1[k[t]]) * qk;
      dfi += wi[t] * d;
      dfj += wj[t] * d;
    }
  };
  acc(su, g1u, qu, q.z, vbu);
  acc(sv, g1v, qv, q.w, vbv);
  float4 a = pbar[p];
  a.x += inv_or_div(dfi, g.dx, g.pow2_x, g.inv_dx);
  a.y += inv_or_div(dfj, g.dx, g.pow2_x, g.inv_dx);
  a.z += vbu;
  a.w += vbv;
  pbar[p] = a;
}

// ---------------------------------------------------------------- density loss
// Particle density on the cell grid: every particle splats weight 1 at (x / dx,
// y / dy) with Grid::addToValueInterpolated's weights.
// FX: the terms are added as 64-bit fixed-point integers (fx_term) into rho_fx, so
// the density does not depend on the order of the atomics (fsg_params.deterministic).
template <bool FX>
__global__ void __launch_bounds__(kAT) k_density(Geom g, const float4* __restrict__ pt,
                                                 int64_t np, float* __restrict__ rho,
                                                 unsigned long long* __restrict__ rho_fx) {
  const int64_t p = gtid();
  if (p >= np) return;
  const float4 q = pt[p];
  const SplatCorners s = splat_corners(div_x(q.x, g), div_y(q.y, g), g.nx, g.ny);
  const float a0 = (1.0f - s.fj), a1 = s.fj;
  auto add = [&](int c, float t) {
    if (FX)
      atomicAdd(rho_fx + c, (unsigned long long)fx_term(t));
    else
      unsafeAtomicAdd(&rho[c], t);
  };
  add(s.i + s.j * g.nx, (1.0f - s.fi) * a0);
  add(s.i1 + s.j * g.nx, s.fi * a0);
  add(s.i + s.j1 * g.nx, (1.0f - s.fi) * a1);
  add(s.i1 + s.j1 * g.nx, s.fi * a1);
}

// pbar = (d rho / d X)^T rhobar, zero velocity adjoints.
__global__ void __launch_bounds__(kAT) k_density_adj(Geom g, const float4* __restrict__ pt,
                                                     int64_t np, const float* __restrict__ rb,
                                                     float4* __restrict__ pbar) {
  const int64_t p = gtid();
  if (p >= np) return;
  const float4 q = pt[p];
  const SplatCorners s = splat_corners(div_x(q.x, g), div_y(q.y, g), g.nx, g.ny);
  const int nx = g.nx;
  const float r00 = rb[s.i + s.j * nx], r10 = rb[s.i1 + s.j * nx];
  const float r01 = rb[s.i + s.j1 * nx], r11 = rb[s.i1 + s.j1 * nx];
  const float dfi = (1.0f - s.fj) * (r10 - r00) + s.fj * (r11 - r01);
  const float dfj = (1.0f - s.fi) * (r01 - r00) + s.fi * (r11 - r10);
  pbar[p] = make_float4(inv_or_div(dfi, g.dx, g.pow2_x, g.inv_dx),
                        inv_or_div(dfj, g.dy, g.pow2_y, g.inv_dy), 0.f, 0.f);
}

// ---------------------------------------------------------------- semi-Lagrangian
// FluidSolver::stepSemiLagrangian (src/FluidSolver.cpp:99-134) with its decisions
// held (labels, both Dirichlet passes).  Names:
//   F0      the front at the step's start (taped copy), the carried velocity state
//   S       advectVelocitySemiLagrangian's scatter into the zeroed back (:709-772)
//   A       F0 + gravity at LIQUID cells (:276-295), then the first Dirichlet pass:
//           the pressure update's source and the back buffer after the step
//   F1      the front after the step: Dirichlet(A - dt/density grad p) at faces next
//           to LIQUID, Dirichlet(S) elsewhere (the pressure update keeps the back)
//   X1      X0 + the ODE displacement through F1 (advectParticlesWithGrid, :774-791)
// The particles' velocities are not read or written by the step: their adjoints pass.

// Adjoint of one velocity sample f(x, y) = (u(x, y), v(x, y)) of VelField::at with
// adjoint (kbx, kby): position adjoint added to (xbar, ybar), the transposed bilinear
// samples added to the grids' adjoints (au, av) by atomics.
__device__ __forceinline__ void vel_at_adj(const Geom& g, const float* __restrict__ u,
                                           const float* __restrict__ v, float x, float y,
                                           float kbx, float kby, float& xbar, float& ybar,
                                           float* au, float* av) {
  const SplatCorners cu = interp_corners(div_x(x, g), div_y(y - g.dy * 0.5f, g), g.nx, g.ny);
  const SplatCorners cv = interp_corners(div_x(x - g.dx * 0.5f, g), div_y(y, g), g.nx, g.ny);
  float ui, uj, vi, vj;
  interp_grad(u, g.nx, cu, ui, uj);
  interp_grad(v, g.nx, cv, vi, vj);
  xbar += inv_or_div(kbx * ui + kby * vi, g.dx, g.pow2_x, g.inv_dx);
  ybar += inv_or_div(kbx * uj + kby * vj, g.dy, g.pow2_y, g.inv_dy);
  if (kbx != 0.f) splat_add(au, g.nx, cu, kbx);
  if (kby != 0.f) splat_add(av, g.nx, cv, kby);
}

// Reverse of ode_disp (fsg_device.h): the displacement (ddx, ddy) of one ODE step
// from (x, y) through (u, v) with adjoint (dbx, dby).  The stages are recomputed with
// ode_disp's own float expressions, then taken in reverse order; each stage's sample
// gives its position adjoint (carried to the stages it was formed from) and adds its
// transposed sample to (au, av).  Returns the start position's adjoint.
__device__ __forceinline__ void ode_disp_adj(const Geom& g, const float* __restrict__ u,
                                             const float* __restrict__ v, int integrator,
                                             float x, float y, float h, float dbx, float dby,
                                             float& xbar, float& ybar, float* au, float* av) {
  const VelField f{u, v, g};
  xbar = 0.f;
  ybar = 0.f;
  if (integrator == kRK4) {
    // f1 at x, f2 = f3 at x + h/2, f4 at x + h (scalar shifts): each stage is a
    // function of x alone, weighted h/6, 4h/6, h/6
    const float hh = h / 2.0f;
    const float w1 = h / 6.0f, w2 = 4.0f * h / 6.0f;
    vel_at_adj(g, u, v, x, y, w1 * dbx, w1 * dby, xbar, ybar, au, av);
    vel_at_adj(g, u, v, x + hh, y + hh, w2 * dbx, w2 * dby, xbar, ybar, au, av);
    vel_at_adj(g, u, v, x + h, y + h, w1 * dbx, w1 * dby, xbar, ybar, au, av);
    return;
  }
  if (integrator == kEuler) {
    vel_at_adj(g, u, v, x + h, y + h, h * dbx, h * dby, xbar, ybar, au, av);
    return;
  }
  float k1x, k1y;
  f.at(x, y, k1x, k1y);
  if (integrator == kRK2) {
    const float mx = x + k1x * h * 0.5f, my = y + k1y * h * 0.5f;
    float mxb = 0.f, myb = 0.f;
    vel_at_adj(g, u, v, mx, my, h * dbx, h * dby, mxb, myb, au, av);
    // m = x + k1 h/2
    xbar += mxb;
    ybar += myb;
    vel_at_adj(g, u, v, x, y, 0.5f * h * mxb, 0.5f * h * myb, xbar, ybar, au, av);
    return;
  }
  // RK3: k2 at a = x + k1 h/2, k3 at b = x + k2 h 3/4, d = (2 k1 + 3 k2 + 4 k3) h/9
  const float ax = x + k1x * h * 1.0f / 2.0f, ay = y + k1y * h * 1.0f / 2.0f;
  float k2x, k2y;
  f.at(ax, ay, k2x, k2y);
  const float bx = x + k2x * h * 3.0f / 4.0f, by = y + k2y * h * 3.0f / 4.0f;
  const float c = h / 9.0f;
  float bxb = 0.f, byb = 0.f;
  vel_at_adj(g, u, v, bx, by, 4.0f * c * dbx, 4.0f * c * dby, bxb, byb, au, av);
  xbar += bxb;
  ybar += byb;
  const float k2bx = 3.0f * c * dbx + 0.75f * h * bxb, k2by = 3.0f * c * dby + 0.75f * h * byb;
  float axb = 0.f, ayb = 0.f;
  vel_at_adj(g, u, v, ax, ay, k2bx, k2by, axb, ayb, au, av);
  xbar += axb;
  ybar += ayb;
  const float k1bx = 2.0f * c * dbx + 0.5f * h * axb, k1by = 2.0f * c * dby + 0.5f * h * ayb;
  vel_at_adj(g, u, v, x, y, k1bx, k1by, xbar, ybar, au, av);
}

// advectParticlesWithGrid: X1 = X0 + disp(F1, X0, dt).  pbar (X1bar, V1bar) in,
// (X0bar, V1bar) out; F1's adjoint from the samples added to (su, sv).
__global__ void __launch_bounds__(kAT) k_adj_advect_grid(Geom g, const float4* __restrict__ x0,
                                                         int64_t np, const float* __restrict__ uf,
                                                         const float* __restrict__ vf,
                                                         int integrator, float dt,
                                                         float4* __restrict__ pbar,
                                                         float* __restrict__ su,
                                                         float* __restrict__ sv) {
  const int64_t p = gtid();
  if (p >= np) return;
  const float4 q = x0[p];
  const float4 a = pbar[p];
  float xb, yb;
  ode_disp_adj(g, uf, vf, integrator, q.x, q.y, dt, a.x, a.y, xb, yb, su, sv);
  pbar[p] = make_float4(a.x + xb, a.y + yb, a.z, a.w);
}

// Gravity and the first Dirichlet pass: A = Z1(F0 + g dt at LIQUID cells), the zeroing
// decisions recomputed from the taped F0 exactly as k_face_update takes them.
// F0bar := Abar where the face was not zeroed (the scatter's terms are added after).
__global__ void __launch_bounds__(kAT) k_adj_sl_dirichlet(Geom g, const uint8_t* __restrict__ label,
                                                          const float* __restrict__ f0u,
                                                          const float* __restrict__ f0v,
                                                          const float* __restrict__ abu,
                                                          const float* __restrict__ abv,
                                                          float gx_dt, float gy_dt, int closed,
                                                          float* __restrict__ fu,
                                                          float* __restrict__ fv) {
  const int64_t c = gtid();
  if (c >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(c % g.nx), j = (int)(c / g.nx);
  const uint8_t L = label[c];
  const uint8_t Lim = label[clampi(i - 1, 0, g.nx - 1) + j * g.nx];
  const uint8_t Ljm = label[i + clampi(j - 1, 0, g.ny - 1) * g.nx];
  float uu = f0u[c], vv = f0v[c];
  if (L == kLiquid) {
    uu = uu + gx_dt;
    vv = vv + gy_dt;
  }
  bool zu, zv;
  if (closed) {
    zu = Lim == kSolid || L == kSolid;
    zv = Ljm == kSolid || L == kSolid;
  } else {
    zu = (Lim == kSolid && uu < 0.f) || (L == kSolid && uu > 0.f);
    zv = (Ljm == kSolid && vv < 0.f) || (L == kSolid && vv > 0.f);
  }
  fu[c] = zu ? 0.f : abu[c];
  fv[c] = zv ? 0.f : abv[c];
}

// The scatter S = sum over faces next to LIQUID of splat(b_f, val_f), b_f = x_f +
// disp(F0, x_f, -dt), val_f = F0 sampled at x_f.  With Sbar: val_f's adjoint is
// Sbar gathered with the splat's weights at b_f, b_f's is val_f times the weights'
// derivative; both go back to F0 (the sample at x_f, the ODE stages of the trace)
// into (fu, fv) by atomics.
__global__ void __launch_bounds__(kAT) k_adj_sl_scatter(Geom g, const uint8_t* __restrict__ label,
                                                        const float* __restrict__ f0u,
                                                        const float* __restrict__ f0v,
                                                        const float* __restrict__ sbu,
                                                        const float* __restrict__ sbv,
                                                        int integrator, float dt,
                                                        float* __restrict__ fu,
                                                        float* __restrict__ fv) {
  const int64_t c = gtid();
  if (c >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(c % g.nx), j = (int)(c / g.nx);
  const int nx = g.nx, ny = g.ny;
  const uint8_t L = label[c];
  const uint8_t Lim = label[clampi(i - 1, 0, nx - 1) + j * nx];
  const uint8_t Ljm = label[i + clampi(j - 1, 0, ny - 1) * nx];
  const VelField f{f0u, f0v, g};
  for (int comp = 0; comp < 2; ++comp) {
    const bool on = comp == 0 ? (L == kLiquid || Lim == kLiquid) : (L == kLiquid || Ljm == kLiquid);
    if (!on) continue;
    const float xp = comp == 0 ? (float)i * g.dx : ((float)i + 0.5f) * g.dx;
    const float yp = comp == 0 ? ((float)j + 0.5f) * g.dy : (float)j * g.dy;
    float ddx, ddy;
    ode_disp(f, integrator, xp, yp, -dt, ddx, ddy);
    const float xb = xp + ddx, yb = yp + ddy;
    const float* sb = comp == 0 ? sbu : sbv;
    const SplatCorners s = comp == 0
        ? splat_corners(div_x(xb, g), div_y(yb - g.dy * 0.5f, g), nx, ny)
        : splat_corners(div_x(xb - g.dx * 0.5f, g), div_y(yb, g), nx, ny);
    const float s00 = sb[s.i + s.j * nx], s10 = sb[s.i1 + s.j * nx];
    const float s01 = sb[s.i + s.j1 * nx], s11 = sb[s.i1 + s.j1 * nx];
    const float valbar = (1.0f - s.fj) * ((1.0f - s.fi) * s00 + s.fi * s10) +
                         s.fj * ((1.0f - s.fi) * s01 + s.fi * s11);
    float dfi, dfj;
    interp_grad(sb, nx, s, dfi, dfj);
    const float val = comp == 0 ? vel_x_at(f0u, g, xp, yp) : vel_y_at(f0v, g, xp, yp);
    const float bbx = inv_or_div(val * dfi, g.dx, g.pow2_x, g.inv_dx);
    const float bby = inv_or_div(val * dfj, g.dy, g.pow2_y, g.inv_dy);
    if (valbar != 0.f) {
      const SplatCorners o = comp == 0
          ? interp_corners(div_x(xp, g), div_y(yp - g.dy * 0.5f, g), nx, ny)
          : interp_corners(div_x(xp - g.dx * 0.5f, g), div_y(yp, g), nx, ny);
      splat_add(comp == 0 ? fu : fv, nx, o, valbar);
    }
    if (bbx != 0.f || bby != 0.f) {
      float xs, ys;  // the trace starts at the fixed face position: its adjoint is dropped
      ode_disp_adj(g, f0u, f0v, integrator, xp, yp, -dt, bbx, bby, xs, ys, fu, fv);
    }
  }
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(int64_t n) { return (unsigned)((n + kAT - 1) / kAT); }

void launch_adj_g2p(hipStream_t s, const Geom& g, const float4* x0, int64_t np, const float* hu,
                    const float* hv, const float* du, const float* dv, float r, float dt,
                    const uint8_t* reject, float4* pbar, float* su, float* sv) {
  if (np > 0)
    k_adj_g2p<<<nblk(np), kAT, 0, s>>>(g, x0, np, hu, hv, du, dv, r, dt, reject, pbar, su, sv);
}

void launch_adj_post(hipStream_t s, const Geom& g, const uint8_t* label, const uint8_t* zmask,
                     float r, int flip, const float* su, const float* sv, const float* fu,
                     const float* fv, float* eu, float* ev, float* g4u, float* g4v, float* g1u,
                     float* g1v) {
  k_adj_post<<<nblk((int64_t)g.nx * g.ny), kAT, 0, s>>>(g, label, zmask, r, flip, su, sv, fu, fv,
                                                        eu, ev, g4u, g4v, g1u, g1v);
}

void launch_adj_pgrad(hipStream_t s, const Geom& g, const uint8_t* label, const float* eu,
                      const float* ev, float dt_over_density, double* pbar) {
  k_adj_pgrad<<<nblk((int64_t)g.nx * g.ny), kAT, 0, s>>>(g, label, eu, ev, dt_over_density, pbar);
}

void launch_adj_div(hipStream_t s, const Geom& g, const uint8_t* inf, const double* bbar,
                    const float* nu, const float* nv, float* eu, float* ev) {
  k_adj_div<<<nblk((int64_t)g.nx * g.ny), kAT, 0, s>>>(g, inf, bbar, nu, nv, eu, ev);
}

void launch_adj_mask_init(hipStream_t s, const Geom& g, const uint8_t* label, uint8_t* r0,
                          uint8_t* b0) {
  k_adj_mask_init<<<nblk((int64_t)g.nx * g.ny), kAT, 0, s>>>(g, label, r0, b0);
}

void launch_adj_mask_step(hipStream_t s, const Geom& g, const uint8_t* label, const uint8_t* rk,
                          const uint8_t* bk, uint8_t* rn, uint8_t* bn) {
  k_adj_mask_step<<<nblk((int64_t)g.nx * g.ny), kAT, 0, s>>>(g, label, rk, bk, rn, bn);
}

void launch_adj_ext_sweep(hipStream_t s, const Geom& g, const uint8_t* label, const uint8_t* rk,
                          const float* inu, const float* inv, float* outu, float* outv) {
  k_adj_ext_sweep<<<nblk((int64_t)g.nx * g.ny), kAT, 0, s>>>(g, label, rk, inu, inv, outu, outv);
}

void launch_adj_pre(hipStream_t s, const Geom& g, const uint8_t* label, const uint8_t* m0,
                    const float* e0u, const float* e0v, const float* g4u, const float* g4v,
                    const float* g1u, const float* g1v, const float* prev_u,
                    const float* prev_v, const float4* acc, float gx_dt, float gy_dt,
                    int closed, float* qu, float* qv, float* nu, float* nv) {
  k_adj_pre<<<nblk((int64_t)g.nx * g.ny), kAT, 0, s>>>(g, label, m0, e0u, e0v, g4u, g4v, g1u,
                                                       g1v, prev_u, prev_v, acc, gx_dt, gy_dt,
                                                       closed, qu, qv, nu, nv);
}

void launch_adj_p2g(hipStream_t s, const Geom& g, const float4* x0, int64_t np, const float* g1u,
                    const float* g1v, const float* qu, const float* qv, float4* pbar) {
  if (np > 0) k_adj_p2g<<<nblk(np), kAT, 0, s>>>(g, x0, np, g1u, g1v, qu, qv, pbar);
}

void launch_density(hipStream_t s, const Geom& g, const float4* pt, int64_t np, float* rho,
                    long long* rho_fx) {
  if (rho_fx) {
    if (np > 0)
      k_density<true><<<nblk(np), kAT, 0, s>>>(g, pt, np, rho,
                                               reinterpret_cast<unsigned long long*>(rho_fx));
    launch_fx_to_float(s, (int64_t)g.nx * g.ny, rho_fx, rho);
    return;
  }
  if (np > 0) k_density<false><<<nblk(np), kAT, 0, s>>>(g, pt, np, rho, nullptr);
}

void launch_density_adj(hipStream_t s, const Geom& g, const float4* pt, int64_t np,
                        const float* rb, float4* pbar) {
  if (np > 0) k_density_adj<<<nblk(np), kAT, 0, s>>>(g, pt, np, rb, pbar);
}

void launch_adj_advect_grid(hipStream_t s, const Geom& g, const float4* x0, int64_t np,
                            const float* uf, const float* vf, int integrator, float dt,
                            float4* pbar, float* su, float* sv) {
  if (np > 0)
    k_adj_advect_grid<<<nblk(np), kAT, 0, s>>>(g, x0, np, uf, vf, integrator, dt, pbar, su, sv);
}

void launch_adj_sl(hipStream_t s, const Geom& g, const uint8_t* label, const float* f0u,
                   const float* f0v, const float* abu, const float* abv, const float* sbu,
                   const float* sbv, float gx_dt, float gy_dt, int closed, int integrator,
                   float dt, float* fu, float* fv) {
  const unsigned nb = nblk((int64_t)g.nx * g.ny);
  k_adj_sl_dirichlet<<<nb, kAT, 0, s>>>(g, label, f0u, f0v, abu, abv, gx_dt, gy_dt, closed, fu,
                                        fv);
  k_adj_sl_scatter<<<nb, kAT, 0, s>>>(g, label, f0u, f0v, sbu, sbv, integrator, dt, fu, fv);
}

}  // namespace fsg
