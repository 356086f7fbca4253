/*
 * oracle.c -- CPU restatement of kbladin/Fluid_Simulation's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled twice: -DREAL=float
 * -DSFX=_f32 and -DREAL=double -DSFX=_f64.  Built with -O2 -ffp-contract=off so
 * every float operation rounds exactly where the reference's does (gcc x86-64
 * without FMA, as the reference is built, CMakeLists.txt:24).
 *
 * Expressions that mix a double literal with MyFloat (e.g. `y - 0.5 * dy`,
 * `(j + 0.5) * dy`, `w > 0.000001`, `1 / pow(dx, 2)`) are written with the same
 * C promotions, so the f32 build rounds like the reference's float build.
 *
 * Parity pins (what checks each function; the tests live in tests/):
 *   classify_particles  hand-placed particles incl. exact integers, k+eps, out of range
 *   p2g                 O(C*P) tent-function brute force, uniform-velocity field,
 *                       conservation of splatted mass
 *   gravity/dirichlet   per-face truth tables
 *   extend              hand-computed 5x5 cases; valid faces unchanged; layer bound
 *   pressure            scipy spsolve on an independently assembled Laplacian;
 *                       hydrostatic column closed form; post-update divergence
 *   cg (Eigen restated) residual bound; equals direct solve when converged;
 *                       the truncated 100-iteration iterate of Eigen's SIMD build is
 *                       parity unpinned (Eigen is absent)
 *   interp / splat      exact on affine (interp) fields; partition of unity (splat)
 *   rk3                 stability polynomial of a linear field = cubic Taylor poly
 *   g2p / advect        closed forms on affine fields; rejection decision table
 *   full step           the buffer state machine of src/FluidSolver.cpp:211-251 by
 *                       invariants derived from the source (prev = the brute-force
 *                       P2G grid over the stale back buffer; non-union faces; the
 *                       no-fluid early return)
 *   decision replay     (test support, no counterpart in the reference) a step that
 *                       replays the decisions it recorded itself is the same step,
 *                       bit for bit; a replayed decision that differs changes it
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifndef REAL
#define REAL float
#endif
#ifndef SFX
#define SFX _f32
#endif
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SFX)
#define STATE CAT(orc_state, SFX)

/* MathDefinitions.h:16-19: CLAMP takes MyFloat, so integer clamps round-trip
 * through REAL. */
static inline int clampi(int v, int lo, int hi) {
  REAL d = (REAL)v, mn = (REAL)lo, mx = (REAL)hi;
  const REAL t = d < mn ? mn : d;
  return (int)(t > mx ? mx : t);
}

#define IDX(s, i, j) ((size_t)(i) + (size_t)(j) * (size_t)(s)->nx)

static inline REAL *U_F(STATE *s) { return s->front ? s->u1 : s->u0; }
static inline REAL *V_F(STATE *s) { return s->front ? s->v1 : s->v0; }
static inline REAL *U_B(STATE *s) { return s->front ? s->u0 : s->u1; }
static inline REAL *V_B(STATE *s) { return s->front ? s->v0 : s->v1; }
static inline void swap_vel(STATE *s) { s->front ^= 1; } /* MacGrid.cpp:89-93 */

/* MacGrid::cellType, include/MacGrid.h:92-97: clamped lookup. */
static inline int cell(const STATE *s, int i, int j) {
  i = clampi(i, 0, s->nx - 1);
  j = clampi(j, 0, s->ny - 1);
  return s->label[IDX(s, i, j)];
}

/* Grid<T>::valueInterpolated, include/Grid.h:117-144 (truncating casts, fraction
 * taken before clamping, x-lerp then y-lerp). */
REAL F(orc_interp)(const REAL *g, int32_t nx, int32_t ny, REAL dx, REAL dy, REAL x,
                   REAL y) {
  int i = (int)(x / dx);
  int j = (int)(y / dy);
  REAL i_frac = x / dx - (REAL)i;
  REAL j_frac = y / dy - (REAL)j;
  i = clampi(i, 0, nx - 1);
  j = clampi(j, 0, ny - 1);
  int i1 = clampi(i + 1, 0, nx - 1);
  int j1 = clampi(j + 1, 0, ny - 1);
  REAL v00 = g[(size_t)i + (size_t)j * nx];
  REAL v10 = g[(size_t)i1 + (size_t)j * nx];
  REAL v01 = g[(size_t)i + (size_t)j1 * nx];
  REAL v11 = g[(size_t)i1 + (size_t)j1 * nx];
  REAL v0 = (1 - i_frac) * v00 + i_frac * v10;
  REAL v1 = (1 - i_frac) * v01 + i_frac * v11;
  return (1 - j_frac) * v0 + j_frac * v1;
}

/* Grid<T>::addToValueInterpolated, include/Grid.h:152-184 (i+1 taken before the
 * clamp; spread in y, then in x; four += in the order 00, 10, 01, 11). */
void F(orc_splat)(REAL *g, int32_t nx, int32_t ny, REAL dx, REAL dy, REAL x, REAL y,
                  REAL value) {
  int i = (int)(x / dx);
  int j = (int)(y / dy);
  int i1 = i + 1, j1 = j + 1;
  REAL i_frac = x / dx - (REAL)i;
  REAL j_frac = y / dy - (REAL)j;
  i = clampi(i, 0, nx - 1);
  j = clampi(j, 0, ny - 1);
  i1 = clampi(i1, 0, nx - 1);
  j1 = clampi(j1, 0, ny - 1);
  REAL val0 = (1 - j_frac) * value;
  REAL val1 = j_frac * value;
  REAL v00 = (1 - i_frac) * val0, v10 = i_frac * val0;
  REAL v01 = (1 - i_frac) * val1, v11 = i_frac * val1;
  g[(size_t)i + (size_t)j * nx] += v00;
  g[(size_t)i1 + (size_t)j * nx] += v10;
  g[(size_t)i + (size_t)j1 * nx] += v01;
  g[(size_t)i1 + (size_t)j1 * nx] += v11;
}

/* MacGrid::velXInterpolated / velYInterpolated, include/MacGrid.h:66-79: the
 * offset `y - _DELTA_Y * 0.5` is formed in double and rounded to MyFloat. */
static inline REAL vel_x_at(const STATE *s, const REAL *u, REAL x, REAL y) {
  return F(orc_interp)(u, s->nx, s->ny, s->dx, s->dy, x, (REAL)(y - s->dy * 0.5));
}
static inline REAL vel_y_at(const STATE *s, const REAL *v, REAL x, REAL y) {
  return F(orc_interp)(v, s->nx, s->ny, s->dx, s->dy, (REAL)(x - s->dx * 0.5), y);
}

/* ------------------------------------------------------------------ classify */

/* MacGrid::clearCellTypeBuffer, src/MacGrid.cpp:32-50, plus the optional static
 * solid mask (an extension: the reference has no interior solids). */
static void clear_labels(STATE *s) {
  for (int j = 0; j < s->ny; ++j)
    for (int i = 0; i < s->nx; ++i) {
      int border = (i == 0 || j == 0 || i == s->nx - 1 || j == s->ny - 1);
      uint8_t l = border ? ORC_SOLID : ORC_AIR;
      if (s->solid_mask && s->solid_mask[IDX(s, i, j)]) l = ORC_SOLID;
      s->label[IDX(s, i, j)] = l;
    }
}

/* FluidDomain::classifyCells(MarkerParticleSet&), src/FluidDomain.cpp:150-180:
 * cell = (int)((x / lengthX) * sizeX), clamped; the border (and static solids)
 * are re-marked SOLID afterwards. */
static void classify_particles(STATE *s) {
  if (s->replay_label) { /* decision replay: the recorded labels */
    memcpy(s->label, s->replay_label, (size_t)s->nx * s->ny);
    return;
  }
  clear_labels(s);
  const REAL len_x = (REAL)s->nx * s->dx; /* GridInterface::lengthX, Grid.h:54 */
  const REAL len_y = (REAL)s->ny * s->dy;
  for (int64_t p = 0; p < s->n_particles; ++p) {
    const REAL *q = s->particles + 4 * p;
    int x = (int)((q[0] / len_x) * (REAL)s->nx);
    int y = (int)((q[1] / len_y) * (REAL)s->ny);
    x = clampi(x, 0, s->nx - 1);
    y = clampi(y, 0, s->ny - 1);
    s->label[IDX(s, x, y)] = ORC_LIQUID;
  }
  for (int j = 0; j < s->ny; ++j)
    for (int i = 0; i < s->nx; ++i) {
      int border = (i == 0 || j == 0 || i == s->nx - 1 || j == s->ny - 1);
      if (border || (s->solid_mask && s->solid_mask[IDX(s, i, j)]))
        s->label[IDX(s, i, j)] = ORC_SOLID;
    }
}

/* FluidDomain::classifyCells(LevelSet&), src/FluidDomain.cpp:129-148. */
static void classify_levelset(STATE *s) {
  clear_labels(s);
  for (int j = 0; j < s->ny; ++j)
    for (int i = 0; i < s->nx; ++i) {
      if (s->phi[IDX(s, i, j)] < 0) s->label[IDX(s, i, j)] = ORC_LIQUID;
      int border = (i == 0 || j == 0 || i == s->nx - 1 || j == s->ny - 1);
      if (border || (s->solid_mask && s->solid_mask[IDX(s, i, j)]))
        s->label[IDX(s, i, j)] = ORC_SOLID;
    }
}

/* ----------------------------------------------------------------------- P2G */

/* FluidSolver::transferVelocityToGridSpread, src/FluidSolver.cpp:873-919: zero the
 * four accumulators, splat u at (x, y - dy/2) and v at (x - dx/2, y) per particle
 * in particle order, then back := sum / weight where weight > 1e-6 (other faces
 * keep the stale back value), then swap. */
static void p2g(STATE *s) {
  const size_t C = (size_t)s->nx * s->ny;
  memset(s->acc_u, 0, C * sizeof(REAL));
  memset(s->acc_v, 0, C * sizeof(REAL));
  memset(s->acc_wu, 0, C * sizeof(REAL));
  memset(s->acc_wv, 0, C * sizeof(REAL));
  /* The solver's pool copy sets delta_y := delta_x (src/FluidSolver.cpp:56-63);
   * the accumulators are grids of that pool, so they splat with dx on both axes
   * while the half-cell offsets use the MacGrid's own dx, dy. */
  const REAL pool_dy = s->dx;
  for (int64_t p = 0; p < s->n_particles; ++p) {
    const REAL *q = s->particles + 4 * p;
    const REAL xu = q[0], yu = (REAL)(q[1] - 0.5 * s->dy);
    const REAL xv = (REAL)(q[0] - 0.5 * s->dx), yv = q[1];
    F(orc_splat)(s->acc_u, s->nx, s->ny, s->dx, pool_dy, xu, yu, q[2]);
    F(orc_splat)(s->acc_v, s->nx, s->ny, s->dx, pool_dy, xv, yv, q[3]);
    F(orc_splat)(s->acc_wu, s->nx, s->ny, s->dx, pool_dy, xu, yu, (REAL)1.0);
    F(orc_splat)(s->acc_wv, s->nx, s->ny, s->dx, pool_dy, xv, yv, (REAL)1.0);
  }
  REAL *ub = U_B(s), *vb = V_B(s);
  for (size_t c = 0; c < C; ++c) {
    int nu = s->acc_wu[c] > 0.000001, nv = s->acc_wv[c] > 0.000001;
    if (s->replay_face) { /* decision replay (a zero weight cannot be normalised) */
      nu = (s->replay_face[c] & 1) && s->acc_wu[c] != 0;
      nv = (s->replay_face[c] & 2) && s->acc_wv[c] != 0;
    }
    if (nu) ub[c] = s->acc_u[c] / s->acc_wu[c];
    if (nv) vb[c] = s->acc_v[c] / s->acc_wv[c];
    if (s->record_face) s->record_face[c] = (uint8_t)(nu | (nv << 1));
  }
  swap_vel(s);
}

/* MacGrid::updatePreviousVelocityBuffer, src/MacGrid.cpp:52-56. */
static void save_prev(STATE *s) {
  const size_t C = (size_t)s->nx * s->ny;
  memcpy(s->u_prev, U_F(s), C * sizeof(REAL));
  memcpy(s->v_prev, V_F(s), C * sizeof(REAL));
}

/* MacGrid::updateVelocityDiffBuffer, src/MacGrid.cpp:58-70. */
static void save_diff(STATE *s) {
  const size_t C = (size_t)s->nx * s->ny;
  const REAL *u = U_F(s), *v = V_F(s);
  for (size_t c = 0; c < C; ++c) {
    s->u_diff[c] = u[c] - s->u_prev[c];
    s->v_diff[c] = v[c] - s->v_prev[c];
  }
}

/* FluidSolver::addExternalAcceleration, src/FluidSolver.cpp:276-295: only the left
 * and bottom faces of LIQUID cells get a*dt. */
static void gravity(STATE *s, REAL dt) {
  REAL *u = U_F(s), *v = V_F(s);
  for (int j = 0; j < s->ny; ++j)
    for (int i = 0; i < s->nx; ++i)
      if (cell(s, i, j) == ORC_LIQUID) {
        size_t c = IDX(s, i, j);
        u[c] = u[c] + s->gravity_x * dt;
        v[c] = v[c] + s->gravity_y * dt;
      }
}

/* FluidSolver::enforceDirichlet, src/FluidSolver.cpp:297-321.  closed_walls (not in
 * the reference): every face with a SOLID cell on either side is zeroed. */
static void dirichlet(STATE *s) {
  REAL *u = U_F(s), *v = V_F(s);
  /* decision bits of this pass: the first pass of a step 4 / 8, later ones 16 / 32 */
  const int k = s->dirichlet_pass > 0 ? 1 : 0;
  const uint8_t bu = (uint8_t)(4 << (2 * k)), bv = (uint8_t)(8 << (2 * k));
  s->dirichlet_pass++;
  for (int j = 0; j < s->ny; ++j)
    for (int i = 0; i < s->nx; ++i) {
      int im = clampi(i - 1, 0, s->nx - 1);
      int jm = clampi(j - 1, 0, s->ny - 1);
      size_t c = IDX(s, i, j);
      int zu, zv;
      if (s->closed_walls) {
        zu = cell(s, im, j) == ORC_SOLID || cell(s, i, j) == ORC_SOLID;
        zv = cell(s, i, jm) == ORC_SOLID || cell(s, i, j) == ORC_SOLID;
      } else {
        zu = (cell(s, im, j) == ORC_SOLID && u[c] < 0) ||
             (cell(s, i, j) == ORC_SOLID && u[c] > 0);
        zv = (cell(s, i, jm) == ORC_SOLID && v[c] < 0) ||
             (cell(s, i, j) == ORC_SOLID && v[c] > 0);
      }
      if (s->replay_face) { /* decision replay */
        zu = (s->replay_face[c] & bu) != 0;
        zv = (s->replay_face[c] & bv) != 0;
      }
      if (zu) u[c] = 0;
      if (zv) v[c] = 0;
      if (s->record_face) s->record_face[c] |= (uint8_t)((zu ? bu : 0) | (zv ? bv : 0));
    }
}

/* ------------------------------------------------------------------ extension */

/* FluidSolver::extendVelocityAvarageing, src/FluidSolver.cpp:625-707 (not called by
 * the reference's steps).  One mask per cell, the x pair (:631-643): 1 at LIQUID
 * cells in both masks; back := front for u and v everywhere.  Each sweep, in raster
 * order (j outer, i inner), a non-SOLID cell with front mask 0 averages the
 * cell-centred back velocities ((f(i)+f(i+1))/2 per axis, MacGrid::velXBackBuffer)
 * of its front-mask-1 neighbours in the order (i-1,j), (i,j-1), (i,j+1), (i+1,j),
 * writes the average to both of its faces per axis (setVelXBackBuffer: (i,j) and
 * (i+1,j); setVelYBackBuffer: (i,j) and (i,j+1)) and sets its back mask.  The
 * writes land in the buffer the sweep reads, so later cells see them.  Masks swap
 * per sweep (both pairs, :72-76), velocities at the end.  Neighbours and faces
 * outside the grid (reachable only through non-SOLID border cells, an assert in
 * the reference) are skipped. */
static void extend_average(STATE *s, int n_iter) {
  const int nx = s->nx, ny = s->ny;
  REAL *uf = U_F(s), *vf = V_F(s), *ub = U_B(s), *vb = V_B(s);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t c = IDX(s, i, j);
      uint8_t m = cell(s, i, j) == ORC_LIQUID ? 1 : 0;
      (s->mask_front ? s->mask_u1 : s->mask_u0)[c] = m;
      (s->mask_front ? s->mask_u0 : s->mask_u1)[c] = m;
      ub[c] = uf[c];
      vb[c] = vf[c];
    }
  for (int it = 0; it < n_iter; ++it) {
    const uint8_t *mf = s->mask_front ? s->mask_u1 : s->mask_u0;
    uint8_t *mb = s->mask_front ? s->mask_u0 : s->mask_u1;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t c = IDX(s, i, j);
        if (mf[c] != 0 || cell(s, i, j) == ORC_SOLID) continue;
        REAL su = 0, sv = 0;
        int n = 0;
        const int ni[4] = {i - 1, i, i, i + 1}, nj[4] = {j, j - 1, j + 1, j};
        for (int q = 0; q < 4; ++q) {
          const int a = ni[q], b = nj[q];
          if (a < 0 || a >= nx || b < 0 || b >= ny || mf[IDX(s, a, b)] != 1) continue;
          const REAL u1 = a + 1 < nx ? ub[IDX(s, a + 1, b)] : ub[IDX(s, a, b)];
          const REAL v1 = b + 1 < ny ? vb[IDX(s, a, b + 1)] : vb[IDX(s, a, b)];
          su += (ub[IDX(s, a, b)] + u1) / 2;
          sv += (vb[IDX(s, a, b)] + v1) / 2;
          n++;
        }
        if (n > 0) {
          su /= n;
          sv /= n;
          ub[c] = su;
          if (i + 1 < nx) ub[IDX(s, i + 1, j)] = su;
          vb[c] = sv;
          if (j + 1 < ny) vb[IDX(s, i, j + 1)] = sv;
          mb[c] = 1;
        }
      }
    s->mask_front ^= 1;
  }
  swap_vel(s);
}

/* FluidSolver::extendVelocityIndividual, src/FluidSolver.cpp:485-622.
 * Init (:490-530) including the :527 quirk (the v-invalid branch zeroes the front
 * *u*), then n sweeps of neighbour averaging over the back buffer with the front
 * mask deciding, back mask marking, mask swap per sweep (:619), then a velocity
 * swap (:621).  ext_info (optional) records (layer << 4) | neighbour bits with
 * neighbour bit order (i-1,j), (i,j-1), (i,j+1), (i+1,j); layer 15 = never. */
static void extend(STATE *s, int n_iter) {
  uint8_t *mxf = s->mask_front ? s->mask_u1 : s->mask_u0;
  uint8_t *mxb = s->mask_front ? s->mask_u0 : s->mask_u1;
  uint8_t *myf = s->mask_front ? s->mask_v1 : s->mask_v0;
  uint8_t *myb = s->mask_front ? s->mask_v0 : s->mask_v1;
  REAL *uf = U_F(s), *vf = V_F(s), *ub = U_B(s), *vb = V_B(s);
  const int nx = s->nx, ny = s->ny;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t c = IDX(s, i, j);
      if (cell(s, i, j) == ORC_LIQUID || cell(s, i - 1, j) == ORC_LIQUID) {
        mxf[c] = 1; mxb[c] = 1;
        ub[c] = uf[c];
        if (s->ext_info_u) s->ext_info_u[c] = 0;
      } else {
        mxf[c] = 0; mxb[c] = 0;
        ub[c] = 0;
        uf[c] = 0;
        if (s->ext_info_u) s->ext_info_u[c] = 0xF0;
      }
      if (cell(s, i, j) == ORC_LIQUID || cell(s, i, j - 1) == ORC_LIQUID) {
        myf[c] = 1; myb[c] = 1;
        vb[c] = vf[c];
        if (s->ext_info_v) s->ext_info_v[c] = 0;
      } else {
        myf[c] = 0; myb[c] = 0;
        vb[c] = 0;
        uf[c] = 0; /* src/FluidSolver.cpp:527 writes the x buffer here */
        if (s->ext_info_v) s->ext_info_v[c] = 0xF0;
      }
    }
  for (int it = 0; it < n_iter; ++it) {
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t c = IDX(s, i, j);
        if (mxf[c] == 0 && cell(s, i, j) != ORC_SOLID && cell(s, i - 1, j) != ORC_SOLID) {
          REAL acc = 0;
          int n = 0, bits = 0;
          if (mxf[IDX(s, i - 1, j)] == 1) { acc += ub[IDX(s, i - 1, j)]; n++; bits |= 1; }
          if (mxf[IDX(s, i, j - 1)] == 1) { acc += ub[IDX(s, i, j - 1)]; n++; bits |= 2; }
          if (mxf[IDX(s, i, j + 1)] == 1) { acc += ub[IDX(s, i, j + 1)]; n++; bits |= 4; }
          if (mxf[IDX(s, i + 1, j)] == 1) { acc += ub[IDX(s, i + 1, j)]; n++; bits |= 8; }
          if (n > 0) {
            acc /= (REAL)n;
            ub[c] = acc;
            mxb[c] = 1;
            if (s->ext_info_u) s->ext_info_u[c] = (uint8_t)(((it + 1) << 4) | bits);
          }
        }
        if (myf[c] == 0 && cell(s, i, j) != ORC_SOLID && cell(s, i, j - 1) != ORC_SOLID) {
          REAL acc = 0;
          int n = 0, bits = 0;
          if (myf[IDX(s, i - 1, j)] == 1) { acc += vb[IDX(s, i - 1, j)]; n++; bits |= 1; }
          if (myf[IDX(s, i, j - 1)] == 1) { acc += vb[IDX(s, i, j - 1)]; n++; bits |= 2; }
          if (myf[IDX(s, i, j + 1)] == 1) { acc += vb[IDX(s, i, j + 1)]; n++; bits |= 4; }
          if (myf[IDX(s, i + 1, j)] == 1) { acc += vb[IDX(s, i + 1, j)]; n++; bits |= 8; }
          if (n > 0) {
            acc /= (REAL)n;
            vb[c] = acc;
            myb[c] = 1;
            if (s->ext_info_v) s->ext_info_v[c] = (uint8_t)(((it + 1) << 4) | bits);
          }
        }
      }
    /* FluidSolverMemoryPool::swapValidMaskBuffer, src/FluidSolver.cpp:72-76 */
    uint8_t *t;
    t = mxf; mxf = mxb; mxb = t;
    t = myf; myf = myb; myb = t;
    s->mask_front ^= 1;
  }
  swap_vel(s);
}

/* ------------------------------------------------------------------- pressure */

/* Restatement of Eigen's ConjugateGradient (IterativeLinearSolvers, default
 * Lower + DiagonalPreconditioner), as called at src/FluidSolver.cpp:81,424-425:
 * x0 = 0; stop when |r|^2 < max(tol^2 |b|^2, min normal) or after max_iter
 * iterations.  The matrix is applied matrix-free from the labels: diagonal
 * -(#non-SOLID neighbours)/dx^2, 1/dx^2 for LIQUID neighbours
 * (src/FluidSolver.cpp:379-410; deltaX used for both axes).  Products sum the
 * diagonal first, then (i-1), (i+1), (j-1), (j+1).  Dots are sequential sums. */
typedef struct {
  int F;
  const int32_t *cell_of; /* compact index -> cell */
  const int32_t *fidx;    /* cell -> compact index or -1 */
  REAL *diag, off;
  int nx;
  const uint8_t *label;
} lap_t;

static void lap_apply(const lap_t *L, const REAL *p, REAL *out) {
  for (int k = 0; k < L->F; ++k) {
    int c = L->cell_of[k];
    REAL acc = L->diag[k] * p[k];
    int nb[4] = {c - 1, c + 1, c - L->nx, c + L->nx};
    for (int q = 0; q < 4; ++q)
      if (L->label[nb[q]] == ORC_LIQUID) acc += L->off * p[L->fidx[nb[q]]];
    out[k] = acc;
  }
}

static REAL dot(const REAL *a, const REAL *b, int n) {
  REAL s = 0;
  for (int k = 0; k < n; ++k) s += a[k] * b[k];
  return s;
}

static void cg_solve(const lap_t *L, const REAL *b, REAL *x, int max_iter, double tol_d,
                     int project_mean, int *iters_out, double *err_out) {
  const int n = L->F;
  REAL *r = (REAL *)malloc(sizeof(REAL) * n * 4 + 16);
  REAL *pd = r + n, *z = r + 2 * n, *tmp = r + 3 * n;
  REAL *invd = (REAL *)malloc(sizeof(REAL) * n + 16);
  for (int k = 0; k < n; ++k) invd[k] = L->diag[k] != 0 ? (REAL)1 / L->diag[k] : (REAL)1;
  for (int k = 0; k < n; ++k) { x[k] = 0; r[k] = b[k]; }
  if (project_mean) {
    REAL m = 0;
    for (int k = 0; k < n; ++k) m += r[k];
    m = m / (REAL)n;
    for (int k = 0; k < n; ++k) r[k] -= m;
  }
  const REAL tol = (REAL)tol_d;
  const REAL min_normal = sizeof(REAL) == 4 ? (REAL)FLT_MIN : (REAL)DBL_MIN;
  REAL rhs_norm2 = dot(r, r, n);
  if (rhs_norm2 == 0) {
    *iters_out = 0; *err_out = 0;
    free(r); free(invd);
    return;
  }
  REAL threshold = tol * tol * rhs_norm2;
  if (threshold < min_normal) threshold = min_normal;
  REAL res_norm2 = dot(r, r, n);
  if (res_norm2 < threshold) {
    *iters_out = 0; *err_out = sqrt((double)(res_norm2 / rhs_norm2));
    free(r); free(invd);
    return;
  }
  for (int k = 0; k < n; ++k) pd[k] = invd[k] * r[k];
  REAL abs_new = dot(r, pd, n);
  int i = 0;
  while (i < max_iter) {
    lap_apply(L, pd, tmp);
    REAL alpha = abs_new / dot(pd, tmp, n);
    for (int k = 0; k < n; ++k) x[k] += alpha * pd[k];
    for (int k = 0; k < n; ++k) r[k] -= alpha * tmp[k];
    res_norm2 = dot(r, r, n);
    if (res_norm2 < threshold) break;
    for (int k = 0; k < n; ++k) z[k] = invd[k] * r[k];
    REAL abs_old = abs_new;
    abs_new = dot(r, z, n);
    REAL beta = abs_new / abs_old;
    for (int k = 0; k < n; ++k) pd[k] = z[k] + beta * pd[k];
    i++;
  }
  *iters_out = i;
  *err_out = sqrt((double)(res_norm2 / rhs_norm2));
  free(r); free(invd);
}

/* FluidSolver::pressureSolve, src/FluidSolver.cpp:323-483.  Index LIQUID cells in
 * row-major order (:331-346); return without a swap when there are none
 * (:347-350); b = divVelX + divVelY (:413, unscaled); solve; write the back
 * buffer at faces where (i,j), (i-1,j) or (i,j-1) is LIQUID (:439) with
 * p = 0 for non-LIQUID cells (:455-460) and vel - dt/density * dp / dx
 * (:471-474); swap (:482).  The per-cell particle count (:352-360) feeds the term
 * k * (n > 9 ? n - 8 : 0) added to each LIQUID pressure (:441-453); the
 * reference's k is 0.  Its count index (int)((x / lengthX) * sizeX) and the
 * (i-1, j) / (i, j-1) lookups are clamped here (the reference asserts on them). */
static void pressure(STATE *s, REAL dt) {
  const int nx = s->nx, ny = s->ny;
  const size_t C = (size_t)nx * ny;
  int F_ = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t c = IDX(s, i, j);
      if (cell(s, i, j) == ORC_LIQUID) s->fluid_index[c] = F_++;
      else s->fluid_index[c] = -1;
    }
  s->n_fluid = F_;
  if (s->pressure) memset(s->pressure, 0, C * sizeof(REAL));
  if (s->rhs) memset(s->rhs, 0, C * sizeof(REAL));
  s->cg_iterations = 0;
  s->cg_error = 0;
  if (F_ == 0) return;
  int32_t *cell_of = (int32_t *)malloc(sizeof(int32_t) * F_);
  REAL *diag = (REAL *)malloc(sizeof(REAL) * F_);
  REAL *b = (REAL *)malloc(sizeof(REAL) * F_);
  REAL *x = (REAL *)malloc(sizeof(REAL) * F_);
  const REAL off = (REAL)(1 / pow((double)s->dx, 2));
  const REAL *uf = U_F(s), *vf = V_F(s);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t c = IDX(s, i, j);
      int k = s->fluid_index[c];
      if (k < 0) continue;
      cell_of[k] = (int32_t)c;
      int n_non_solid = 0;
      if (cell(s, i - 1, j) != ORC_SOLID) n_non_solid++;
      if (cell(s, i + 1, j) != ORC_SOLID) n_non_solid++;
      if (cell(s, i, j - 1) != ORC_SOLID) n_non_solid++;
      if (cell(s, i, j + 1) != ORC_SOLID) n_non_solid++;
      diag[k] = (REAL)(-n_non_solid / pow((double)s->dx, 2));
      /* MacGrid::divVelX/divVelY, include/MacGrid.h:98-111 */
      REAL dvx = (uf[IDX(s, i + 1, j)] - uf[c]) / s->dx;
      REAL dvy = (vf[IDX(s, i, j + 1)] - vf[c]) / s->dy;
      b[k] = dvx + dvy;
    }
  lap_t L = {F_, cell_of, s->fluid_index, diag, off, nx, s->label};
  int iters = 0;
  double err = 0;
  cg_solve(&L, b, x, s->cg_max_iter, s->cg_tol, s->project_mean, &iters, &err);
  s->cg_iterations = iters;
  s->cg_error = err;
  if (s->pressure)
    for (int k = 0; k < F_; ++k) s->pressure[cell_of[k]] = x[k];
  if (s->rhs)
    for (int k = 0; k < F_; ++k) s->rhs[cell_of[k]] = b[k];

  REAL *ub = U_B(s), *vb = V_B(s);
  const REAL kvol = s->particle_pressure_k;
  int32_t *npart = (int32_t *)calloc(C, sizeof(int32_t));
  for (int64_t q = 0; q < s->n_particles; ++q) {
    const REAL *pt = s->particles + 4 * q;
    int cx = (int)((pt[0] / ((REAL)nx * s->dx)) * (REAL)nx);
    int cy = (int)((pt[1] / ((REAL)ny * s->dy)) * (REAL)ny);
    npart[IDX(s, clampi(cx, 0, nx - 1), clampi(cy, 0, ny - 1))]++;
  }
#define PPRESS(c) (kvol * (REAL)(npart[c] > 9 ? npart[c] - 8 : 0))
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int im = clampi(i - 1, 0, nx - 1);
      int jm = clampi(j - 1, 0, ny - 1);
      int idx = s->fluid_index[IDX(s, i, j)];
      int idx_im = s->fluid_index[IDX(s, im, j)];
      int idx_jm = s->fluid_index[IDX(s, i, jm)];
      if (idx >= 0 || idx_im >= 0 || idx_jm >= 0) {
        REAL p = idx >= 0 ? x[idx] + PPRESS(IDX(s, i, j)) : 0;
        REAL p_im = idx_im >= 0 ? x[idx_im] + PPRESS(IDX(s, im, j)) : 0;
        REAL p_jm = idx_jm >= 0 ? x[idx_jm] + PPRESS(IDX(s, i, jm)) : 0;
        REAL pdx = p - p_im;
        REAL pdy = p - p_jm;
        size_t c = IDX(s, i, j);
        REAL vx = uf[c], vy = vf[c];
        ub[c] = vx - dt / s->density * pdx / s->dx;
        vb[c] = vy - dt / s->density * pdy / s->dy;
      }
    }
  swap_vel(s);
#undef PPRESS
  free(npart);
  free(cell_of); free(diag); free(b); free(x);
}

/* ------------------------------------------------------------------------ G2P */

/* transferVelocityToParticles{PIC,FLIP,PICFLIP}, src/FluidSolver.cpp:921-963. */
static void g2p(STATE *s, int mode) {
  const REAL *u = U_F(s), *v = V_F(s);
  const REAL r = s->pic_ratio;
  for (int64_t p = 0; p < s->n_particles; ++p) {
    REAL *q = s->particles + 4 * p;
    const REAL x = q[0], y = q[1];
    if (mode == 0) {
      REAL pu = vel_x_at(s, u, x, y);
      REAL pv = vel_y_at(s, v, x, y);
      q[2] = pu; q[3] = pv;
    } else if (mode == 1) {
      REAL fu = q[2] + vel_x_at(s, s->u_diff, x, y);
      REAL fv = q[3] + vel_y_at(s, s->v_diff, x, y);
      q[2] = fu; q[3] = fv;
    } else {
      REAL pu = vel_x_at(s, u, x, y);
      REAL pv = vel_y_at(s, v, x, y);
      REAL fu = q[2] + vel_x_at(s, s->u_diff, x, y);
      REAL fv = q[3] + vel_y_at(s, s->v_diff, x, y);
      q[2] = pu * r + fu * (1 - r);
      q[3] = pv * r + fv * (1 - r);
    }
  }
}

/* MarkerParticleSet::advect / advectAndEnsureOutsideObstacles,
 * src/MarkerParticleSet.cpp:40-62 + MarkerParticle::advect (.h:37-41).  The cell
 * test uses (int)(pos / delta), a different formula from classifyCells. */
static void advect_particles(STATE *s, REAL dt, int reject) {
  for (int64_t p = 0; p < s->n_particles; ++p) {
    REAL *q = s->particles + 4 * p;
    q[0] += q[2] * dt;
    q[1] += q[3] * dt;
    uint8_t flag = 0;
    if (reject) {
      int x = (int)(q[0] / s->dx);
      int y = (int)(q[1] / s->dy);
      int pushed = cell(s, x, y) == ORC_SOLID;
      if (s->replay_reject) pushed = s->replay_reject[p] != 0; /* decision replay */
      if (pushed) {
        q[0] += q[2] * -dt;
        q[1] += q[3] * -dt;
        flag = 1;
      }
    }
    if (s->reject_flag) s->reject_flag[p] = flag;
  }
}

/* ------------------------------------------------------------ ODE / advection */

/* FluidSolver::getAdvectedPosition (src/FluidSolver.cpp:793-814) returns
 * pos + step(f, pos, h), f = front-buffer velocity.  Steps:
 *   RK3   include/OdeSolver.h:102-113: k2 at x + k1*h*1/2, k3 at x + k2*h*3/4,
 *         (2 k1 + 3 k2 + 4 k3) * h * 1/9, in Vec2 MyFloat ops (FluidSolver.h:120-141)
 *   Euler include/OdeSolver.h:78-86: f(x + h) * h with the scalar shift of both
 *         coordinates (Vec2::operator+(MyFloat))
 *   RK2   midpoint (new; no counterpart in the reference): k2 at x + k1*h*(1/2),
 *         displacement k2*h.
 *   RK4   include/OdeSolver.h:117-142 with the Vec2 + MyFloat scalar shift its
 *         stages imply: f1 = f(x), f2 = f3 = f(x + h/2), f4 = f(x + h),
 *         (f1 + f2*2 + f3*2 + f4) * h / 6 (Vec2 * MyFloat order; the template itself
 *         does not compile with the solver's Vec2). */
static void rk_disp(const STATE *s, const REAL *u, const REAL *v, REAL x, REAL y, REAL h,
                    REAL *dx_out, REAL *dy_out) {
  const REAL one = (REAL)1.0;
  if (s->integrator == ORC_RK4) {
    const REAL hh = h / (REAL)2.0;
    REAL f1x = vel_x_at(s, u, x, y), f1y = vel_y_at(s, v, x, y);
    REAL f2x = vel_x_at(s, u, x + hh, y + hh), f2y = vel_y_at(s, v, x + hh, y + hh);
    REAL f4x = vel_x_at(s, u, x + h, y + h), f4y = vel_y_at(s, v, x + h, y + h);
    *dx_out = (f1x + f2x * (REAL)2 + f2x * (REAL)2 + f4x) * h / (REAL)6;
    *dy_out = (f1y + f2y * (REAL)2 + f2y * (REAL)2 + f4y) * h / (REAL)6;
    return;
  }
  if (s->integrator == ORC_EULER) {
    REAL px = x + h, py = y + h;
    *dx_out = vel_x_at(s, u, px, py) * h;
    *dy_out = vel_y_at(s, v, px, py) * h;
    return;
  }
  REAL k1x = vel_x_at(s, u, x, y), k1y = vel_y_at(s, v, x, y);
  if (s->integrator == ORC_RK2) {
    REAL mx = x + k1x * h * (REAL)0.5, my = y + k1y * h * (REAL)0.5;
    REAL k2x = vel_x_at(s, u, mx, my), k2y = vel_y_at(s, v, mx, my);
    *dx_out = k2x * h;
    *dy_out = k2y * h;
    return;
  }
  REAL ax = x + k1x * h * one / (REAL)2, ay = y + k1y * h * one / (REAL)2;
  REAL k2x = vel_x_at(s, u, ax, ay), k2y = vel_y_at(s, v, ax, ay);
  REAL bx = x + k2x * h * (REAL)3.0 / (REAL)4, by = y + k2y * h * (REAL)3.0 / (REAL)4;
  REAL k3x = vel_x_at(s, u, bx, by), k3y = vel_y_at(s, v, bx, by);
  *dx_out = (k1x * (REAL)2 + k2x * (REAL)3 + k3x * (REAL)4) * h * one / (REAL)9;
  *dy_out = (k1y * (REAL)2 + k2y * (REAL)3 + k3y * (REAL)4) * h * one / (REAL)9;
}

/* FluidDomain::advectLevelSet, src/FluidDomain.cpp:104-127, on phi (nx*ny, in
 * place): forward Euler, change rate -(vx gx + vy gy) times dt times 5, upwind
 * gradients of LevelSet::computeUpwindGradientX/Y (src/LevelSet.cpp:23-59, clamped
 * indices), cell-centred front velocity (u(i)+u(i+1))/2 (MacGrid::velX).  The
 * reference reads u(nx, j) / v(i, ny) out of range (an assert in its -g build):
 * clamped here.  Its loops run j over sizeX and i over sizeY; cells outside the
 * min(nx, ny) square keep the new level set's initial 0. */
void F(orc_advect_level_set)(STATE *s, REAL *phi, REAL dt) {
  const int nx = s->nx, ny = s->ny, m = nx < ny ? nx : ny;
  REAL *out = (REAL *)calloc((size_t)nx * ny, sizeof(REAL));
  const REAL *u = U_F(s), *v = V_F(s);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      const int i1 = i + 1 < nx ? i + 1 : nx - 1, j1 = j + 1 < ny ? j + 1 : ny - 1;
      const REAL vx = (u[IDX(s, i, j)] + u[IDX(s, i1, j)]) / 2;
      const REAL vy = (v[IDX(s, i, j)] + v[IDX(s, i, j1)]) / 2;
      const REAL p0 = phi[IDX(s, i, j)];
      const REAL gx = vx < 0 ? phi[IDX(s, clampi(i + 1, 0, nx - 1), j)] - p0
                             : p0 - phi[IDX(s, clampi(i - 1, 0, nx - 1), j)];
      const REAL gy = vy < 0 ? phi[IDX(s, i, clampi(j + 1, 0, ny - 1))] - p0
                             : p0 - phi[IDX(s, i, clampi(j - 1, 0, ny - 1))];
      const REAL rate = -(vx * gx + vy * gy);
      out[IDX(s, i, j)] = p0 + rate * dt * 5;
    }
  memcpy(phi, out, sizeof(REAL) * (size_t)nx * ny);
  free(out);
}

void F(orc_rk_displacement)(const STATE *s, REAL x, REAL y, REAL h, REAL *ox, REAL *oy) {
  STATE *m = (STATE *)s;
  rk_disp(s, U_F(m), V_F(m), x, y, h, ox, oy);
}

/* FluidSolver::advectVelocitySemiLagrangian, src/FluidSolver.cpp:709-772
 * (reference scatter form): back := 0; for each face next to LIQUID, backtrace
 * its position by -dt, sample the face's own velocity at its own position and
 * splat it into back at the backtraced point; no swap. */
static void sl_scatter(STATE *s, REAL dt) {
  const int nx = s->nx, ny = s->ny;
  REAL *uf = U_F(s), *vf = V_F(s), *ub = U_B(s), *vb = V_B(s);
  memset(ub, 0, sizeof(REAL) * (size_t)nx * ny);
  memset(vb, 0, sizeof(REAL) * (size_t)nx * ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (cell(s, i, j) == ORC_LIQUID || cell(s, i - 1, j) == ORC_LIQUID) {
        REAL xp = (REAL)i * s->dx;
        REAL yp = (REAL)((j + 0.5) * s->dy);
        REAL ddx, ddy;
        rk_disp(s, uf, vf, xp, yp, -dt, &ddx, &ddy);
        REAL xb = xp + ddx, yb = yp + ddy;
        REAL val = vel_x_at(s, uf, xp, yp);
        F(orc_splat)(ub, nx, ny, s->dx, s->dy, xb, (REAL)(yb - 0.5 * s->dy), val);
      }
      if (cell(s, i, j) == ORC_LIQUID || cell(s, i, j - 1) == ORC_LIQUID) {
        REAL xp = (REAL)((i + 0.5) * s->dx);
        REAL yp = (REAL)j * s->dy;
        REAL ddx, ddy;
        rk_disp(s, uf, vf, xp, yp, -dt, &ddx, &ddy);
        REAL xb = xp + ddx, yb = yp + ddy;
        REAL val = vel_y_at(s, vf, xp, yp);
        F(orc_splat)(vb, nx, ny, s->dx, s->dy, (REAL)(xb - 0.5 * s->dx), yb, val);
      }
    }
}

/* Gather form of semi-Lagrangian advection (new; the textbook method the grid-only
 * workload asks for): for each face next to LIQUID, back := front velocity
 * component sampled at the backtraced face position; other faces copy front;
 * then swap. */
static void sl_gather(STATE *s, REAL dt) {
  const int nx = s->nx, ny = s->ny;
  REAL *uf = U_F(s), *vf = V_F(s), *ub = U_B(s), *vb = V_B(s);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t c = IDX(s, i, j);
      ub[c] = uf[c];
      vb[c] = vf[c];
      if (cell(s, i, j) == ORC_LIQUID || cell(s, i - 1, j) == ORC_LIQUID) {
        REAL xp = (REAL)i * s->dx;
        REAL yp = (REAL)((j + 0.5) * s->dy);
        REAL ddx, ddy;
        rk_disp(s, uf, vf, xp, yp, -dt, &ddx, &ddy);
        ub[c] = vel_x_at(s, uf, xp + ddx, yp + ddy);
      }
      if (cell(s, i, j) == ORC_LIQUID || cell(s, i, j - 1) == ORC_LIQUID) {
        REAL xp = (REAL)((i + 0.5) * s->dx);
        REAL yp = (REAL)j * s->dy;
        REAL ddx, ddy;
        rk_disp(s, uf, vf, xp, yp, -dt, &ddx, &ddy);
        vb[c] = vel_y_at(s, vf, xp + ddx, yp + ddy);
      }
    }
  swap_vel(s);
}

/* FluidSolver::advectParticlesWithGrid, src/FluidSolver.cpp:774-791. */
static void advect_grid(STATE *s, REAL dt) {
  const REAL *uf = U_F(s), *vf = V_F(s);
  for (int64_t p = 0; p < s->n_particles; ++p) {
    REAL *q = s->particles + 4 * p;
    REAL ddx, ddy;
    rk_disp(s, uf, vf, q[0], q[1], dt, &ddx, &ddy);
    q[0] = q[0] + ddx;
    q[1] = q[1] + ddy;
  }
}

/* ------------------------------------------------------------------ dispatch */

void F(orc_defaults)(STATE *s, int32_t nx, int32_t ny, REAL length_x, REAL length_y,
                     REAL density, REAL pic_ratio) {
  memset(s, 0, sizeof(*s));
  s->nx = nx;
  s->ny = ny;
  s->dx = length_x / (REAL)nx; /* MacGrid.cpp:8 (MyFloat / int) */
  s->dy = length_y / (REAL)ny;
  s->density = density;
  s->pic_ratio = pic_ratio;
  s->gravity_x = (REAL)0;
  s->gravity_y = (REAL)-9.82;
  s->ext_layers = 2;
  s->cg_max_iter = 100;
  s->cg_tol = sizeof(REAL) == 4 ? (double)FLT_EPSILON : (double)DBL_EPSILON;
  s->integrator = ORC_RK3;
}

/* FluidSolver::validate, src/FluidSolver.cpp:89-97 (|delta difference| formed in
 * MyFloat, compared with the double 1e-7). */
int F(orc_validate)(const STATE *s, int32_t pool_nx, int32_t pool_ny, REAL pool_dx,
                    REAL pool_dy) {
  REAL ddx = s->dx - pool_dx, ddy = s->dy - pool_dy;
  if (ddx < 0) ddx = -ddx;
  if (ddy < 0) ddy = -ddy;
  return s->nx == pool_nx && s->ny == pool_ny && ddx < 0.0000001 && ddy < 0.0000001;
}

int F(orc_run_phase)(STATE *s, int32_t phase, REAL dt) {
  switch (phase) {
    case ORC_PH_CLASSIFY: classify_particles(s); return 0;
    case ORC_PH_P2G: p2g(s); return 0;
    case ORC_PH_SAVE_PREV: save_prev(s); return 0;
    case ORC_PH_GRAVITY: gravity(s, dt); return 0;
    case ORC_PH_DIRICHLET: dirichlet(s); return 0;
    case ORC_PH_EXTEND: extend(s, s->ext_layers); return 0;
    case ORC_PH_EXTEND_AVERAGE: extend_average(s, s->ext_layers); return 0;
    case ORC_PH_PRESSURE: pressure(s, dt); return 0;
    case ORC_PH_DIFF: save_diff(s); return 0;
    case ORC_PH_G2P_PIC: g2p(s, 0); return 0;
    case ORC_PH_G2P_FLIP: g2p(s, 1); return 0;
    case ORC_PH_G2P_PICFLIP: g2p(s, 2); return 0;
    case ORC_PH_ADVECT: advect_particles(s, dt, 0); return 0;
    case ORC_PH_ADVECT_REJECT: advect_particles(s, dt, 1); return 0;
    case ORC_PH_SL_SCATTER: sl_scatter(s, dt); return 0;
    case ORC_PH_SL_GATHER: sl_gather(s, dt); return 0;
    case ORC_PH_ADVECT_GRID: advect_grid(s, dt); return 0;
    case ORC_PH_CLASSIFY_LEVELSET:
      if (!s->phi) return -1;
      classify_levelset(s);
      return 0;
    default: return -1;
  }
}

/* FluidSolver::step{SemiLagrangian,PIC,FLIP,PICFLIP}, src/FluidSolver.cpp:99-251,
 * plus the grid-only gather scheme (labels from the level set, gather advection,
 * gravity, Dirichlet, projection, Dirichlet). */
int F(orc_step)(STATE *s, int32_t scheme, REAL dt) {
  s->dirichlet_pass = 0;
  if (s->record_face) memset(s->record_face, 0, (size_t)s->nx * s->ny);
  switch (scheme) {
    case ORC_SCHEME_SEMI_LAGRANGIAN:
      classify_particles(s);
      sl_scatter(s, dt);
      gravity(s, dt);
      dirichlet(s);
      pressure(s, dt);
      dirichlet(s);
      advect_grid(s, dt);
      return 0;
    case ORC_SCHEME_PIC:
      classify_particles(s);
      p2g(s);
      gravity(s, dt);
      dirichlet(s);
      extend(s, s->ext_layers);
      pressure(s, dt);
      dirichlet(s);
      g2p(s, 0);
      advect_particles(s, dt, 0);
      return 0;
    case ORC_SCHEME_FLIP:
    case ORC_SCHEME_PICFLIP:
      classify_particles(s);
      p2g(s);
      save_prev(s);
      gravity(s, dt);
      dirichlet(s);
      extend(s, s->ext_layers);
      pressure(s, dt);
      dirichlet(s);
      save_diff(s);
      if (scheme == ORC_SCHEME_FLIP) {
        g2p(s, 1);
        advect_particles(s, dt, 0);
      } else {
        g2p(s, 2);
        advect_particles(s, dt, 1);
      }
      return 0;
    case ORC_SCHEME_GRID_ADVECT:
      if (!s->phi) return -1;
      classify_levelset(s);
      sl_gather(s, dt);
      gravity(s, dt);
      dirichlet(s);
      pressure(s, dt);
      dirichlet(s);
      return 0;
    default: return -1;
  }
}
