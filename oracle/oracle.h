/*
 * oracle.h -- CPU restatement of the per-step hot path of kbladin/Fluid_Simulation.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product (libfsg.so and the C++ classes above it) never calls it.
 *
 * The same source is compiled twice: REAL=float (symbols *_f32, the reference's
 * default MyFloat, include/MathDefinitions.h:7-14) and REAL=double (symbols *_f64,
 * the USE_DOUBLE_PRECISION build used as the FP64 oracle).
 *
 * State mirrors MacGrid (include/MacGrid.h:162-176) + FluidSolverMemoryPool
 * (include/FluidSolver.h:14-43) + the particle AoS (include/MarkerParticleSet.h:50-56).
 * Buffer swaps (src/MacGrid.cpp:89-93, src/FluidSolver.cpp:72-76) are modelled by
 * toggling `front` / `mask_front` instead of swapping pointers.
 *
 * Parity status per function is listed in oracle.c next to each function.
 */
#ifndef FSG_ORACLE_H
#define FSG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LIQUID = 0, ORC_AIR = 1, ORC_SOLID = 2 };       /* include/MacGrid.h:14-17 */
enum { ORC_RK3 = 0, ORC_EULER = 1, ORC_RK2 = 2, ORC_RK4 = 3 }; /* include/OdeSolver.h:78-142 */
enum { ORC_SCHEME_SEMI_LAGRANGIAN = 0, ORC_SCHEME_PIC = 1, ORC_SCHEME_FLIP = 2,
       ORC_SCHEME_PICFLIP = 3, ORC_SCHEME_GRID_ADVECT = 4 };

/* phase ids shared with fsg_run_phase (include/fsg.h) */
enum {
  ORC_PH_CLASSIFY = 1, ORC_PH_P2G = 2, ORC_PH_SAVE_PREV = 3, ORC_PH_GRAVITY = 4,
  ORC_PH_DIRICHLET = 5, ORC_PH_EXTEND = 6, ORC_PH_PRESSURE = 7, ORC_PH_DIFF = 8,
  ORC_PH_G2P_PIC = 9, ORC_PH_G2P_FLIP = 10, ORC_PH_G2P_PICFLIP = 11,
  ORC_PH_ADVECT = 12, ORC_PH_ADVECT_REJECT = 13, ORC_PH_SL_SCATTER = 14,
  ORC_PH_SL_GATHER = 15, ORC_PH_ADVECT_GRID = 16, ORC_PH_CLASSIFY_LEVELSET = 17,
  ORC_PH_EXTEND_AVERAGE = 18
};

#define ORC_DECLARE(R, SFX)                                                         \
  typedef struct orc_state##SFX {                                                   \
    int32_t nx, ny;                                                                 \
    R dx, dy;                 /* MacGrid spacing, length/size in R (MacGrid.cpp:8) */ \
    R density, pic_ratio;                                                           \
    R gravity_x, gravity_y;   /* stepPICFLIP passes (0, -9.82) (FluidSolver.cpp:232) */ \
    int32_t ext_layers;       /* 2 (FluidSolver.cpp:234) */                         \
    int32_t cg_max_iter;      /* 100 (FluidSolver.cpp:81) */                        \
    double cg_tol;            /* Eigen default: epsilon of the scalar type */       \
    int32_t integrator;       /* ORC_RK3 (FluidSolver.h:144) */                      \
    int32_t front;            /* 0: u0/v0 is the front buffer */                    \
    int32_t project_mean;     /* remove mean of b (fully Neumann systems) */        \
    int32_t mask_front;       /* 0: mask_*0 is the front valid mask */              \
    R *u0, *u1, *v0, *v1;                                                           \
    R *u_prev, *v_prev, *u_diff, *v_diff;                                           \
    uint8_t *label;                                                                 \
    const uint8_t *solid_mask;/* optional interior solids (NULL = none) */          \
    const R *phi;             /* optional level set for CLASSIFY_LEVELSET */        \
    int64_t n_particles;                                                            \
    R *particles;             /* AoS x, y, vx, vy */                                \
    uint8_t *reject_flag;     /* optional per particle: 1 if pushed back */         \
    R *acc_u, *acc_wu, *acc_v, *acc_wv;                                             \
    uint8_t *mask_u0, *mask_u1, *mask_v0, *mask_v1;                                 \
    uint8_t *ext_info_u, *ext_info_v; /* optional: (layer<<4)|neighbour bits */     \
    int32_t *fluid_index;                                                           \
    R *pressure;              /* per cell: solved x for LIQUID cells, else 0 */     \
    R *rhs;                   /* per cell: b for LIQUID cells, else 0 */            \
    int32_t n_fluid, cg_iterations;                                                 \
    double cg_error;                                                                \
    R particle_pressure_k;    /* 0 (FluidSolver.cpp:443) */                         \
    int32_t closed_walls;     /* 1: Dirichlet zeroes every face next to SOLID */    \
    /* Decision replay (all NULL: the reference's own decisions).  Set, the next  \
     * step takes its discrete decisions from these records instead of computing  \
     * them: labels (per cell), reject flags (per particle) and per cell a byte  \
     * of face decisions: bit 0 / 1 u / v normalised by P2G, bit 2 / 3 u / v     \
     * zeroed by the first Dirichlet pass, bit 4 / 5 by the second. */          \
    const uint8_t *replay_label;                                                    \
    const uint8_t *replay_face;                                                     \
    const uint8_t *replay_reject;                                                   \
    uint8_t *record_face;     /* optional: the face decisions this step took */     \
    int32_t dirichlet_pass;   /* Dirichlet passes so far in the current step */     \
  } orc_state##SFX;                                                                 \
  void orc_defaults##SFX(orc_state##SFX *s, int32_t nx, int32_t ny, R length_x,     \
                         R length_y, R density, R pic_ratio);                       \
  int orc_run_phase##SFX(orc_state##SFX *s, int32_t phase, R dt);                   \
  int orc_step##SFX(orc_state##SFX *s, int32_t scheme, R dt);                       \
  int orc_validate##SFX(const orc_state##SFX *s, int32_t pool_nx, int32_t pool_ny,  \
                        R pool_dx, R pool_dy);                                      \
  R orc_interp##SFX(const R *grid, int32_t nx, int32_t ny, R dx, R dy, R x, R y);   \
  void orc_splat##SFX(R *grid, int32_t nx, int32_t ny, R dx, R dy, R x, R y, R val);\
  void orc_rk_displacement##SFX(const orc_state##SFX *s, R x, R y, R h, R *ox,      \
                                R *oy);                                             \
  void orc_advect_level_set##SFX(orc_state##SFX *s, R *phi, R dt);

ORC_DECLARE(float, _f32)
ORC_DECLARE(double, _f64)

#ifdef __cplusplus
}
#endif
#endif
