// FluidSolver.h -- the solver of the drop-in C++ API.
//
// Public interface of the reference's include/FluidSolver.h:14-54:
// FluidSolverMemoryPool (sizes and spacing the solver was built for) and
// FluidSolver::step{SemiLagrangian,PIC,FLIP,PICFLIP}(FluidDomain&, MyFloat dt).
// Each step is one fsg_step on the domain's GPU context.  The pool's scratch
// grids of the reference (fluid indices, valid masks, P2G sums) are device buffers
// of libfsg.so here and are not exposed.
//
// Additions (no counterpart in the reference): setParams()/params() select the
// pressure-solver form and tolerances (fsg_params); the defaults reproduce the
// reference's hard-coded values (gravity -9.82, 2 extension layers, CG at most 100
// iterations with Eigen's default tolerance).
#ifndef FSG_FLUID_SOLVER_H
#define FSG_FLUID_SOLVER_H

#include "../../../include/fsg.h"
#include "FluidDomain.h"
#include "MathDefinitions.h"

class FluidSolverMemoryPool : public GridInterface {
 public:
  FluidSolverMemoryPool(int size_x, int size_y, MyFloat delta_x, MyFloat delta_y);
  FluidSolverMemoryPool(const FluidDomain& fluid_domain);
  FluidSolverMemoryPool(const FluidSolverMemoryPool& other);
  ~FluidSolverMemoryPool();

  // the valid-mask pair lives on the device and is swapped there
  void swapValidMaskBuffer() {}
};

class FluidSolver {
 public:
  FluidSolver(FluidSolverMemoryPool mem_pool);
  ~FluidSolver();

  void stepSemiLagrangian(FluidDomain& fluid_domain, MyFloat dt);
  void stepPIC(FluidDomain& fluid_domain, MyFloat dt);
  void stepFLIP(FluidDomain& fluid_domain, MyFloat dt);
  void stepPICFLIP(FluidDomain& fluid_domain, MyFloat dt);

  void setParams(const fsg_params& params);
  const fsg_params& params() const { return _params; }

 private:
  bool validate(const FluidDomain& fluid_domain);
  void step(FluidDomain& fluid_domain, int scheme, MyFloat dt);

  FluidSolverMemoryPool _mem_pool;
  fsg_params _params;
};

#endif
