// FluidSolver.cpp -- each step is one fsg_step on the domain's GPU context.
#include "FluidSolver.h"

#include <cmath>
#include <stdexcept>

namespace {
const char kValidateMsg[] =
    "Memory pool and fluid domain does not match. Initialize the fluid solver with a "
    "memory pool corresponding to the right fluid domain!";
}

FluidSolverMemoryPool::FluidSolverMemoryPool(int size_x, int size_y, MyFloat delta_x,
                                             MyFloat delta_y)
    : GridInterface(size_x, size_y, delta_x, delta_y) {}

FluidSolverMemoryPool::FluidSolverMemoryPool(const FluidDomain& fluid_domain)
    : GridInterface(fluid_domain.sizeX(), fluid_domain.sizeY(), fluid_domain.deltaX(),
                    fluid_domain.deltaY()) {}

// The reference's copy constructor passes deltaX for both spacings
// (src/FluidSolver.cpp:56-63); FluidSolver stores its pool by copy, so a domain
// with deltaX != deltaY fails validate() there.  Kept for the same behaviour.
FluidSolverMemoryPool::FluidSolverMemoryPool(const FluidSolverMemoryPool& other)
    : GridInterface(other.sizeX(), other.sizeY(), other.deltaX(), other.deltaX()) {}

FluidSolverMemoryPool::~FluidSolverMemoryPool() {}

FluidSolver::FluidSolver(FluidSolverMemoryPool mem_pool) : _mem_pool(mem_pool) {
  fsg_params_default(&_params);
}

FluidSolver::~FluidSolver() {}

void FluidSolver::setParams(const fsg_params& params) { _params = params; }

bool FluidSolver::validate(const FluidDomain& fluid_domain) {
  return fluid_domain.sizeX() == _mem_pool.sizeX() && fluid_domain.sizeY() == _mem_pool.sizeY() &&
         std::abs(fluid_domain.deltaX() - _mem_pool.deltaX()) < 0.0000001 &&
         std::abs(fluid_domain.deltaY() - _mem_pool.deltaY()) < 0.0000001;
}

void FluidSolver::step(FluidDomain& fluid_domain, int scheme, MyFloat dt) {
  if (!validate(fluid_domain)) throw std::runtime_error(kValidateMsg);
  FsgContext& ctx = fluid_domain.context();
  ctx.check(fsg_set_params(ctx.handle(), &_params), "fsg_set_params");
  ctx.beforeDevice();
  const int rc = fsg_step(ctx.handle(), scheme, dt);
  ctx.afterDevice();
  ctx.check(rc, "fsg_step");
}

void FluidSolver::stepSemiLagrangian(FluidDomain& fluid_domain, MyFloat dt) {
  step(fluid_domain, FSG_SCHEME_SEMI_LAGRANGIAN, dt);
}
void FluidSolver::stepPIC(FluidDomain& fluid_domain, MyFloat dt) {
  step(fluid_domain, FSG_SCHEME_PIC, dt);
}
void FluidSolver::stepFLIP(FluidDomain& fluid_domain, MyFloat dt) {
  step(fluid_domain, FSG_SCHEME_FLIP, dt);
}
void FluidSolver::stepPICFLIP(FluidDomain& fluid_domain, MyFloat dt) {
  step(fluid_domain, FSG_SCHEME_PICFLIP, dt);
}
