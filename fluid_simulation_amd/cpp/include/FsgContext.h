// FsgContext.h -- the device state behind one FluidDomain and the host mirrors
// that serve the reference API's inline per-element accessors.
//
// The GPU copy is the source of truth.  A host mirror is downloaded on the first
// access after a device operation (one transfer of every grid, or of the
// particles) and, once written through a setter, is uploaded before the next
// device operation.  Not thread-safe, like the reference (FluidSolver.h:117).
#ifndef FSG_CONTEXT_H
#define FSG_CONTEXT_H

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/fsg.h"
#include "MathDefinitions.h"

class MarkerParticle;

class FsgContext {
 public:
  FsgContext(int nx, int ny, MyFloat length_x, MyFloat length_y, MyFloat density,
             MyFloat pic_ratio);
  ~FsgContext();
  FsgContext(const FsgContext&) = delete;
  FsgContext& operator=(const FsgContext&) = delete;

  fsg_domain* handle() const { return h_; }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  void check(int rc, const char* what) const;

  // grid mirror: front/back/prev/diff u and v + labels
  struct GridMirror {
    std::vector<float> u[4], v[4];  // indexed by FSG_VEL_FRONT .. FSG_VEL_DIFF
    std::vector<uint8_t> labels;
  };
  const GridMirror& grids();              // host-fresh view
  GridMirror& gridsForWrite();            // host-fresh view, marked dirty
  // particle mirror (AoS x, y, vx, vy)
  const std::vector<MarkerParticle>& particles();
  std::vector<MarkerParticle>& particlesForWrite();

  // the particle vector of the owning domain's MarkerParticleSet becomes the mirror
  void attachParticles(std::vector<MarkerParticle>* v);
  // MacGrid::swapVelocityBuffers on both sides
  void swapVelocity();

  // around every device operation
  void beforeDevice();
  void afterDevice();

 private:
  fsg_domain* h_ = nullptr;
  int nx_, ny_;
  GridMirror g_;
  bool g_fresh_ = true, g_dirty_ = false;
  std::vector<MarkerParticle>* p_;
  bool p_fresh_ = true, p_dirty_ = false;
};

#endif
