// FsgContext.cpp -- device handle + lazily synchronised host mirrors.
#include "FsgContext.h"

#include "MarkerParticleSet.h"

FsgContext::FsgContext(int nx, int ny, MyFloat length_x, MyFloat length_y, MyFloat density,
                       MyFloat pic_ratio)
    : nx_(nx), ny_(ny), p_(nullptr) {
  const int rc = fsg_create(&h_, nx, ny, length_x, length_y, density, pic_ratio);
  if (rc != FSG_OK)
    throw std::runtime_error("fsg_create failed (" + std::to_string(rc) +
                             "): is a gfx950 GPU visible?");
  const size_t C = (size_t)nx * ny;
  for (int q = 0; q < 4; ++q) {
    g_.u[q].assign(C, 0.f);
    g_.v[q].assign(C, 0.f);
  }
  g_.labels.assign(C, 0);
  g_fresh_ = false;  // labels were initialised on the device
}

FsgContext::~FsgContext() {
  if (h_) fsg_destroy(h_);
}

void FsgContext::check(int rc, const char* what) const {
  if (rc == FSG_OK) return;
  const char* msg = fsg_last_error(h_);
  // FluidSolver::validate's failure keeps the reference's message
  if (rc == FSG_E_VALIDATE) throw std::runtime_error(msg ? msg : what);
  throw std::runtime_error(std::string(what) + " failed: " + (msg ? msg : ""));
}

const FsgContext::GridMirror& FsgContext::grids() {
  if (!g_fresh_) {
    for (int q = 0; q < 4; ++q)
      check(fsg_get_velocity(h_, q, g_.u[q].data(), g_.v[q].data()), "fsg_get_velocity");
    check(fsg_get_labels(h_, g_.labels.data()), "fsg_get_labels");
    g_fresh_ = true;
  }
  return g_;
}

FsgContext::GridMirror& FsgContext::gridsForWrite() {
  grids();
  g_dirty_ = true;
  return g_;
}

const std::vector<MarkerParticle>& FsgContext::particles() {
  if (!p_) throw std::logic_error("FsgContext: no particle set attached");
  if (!p_fresh_) {
    int64_t n = 0;
    check(fsg_num_particles(h_, &n), "fsg_num_particles");
    p_->resize((size_t)n);
    if (n > 0)
      check(fsg_get_particles(h_, reinterpret_cast<float*>(p_->data()), n), "fsg_get_particles");
    p_fresh_ = true;
  }
  return *p_;
}

std::vector<MarkerParticle>& FsgContext::particlesForWrite() {
  particles();
  p_dirty_ = true;
  return *p_;
}

void FsgContext::attachParticles(std::vector<MarkerParticle>* v) {
  p_ = v;
  p_dirty_ = true;  // whatever the host set holds becomes the device state
  p_fresh_ = true;
}

void FsgContext::beforeDevice() {
  if (g_dirty_) {
    for (int q = 0; q < 4; ++q)
      check(fsg_set_velocity(h_, q, g_.u[q].data(), g_.v[q].data()), "fsg_set_velocity");
    check(fsg_set_labels(h_, g_.labels.data()), "fsg_set_labels");
    g_dirty_ = false;
  }
  if (p_ && p_dirty_) {
    check(fsg_set_particles(h_, reinterpret_cast<const float*>(p_->data()), (int64_t)p_->size()),
          "fsg_set_particles");
    p_dirty_ = false;
  }
}

void FsgContext::afterDevice() {
  g_fresh_ = false;
  p_fresh_ = false;
}

void FsgContext::swapVelocity() {
  beforeDevice();  // host writes to the front buffer go up before the flip
  // keep a fresh mirror fresh: the device swap is an index flip, mirror it
  if (g_fresh_) {
    g_.u[FSG_VEL_FRONT].swap(g_.u[FSG_VEL_BACK]);
    g_.v[FSG_VEL_FRONT].swap(g_.v[FSG_VEL_BACK]);
  }
  check(fsg_swap_velocity(h_), "fsg_swap_velocity");
}
