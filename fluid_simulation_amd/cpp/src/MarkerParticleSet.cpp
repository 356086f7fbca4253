// MarkerParticleSet.cpp -- particles on the device when attached to a domain.
#include "MarkerParticleSet.h"

#include "FsgContext.h"
#include "MacGrid.h"

MarkerParticleSet::MarkerParticleSet(int size) { _particles.resize(size); }

MarkerParticleSet::~MarkerParticleSet() {}

// a copy is a detached host set with the same particles
MarkerParticleSet::MarkerParticleSet(const MarkerParticleSet& other)
    : _particles(other.hostRead()), _ctx(nullptr) {}

MarkerParticleSet& MarkerParticleSet::operator=(const MarkerParticleSet& other) {
  if (this != &other) hostWrite() = other.hostRead();
  return *this;
}

void MarkerParticleSet::attach(FsgContext* ctx) {
  _ctx = ctx;
  if (_ctx) _ctx->attachParticles(&_particles);
}

std::vector<MarkerParticle>& MarkerParticleSet::hostRead() const {
  if (_ctx) _ctx->particles();
  return _particles;
}

std::vector<MarkerParticle>& MarkerParticleSet::hostWrite() {
  if (_ctx) _ctx->particlesForWrite();
  return _particles;
}

void MarkerParticleSet::addParticle(MarkerParticle p) { hostWrite().push_back(p); }

void MarkerParticleSet::addLattice(const std::vector<MyFloat>& xs, const std::vector<MyFloat>& ys,
                                   MyFloat vx, MyFloat vy) {
  if (_ctx) {
    _ctx->beforeDevice();
    _ctx->check(fsg_emit_lattice(_ctx->handle(), xs.data(), (int32_t)xs.size(), ys.data(),
                                 (int32_t)ys.size(), vx, vy),
                "fsg_emit_lattice");
    _ctx->afterDevice();
    return;
  }
  for (MyFloat y : ys)
    for (MyFloat x : xs) _particles.push_back(MarkerParticle(x, y, vx, vy));
}
void MarkerParticleSet::reserve(int particle_count) { _particles.reserve(particle_count); }
void MarkerParticleSet::clear() { hostWrite().clear(); }

int MarkerParticleSet::size() const { return (int)hostRead().size(); }

MarkerParticleSet::iterator MarkerParticleSet::begin() { return hostWrite().begin(); }
MarkerParticleSet::iterator MarkerParticleSet::end() { return hostWrite().end(); }
MarkerParticleSet::const_iterator MarkerParticleSet::begin() const { return hostRead().begin(); }
MarkerParticleSet::const_iterator MarkerParticleSet::end() const { return hostRead().end(); }

// MarkerParticleSet::advect (pos += vel * dt)
void MarkerParticleSet::advect(MyFloat dt) {
  if (_ctx) {
    _ctx->beforeDevice();
    _ctx->check(fsg_run_phase(_ctx->handle(), FSG_PH_ADVECT, dt), "fsg_run_phase(ADVECT)");
    _ctx->afterDevice();
    return;
  }
  for (MarkerParticle& p : _particles) p.advect(dt);
}

// advect, then step back out of SOLID cells, tested at (int)(pos / delta).  An
// attached set tests the labels of its own domain's grid on the GPU.
void MarkerParticleSet::advectAndEnsureOutsideObstacles(MyFloat dt, const MacGrid& mac_grid) {
  if (_ctx) {
    _ctx->beforeDevice();
    _ctx->check(fsg_run_phase(_ctx->handle(), FSG_PH_ADVECT_REJECT, dt),
                "fsg_run_phase(ADVECT_REJECT)");
    _ctx->afterDevice();
    return;
  }
  for (MarkerParticle& p : _particles) {
    p.advect(dt);
    const int x = p.posX() / mac_grid.deltaX();
    const int y = p.posY() / mac_grid.deltaY();
    if (mac_grid.cellType(x, y) == SOLID) p.advect(-dt);
  }
}
