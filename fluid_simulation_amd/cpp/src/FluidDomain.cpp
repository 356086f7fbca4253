// FluidDomain.cpp -- domain with device-resident state (drop-in FluidDomain).
#include "FluidDomain.h"

FluidSource::FluidSource(BBox<MyFloat> area, MyFloat delta_x, MyFloat delta_y, MyFloat x_velocity,
                         MyFloat y_velocity, MyFloat time_step, int max_spawns)
    : _area(area),
      _x_velocity(x_velocity),
      _y_velocity(y_velocity),
      _delta_x(delta_x),
      _delta_y(delta_y),
      _time_step(time_step),
      _time_since_last(0.0),
      _max_spawns(max_spawns),
      _n_spawns(0) {}

FluidSource::~FluidSource() {}

// Every time_step seconds, emit a lattice of particles over the area with spacing
// delta / 2.5 (about 6.25 per cell), rows in y, columns in x, accumulated in
// MyFloat -- the reference emitter's lattice (src/FluidDomain.cpp:29-52).  The two
// axes are accumulated here (each value is a chain of float additions); the
// lattice itself is appended on the GPU when the set belongs to a domain.
void FluidSource::update(MarkerParticleSet& particle_set, MyFloat dt) {
  if (isFinished()) return;
  if (_time_since_last >= _time_step) {
    const MyFloat x_step = _delta_x / 2.5;
    const MyFloat y_step = _delta_y / 2.5;
    std::vector<MyFloat> xs, ys;
    for (MyFloat x = _area.x_min; x < _area.x_max; x += x_step) xs.push_back(x);
    for (MyFloat y = _area.y_min; y < _area.y_max; y += y_step) ys.push_back(y);
    particle_set.addLattice(xs, ys, _x_velocity, _y_velocity);
    _time_since_last = 0;
    _n_spawns++;
  }
  _time_since_last += dt;
}

FluidDomain::FluidDomain(int size_x, int size_y, MyFloat length_x, MyFloat length_y,
                         MyFloat density, MyFloat pic_ratio)
    : GridInterface(size_x, size_y, length_x / size_x, length_y / size_y),
      _density(density),
      _pic_ratio(pic_ratio),
      _ctx(std::make_shared<FsgContext>(size_x, size_y, length_x, length_y, density, pic_ratio)),
      _mac_grid(_ctx, size_x, size_y, length_x, length_y),
      _level_set(size_x, size_y, length_x, length_y) {
  _particle_set.attach(_ctx.get());
}

FluidDomain::~FluidDomain() { _particle_set.attach(nullptr); }

void FluidDomain::addFluidSource(FluidSource fluid_source) { _fluid_sources.push_back(fluid_source); }

void FluidDomain::update(MyFloat dt) {
  for (size_t i = 0; i < _fluid_sources.size(); ++i) _fluid_sources[i].update(_particle_set, dt);
}

void FluidDomain::clearFluidSources() { _fluid_sources.clear(); }

void FluidDomain::resetParticleSet() { _particle_set.clear(); }

void FluidDomain::setPicRatio(MyFloat pic_ratio) {
  _pic_ratio = CLAMP(pic_ratio, 0, 1);
  _ctx->check(fsg_set_pic_ratio(_ctx->handle(), _pic_ratio), "fsg_set_pic_ratio");
}

MacGrid& FluidDomain::macGrid() { return _mac_grid; }
LevelSet& FluidDomain::levelSet() { return _level_set; }
MarkerParticleSet& FluidDomain::markerParticleSet() { return _particle_set; }
const MacGrid& FluidDomain::macGrid() const { return _mac_grid; }
const LevelSet& FluidDomain::levelSet() const { return _level_set; }
const MarkerParticleSet& FluidDomain::markerParticleSet() const { return _particle_set; }
const MyFloat FluidDomain::density() const { return _density; }
const MyFloat FluidDomain::picRatio() const { return _pic_ratio; }

// FluidDomain::advectLevelSet (src/FluidDomain.cpp:104-127) on the GPU with the
// front velocity: forward Euler with the reference's factor 5 and first-order
// upwind gradients.  The face past the last column / row, which the reference
// reads out of range (an assert in its -g build), is clamped; on a non-square grid
// only the min(nx, ny) square its swapped loop bounds reach is advected and the
// rest becomes 0, as in the reference's fresh level set.
void FluidDomain::advectLevelSet(MyFloat dt) {
  _ctx->beforeDevice();
  _ctx->check(fsg_set_level_set(_ctx->handle(), _level_set.raw()), "fsg_set_level_set");
  _ctx->check(fsg_advect_level_set(_ctx->handle(), dt), "fsg_advect_level_set");
  _ctx->check(fsg_get_level_set(_ctx->handle(), _level_set.raw()), "fsg_get_level_set");
}

void FluidDomain::classifyCells(LevelSet& level_set) {
  _ctx->beforeDevice();
  _ctx->check(fsg_set_level_set(_ctx->handle(), level_set.raw()), "fsg_set_level_set");
  _ctx->check(fsg_run_phase(_ctx->handle(), FSG_PH_CLASSIFY_LEVELSET, 0.f),
              "fsg_run_phase(CLASSIFY_LEVELSET)");
  _ctx->afterDevice();
}

// Like the reference, labels from the domain's own particle set whatever set is
// passed (src/FluidDomain.cpp:157).
void FluidDomain::classifyCells(MarkerParticleSet&) {
  _ctx->beforeDevice();
  _ctx->check(fsg_run_phase(_ctx->handle(), FSG_PH_CLASSIFY, 0.f), "fsg_run_phase(CLASSIFY)");
  _ctx->afterDevice();
}
