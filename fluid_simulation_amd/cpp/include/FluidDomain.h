// FluidDomain.h -- simulation domain of the drop-in C++ API.
//
// Public interface of the reference's include/FluidDomain.h:10-76.  The domain
// owns one device context (libfsg.so handle) shared by its MacGrid and its
// MarkerParticleSet; FluidSource emission and the level set stay on the host.
#ifndef FSG_FLUID_DOMAIN_H
#define FSG_FLUID_DOMAIN_H

#include <memory>
#include <vector>

#include "FsgContext.h"
#include "LevelSet.h"
#include "MacGrid.h"
#include "MarkerParticleSet.h"
#include "MathDefinitions.h"

class FluidSource {
 public:
  FluidSource(BBox<MyFloat> area, MyFloat delta_x, MyFloat delta_y, MyFloat x_velocity,
              MyFloat y_velocity, MyFloat time_step, int max_spawns);
  ~FluidSource();
  void update(MarkerParticleSet& particle_set, MyFloat dt);
  inline bool isFinished() { return _max_spawns != -1 && _n_spawns >= _max_spawns; }
  inline void resetSpawns() { _n_spawns = 0; }

 private:
  BBox<MyFloat> _area;
  MyFloat _x_velocity;
  MyFloat _y_velocity;
  MyFloat _delta_x;
  MyFloat _delta_y;
  MyFloat _time_step;
  MyFloat _time_since_last;
  int _max_spawns;
  int _n_spawns;
};

class FluidDomain : public GridInterface {
 public:
  FluidDomain(int size_x, int size_y, MyFloat length_x, MyFloat length_y, MyFloat density,
              MyFloat pic_ratio);
  ~FluidDomain();

  void addFluidSource(FluidSource fluid_source);
  void update(MyFloat dt);
  void clearFluidSources();
  void resetParticleSet();
  void setPicRatio(MyFloat pic_ratio);

  MacGrid& macGrid();
  LevelSet& levelSet();
  MarkerParticleSet& markerParticleSet();
  const MacGrid& macGrid() const;
  const LevelSet& levelSet() const;
  const MarkerParticleSet& markerParticleSet() const;
  const MyFloat density() const;
  const MyFloat picRatio() const;

  void advectLevelSet(MyFloat dt);

  void classifyCells(LevelSet& levelSet);
  void classifyCells(MarkerParticleSet& particle_set);

  // the device context (libfsg.so handle) of this domain
  FsgContext& context() const { return *_ctx; }

 private:
  MyFloat _density;
  MyFloat _pic_ratio;
  std::shared_ptr<FsgContext> _ctx;
  MacGrid _mac_grid;
  LevelSet _level_set;
  MarkerParticleSet _particle_set;
  std::vector<FluidSource> _fluid_sources;
};

#endif
