// MarkerParticleSet.h -- marker particles of the drop-in C++ API.
//
// Public interface of the reference's include/MarkerParticleSet.h:9-78.  A set
// owned by a FluidDomain is attached to the domain's device context: its vector is
// the host mirror of the GPU particle array (AoS x, y, vx, vy, 16 bytes each, the
// reference's own layout), and advect / advectAndEnsureOutsideObstacles run on
// the GPU.  A free-standing set is plain host storage with host advection.
#ifndef FSG_MARKER_PARTICLE_SET_H
#define FSG_MARKER_PARTICLE_SET_H

#include <vector>

#include "MathDefinitions.h"

class MacGrid;
class FsgContext;

class MarkerParticle {
 public:
  MarkerParticle() : _pos_x(0), _pos_y(0), _vel_x(0), _vel_y(0) {}
  MarkerParticle(MyFloat pos_x, MyFloat pos_y, MyFloat vel_x = 0, MyFloat vel_y = 0)
      : _pos_x(pos_x), _pos_y(pos_y), _vel_x(vel_x), _vel_y(vel_y) {}
  ~MarkerParticle() {}

  inline MyFloat posX() const { return _pos_x; }
  inline MyFloat posY() const { return _pos_y; }
  inline MyFloat velX() const { return _vel_x; }
  inline MyFloat velY() const { return _vel_y; }
  inline void setPosition(MyFloat pos_x, MyFloat pos_y) {
    _pos_x = pos_x;
    _pos_y = pos_y;
  }
  inline void setVelocity(MyFloat vel_x, MyFloat vel_y) {
    _vel_x = vel_x;
    _vel_y = vel_y;
  }
  inline void advect(MyFloat dt) {
    _pos_x += _vel_x * dt;
    _pos_y += _vel_y * dt;
  }

 private:
  MyFloat _pos_x;
  MyFloat _pos_y;
  MyFloat _vel_x;
  MyFloat _vel_y;
};

static_assert(sizeof(MarkerParticle) == 4 * sizeof(float), "AoS layout of the C ABI");

class MarkerParticleSet {
 public:
  MarkerParticleSet(int size = 0);
  ~MarkerParticleSet();
  MarkerParticleSet(const MarkerParticleSet& other);
  MarkerParticleSet& operator=(const MarkerParticleSet& other);

  void addParticle(MarkerParticle p);
  void reserve(int particle_count);
  void clear();
  void advect(MyFloat dt);
  void advectAndEnsureOutsideObstacles(MyFloat dt, const MacGrid& mac_grid);

  typedef std::vector<MarkerParticle>::iterator iterator;
  typedef std::vector<MarkerParticle>::const_iterator const_iterator;

  int size() const;
  // mutable iteration may change particles: the mirror is uploaded before the
  // next device operation
  iterator begin();
  iterator end();
  const_iterator begin() const;
  const_iterator end() const;

  // FluidDomain attaches its set to its device context
  void attach(FsgContext* ctx);
  // Appends the lattice (xs[i], ys[j], vx, vy), j outer, i inner (FluidSource's
  // emission order); on the GPU when attached.
  void addLattice(const std::vector<MyFloat>& xs, const std::vector<MyFloat>& ys, MyFloat vx,
                  MyFloat vy);

 private:
  std::vector<MarkerParticle>& hostRead() const;
  std::vector<MarkerParticle>& hostWrite();

  mutable std::vector<MarkerParticle> _particles;
  FsgContext* _ctx = nullptr;
};

#endif
