// MacGrid.h -- staggered MAC grid of the drop-in C++ API.
//
// Public interface of the reference's include/MacGrid.h:14-177: front/back u and
// v face grids (u(i,j) the left face of cell (i,j), v(i,j) the bottom face), prev
// and diff grids, cell labels.  The grids live on the GPU (libfsg.so); the inline
// accessors read a host mirror that is downloaded after device work and uploaded
// after host writes.  A MacGrid built on its own owns a device context; the one
// inside a FluidDomain shares the domain's.
#ifndef FSG_MAC_GRID_H
#define FSG_MAC_GRID_H

#include <memory>

#include "../../../include/fsg.h"
#include "FsgContext.h"
#include "Grid.h"
#include "MathDefinitions.h"

enum CellType { LIQUID, AIR, SOLID };  // include/MacGrid.h:14-17 (= FSG_LIQUID/AIR/SOLID)

class MacGrid : public GridInterface {
 public:
  MacGrid(int size_x, int size_y, MyFloat length_x, MyFloat length_y);
  MacGrid(std::shared_ptr<FsgContext> ctx, int size_x, int size_y, MyFloat length_x,
          MyFloat length_y);
  ~MacGrid();

  void clearCellTypeBuffer();
  void updatePreviousVelocityBuffer();
  void updateVelocityDiffBuffer();

  // Getters
  inline MyFloat velX(int i, int j) const {
    return (f(FSG_VEL_FRONT, 0, i, j) + f(FSG_VEL_FRONT, 0, i + 1, j)) / 2;
  }
  inline MyFloat velY(int i, int j) const {
    return (f(FSG_VEL_FRONT, 1, i, j) + f(FSG_VEL_FRONT, 1, i, j + 1)) / 2;
  }
  inline MyFloat velXHalfIndexed(int i, int j) const { return f(FSG_VEL_FRONT, 0, i, j); }
  inline MyFloat velYHalfIndexed(int i, int j) const { return f(FSG_VEL_FRONT, 1, i, j); }
  inline MyFloat velXBackBufferHalfIndexed(int i, int j) const { return f(FSG_VEL_BACK, 0, i, j); }
  inline MyFloat velYBackBufferHalfIndexed(int i, int j) const { return f(FSG_VEL_BACK, 1, i, j); }
  inline MyFloat velXBackBuffer(int i, int j) const {
    return (f(FSG_VEL_BACK, 0, i, j) + f(FSG_VEL_BACK, 0, i + 1, j)) / 2;
  }
  inline MyFloat velYBackBuffer(int i, int j) const {
    return (f(FSG_VEL_BACK, 1, i, j) + f(FSG_VEL_BACK, 1, i, j + 1)) / 2;
  }
  inline MyFloat velXInterpolated(MyFloat x, MyFloat y) const {
    return sample(FSG_VEL_FRONT, 0, x, y - _DELTA_Y * 0.5);
  }
  inline MyFloat velYInterpolated(MyFloat x, MyFloat y) const {
    return sample(FSG_VEL_FRONT, 1, x - _DELTA_X * 0.5, y);
  }
  inline MyFloat velXDiffInterpolated(MyFloat x, MyFloat y) const {
    return sample(FSG_VEL_DIFF, 0, x, y - _DELTA_Y * 0.5);
  }
  inline MyFloat velYDiffInterpolated(MyFloat x, MyFloat y) const {
    return sample(FSG_VEL_DIFF, 1, x - _DELTA_X * 0.5, y);
  }
  inline CellType cellType(int i, int j) const {
    i = CLAMP(i, 0, _SIZE_X - 1);
    j = CLAMP(j, 0, _SIZE_Y - 1);
    return static_cast<CellType>(_ctx->grids().labels[twoDToLinear(i, j)]);
  }
  inline MyFloat divVelX(int i, int j) const {
    return (f(FSG_VEL_FRONT, 0, i + 1, j) - f(FSG_VEL_FRONT, 0, i, j)) / _DELTA_X;
  }
  inline MyFloat divVelY(int i, int j) const {
    return (f(FSG_VEL_FRONT, 1, i, j + 1) - f(FSG_VEL_FRONT, 1, i, j)) / _DELTA_Y;
  }

  // Setters
  inline void setVelXHalfIndexed(int i, int j, MyFloat vel_x) { w(FSG_VEL_FRONT, 0, i, j) = vel_x; }
  inline void setVelYHalfIndexed(int i, int j, MyFloat vel_y) { w(FSG_VEL_FRONT, 1, i, j) = vel_y; }
  inline void setVelXBackBuffer(int i, int j, MyFloat vel_x) {
    w(FSG_VEL_BACK, 0, i, j) = vel_x;
    w(FSG_VEL_BACK, 0, i + 1, j) = vel_x;
  }
  inline void setVelYBackBuffer(int i, int j, MyFloat vel_y) {
    w(FSG_VEL_BACK, 1, i, j) = vel_y;
    w(FSG_VEL_BACK, 1, i, j + 1) = vel_y;
  }
  inline void setVelXBackBufferHalfIndexed(int i, int j, MyFloat vel_x) {
    w(FSG_VEL_BACK, 0, i, j) = vel_x;
  }
  inline void setVelYBackBufferHalfIndexed(int i, int j, MyFloat vel_y) {
    w(FSG_VEL_BACK, 1, i, j) = vel_y;
  }
  inline void setCellType(int i, int j, CellType cell_type) {
    _ctx->gridsForWrite().labels[twoDToLinear(i, j)] = static_cast<uint8_t>(cell_type);
  }
  inline void addToVelXInterpolated(MyFloat x, MyFloat y, MyFloat vel_x) {
    fsgSplatBilinear(_ctx->gridsForWrite().u[FSG_VEL_BACK].data(), _SIZE_X, _SIZE_Y, _DELTA_X,
                     _DELTA_Y, x, y - 0.5 * _DELTA_Y, vel_x);
  }
  inline void addToVelYInterpolated(MyFloat x, MyFloat y, MyFloat vel_y) {
    fsgSplatBilinear(_ctx->gridsForWrite().v[FSG_VEL_BACK].data(), _SIZE_X, _SIZE_Y, _DELTA_X,
                     _DELTA_Y, x - 0.5 * _DELTA_X, y, vel_y);
  }

  void swapVelocityBuffers();

  FsgContext& context() const { return *_ctx; }

 private:
  inline MyFloat f(int which, int comp, int i, int j) const {
    const FsgContext::GridMirror& g = _ctx->grids();
    return (comp ? g.v[which] : g.u[which])[twoDToLinear(i, j)];
  }
  inline float& w(int which, int comp, int i, int j) {
    FsgContext::GridMirror& g = _ctx->gridsForWrite();
    return (comp ? g.v[which] : g.u[which])[twoDToLinear(i, j)];
  }
  inline MyFloat sample(int which, int comp, MyFloat x, MyFloat y) const {
    const FsgContext::GridMirror& g = _ctx->grids();
    const float* d = (comp ? g.v[which] : g.u[which]).data();
    return fsgSampleBilinear(d, _SIZE_X, _SIZE_Y, _DELTA_X, _DELTA_Y, x, y);
  }

  std::shared_ptr<FsgContext> _ctx;
};

#endif
