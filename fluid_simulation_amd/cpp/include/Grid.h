// Grid.h -- host-side dense 2-D grids of the drop-in C++ API.
//
// Public interface of the reference's include/Grid.h:11-190 (GridInterface,
// Grid<T>, BBox<T>): row-major storage i + j*size_x, world <-> cell maps, bilinear
// sampling with truncating casts (fraction taken before clamping) and bilinear
// splatting.  These host grids serve the level set and the host mirrors of the
// device state; the solver's own grids live on the GPU.
#ifndef FSG_GRID_H
#define FSG_GRID_H

#include <cassert>
#include <vector>

#include "MathDefinitions.h"

class GridInterface {
 public:
  GridInterface(int size_x, int size_y, MyFloat delta_x = 1, MyFloat delta_y = 1)
      : _SIZE_X(size_x), _SIZE_Y(size_y), _DELTA_X(delta_x), _DELTA_Y(delta_y) {}
  ~GridInterface() {}

  inline void linearTo2D(int idx, int* i, int* j) const {
    *i = idx % _SIZE_X;
    *j = idx / _SIZE_X;
  }
  inline int twoDToLinear(int i, int j) const {
    assert(indexIsValid(i, j));
    return i + j * _SIZE_X;
  }
  inline void worldToCell(MyFloat x, MyFloat y, int* i, int* j) const {
    *i = x / _DELTA_X;
    *j = y / _DELTA_Y;
  }
  inline void cellToWorld(int i, int j, MyFloat* x, MyFloat* y) const {
    assert(indexIsValid(i, j));
    *x = i * _DELTA_X;
    *y = j * _DELTA_Y;
  }
  inline bool indexIsValid(int i, int j) const {
    return i >= 0 && i < _SIZE_X && j >= 0 && j < _SIZE_Y;
  }
  inline int sizeX() const { return _SIZE_X; }
  inline int sizeY() const { return _SIZE_Y; }
  inline MyFloat deltaX() const { return _DELTA_X; }
  inline MyFloat deltaY() const { return _DELTA_Y; }
  inline MyFloat lengthX() const { return _SIZE_X * _DELTA_X; }
  inline MyFloat lengthY() const { return _SIZE_Y * _DELTA_Y; }

 protected:
  int _SIZE_X;
  int _SIZE_Y;
  MyFloat _DELTA_X;
  MyFloat _DELTA_Y;
};

// Bilinear sample of a row-major grid at (x, y) in world units (truncating casts,
// fraction before clamping, x then y), shared by Grid<T> and the MacGrid mirror.
inline MyFloat fsgSampleBilinear(const MyFloat* data, int nx, int ny, MyFloat dx, MyFloat dy,
                                 MyFloat x, MyFloat y) {
  const MyFloat xs = x / dx, ys = y / dy;
  int i = xs, j = ys;
  const MyFloat fi = xs - i, fj = ys - j;
  i = CLAMP(i, 0, nx - 1);
  j = CLAMP(j, 0, ny - 1);
  const int i1 = CLAMP(i + 1, 0, nx - 1);
  const int j1 = CLAMP(j + 1, 0, ny - 1);
  const MyFloat a = (1 - fi) * data[i + j * nx] + fi * data[i1 + j * nx];
  const MyFloat b = (1 - fi) * data[i + j1 * nx] + fi * data[i1 + j1 * nx];
  return (1 - fj) * a + fj * b;
}

// Bilinear splat of `value` at (x, y) (i+1 taken before the clamp, y then x).
inline void fsgSplatBilinear(MyFloat* data, int nx, int ny, MyFloat dx, MyFloat dy, MyFloat x,
                             MyFloat y, MyFloat value) {
  const MyFloat xs = x / dx, ys = y / dy;
  int i = xs, j = ys;
  int i1 = i + 1, j1 = j + 1;
  const MyFloat fi = xs - i, fj = ys - j;
  i = CLAMP(i, 0, nx - 1);
  j = CLAMP(j, 0, ny - 1);
  i1 = CLAMP(i1, 0, nx - 1);
  j1 = CLAMP(j1, 0, ny - 1);
  const MyFloat w0 = (1 - fj) * value, w1 = fj * value;
  data[i + j * nx] += (1 - fi) * w0;
  data[i1 + j * nx] += fi * w0;
  data[i + j1 * nx] += (1 - fi) * w1;
  data[i1 + j1 * nx] += fi * w1;
}

template <class T>
class Grid : public GridInterface {
 public:
  Grid(int size_x, int size_y, MyFloat delta_x = 1, MyFloat delta_y = 1)
      : GridInterface(size_x, size_y, delta_x, delta_y), data((size_t)size_x * size_y, T()) {}
  ~Grid() {}

  inline T value(int i, int j) const { return data[twoDToLinear(i, j)]; }
  inline T valueInterpolated(MyFloat x, MyFloat y) const {
    return fsgSampleBilinear(data.data(), _SIZE_X, _SIZE_Y, _DELTA_X, _DELTA_Y, x, y);
  }
  inline T& operator()(int i, int j) { return data[twoDToLinear(i, j)]; }
  inline void addToValueInterpolated(MyFloat x, MyFloat y, T value) {
    fsgSplatBilinear(data.data(), _SIZE_X, _SIZE_Y, _DELTA_X, _DELTA_Y, x, y, value);
  }

  // raw storage (row-major), used to move grids to and from the device
  T* raw() { return data.data(); }
  const T* raw() const { return data.data(); }

 protected:
  std::vector<T> data;
};

template <class T>
struct BBox {
  T x_min, x_max, y_min, y_max;
};

#endif
