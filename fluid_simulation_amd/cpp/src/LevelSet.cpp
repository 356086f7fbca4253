// LevelSet.cpp -- host level set (initialised to 0, first-order upwind gradients).
#include "LevelSet.h"

LevelSet::LevelSet(int size_x, int size_y, MyFloat length_x, MyFloat length_y)
    : Grid<MyFloat>(size_x, size_y, length_x / size_x, length_y / size_y) {}

LevelSet::~LevelSet() {}

MyFloat LevelSet::distance(int from_i, int from_j, int to_i, int to_j) {
  const MyFloat di = to_i - from_i, dj = to_j - from_j;
  return sqrt(di * di + dj * dj);
}

// one-sided difference against the flow direction; indices clamped
MyFloat LevelSet::computeUpwindGradientX(int i, int j, MyFloat vel_x) {
  int ip = i + 1, im = i - 1;
  i = CLAMP(i, 0, _SIZE_X - 1);
  j = CLAMP(j, 0, _SIZE_Y - 1);
  ip = CLAMP(ip, 0, _SIZE_X - 1);
  im = CLAMP(im, 0, _SIZE_X - 1);
  return vel_x < 0 ? (*this)(ip, j) - (*this)(i, j) : (*this)(i, j) - (*this)(im, j);
}

MyFloat LevelSet::computeUpwindGradientY(int i, int j, MyFloat vel_y) {
  int jp = j + 1, jm = j - 1;
  i = CLAMP(i, 0, _SIZE_X - 1);
  j = CLAMP(j, 0, _SIZE_Y - 1);
  jp = CLAMP(jp, 0, _SIZE_Y - 1);
  jm = CLAMP(jm, 0, _SIZE_Y - 1);
  return vel_y < 0 ? (*this)(i, jp) - (*this)(i, j) : (*this)(i, j) - (*this)(i, jm);
}
