// LevelSet.h -- level-set grid of the drop-in C++ API (host storage).
//
// Public interface of the reference's include/LevelSet.h:8-18.  The level set is
// a host grid; FluidDomain::classifyCells(LevelSet&) uploads it and labels the
// cells on the GPU.
#ifndef FSG_LEVELSET_H
#define FSG_LEVELSET_H

#include <math.h>

#include "Grid.h"
#include "MathDefinitions.h"

class LevelSet : public Grid<MyFloat> {
 public:
  LevelSet(int size_x, int size_y, MyFloat length_x, MyFloat length_y);
  ~LevelSet();

  MyFloat distance(int from_i, int from_j, int to_i, int to_j);
  MyFloat computeUpwindGradientX(int i, int j, MyFloat vel_x);
  MyFloat computeUpwindGradientY(int i, int j, MyFloat vel_y);
};

#endif
