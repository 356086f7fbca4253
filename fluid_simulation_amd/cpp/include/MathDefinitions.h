// MathDefinitions.h -- numeric type and helpers of the drop-in C++ API.
//
// Same names and semantics as the reference's include/MathDefinitions.h:7-34:
// MyFloat is float (the HIP path computes in fp32), and CLAMP is a MyFloat function
// so integer clamps round-trip through float exactly as the reference's do.  The
// Eigen vector typedef of the reference is not needed: the pressure solve runs in
// libfsg.so.
#ifndef FSG_MATH_DEFINITIONS_H
#define FSG_MATH_DEFINITIONS_H

#include <math.h>

typedef float MyFloat;

inline MyFloat CLAMP(MyFloat d, MyFloat min, MyFloat max) {
  const MyFloat lo = d < min ? min : d;
  return lo > max ? max : lo;
}

inline MyFloat smoothstep(MyFloat edge0, MyFloat edge1, MyFloat x) {
  x = CLAMP((x - edge0) / (edge1 - edge0), 0.0, 1.0);
  return x * x * (3 - 2 * x);
}

inline MyFloat gaussian(MyFloat x, MyFloat sigma, MyFloat mu) {
  const MyFloat norm = 1 / (sigma * sqrt(2 * M_PI));
  const MyFloat d = x - mu;
  return norm * exp(-(d * d) / (2 * sigma * sigma));
}

#endif
