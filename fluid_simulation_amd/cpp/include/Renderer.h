// Renderer.h -- software renderer of the drop-in C++ API.
//
// Public interface of the reference's include/Renderer.h:12-31.  It draws the host
// mirrors of the device state (labels, face velocities, particles) that MacGrid
// and MarkerParticleSet download on first access after a step.
#ifndef FSG_RENDERER_H
#define FSG_RENDERER_H

#include "Canvas.h"
#include "LevelSet.h"
#include "MacGrid.h"
#include "MarkerParticleSet.h"
#include "MathDefinitions.h"

class Renderer {
 public:
  Renderer(BBox<MyFloat> area);
  ~Renderer();

  void clearCanvas(Canvas& canvas);
  void renderGridVelocitiesToCanvas(const MacGrid& grid, Canvas& canvas);
  void renderGridCellsToCanvas(const MacGrid& grid, Canvas& canvas);
  void renderLevelSetFunctionValuesToCanvas(const LevelSet& level_set, Canvas& canvas);
  void renderParticlesToCanvas(const MarkerParticleSet& particle_set, Canvas& canvas);
  void renderMetaBallsToCanvas(const MarkerParticleSet& particle_set, Canvas& canvas);

  void writeCanvasToPpm(const char* file_path, Canvas& canvas);

 private:
  // world -> pixel map of the area: pixel = -translate + scale * world
  struct View {
    MyFloat scale_x, scale_y, translate_x, translate_y;
  };
  View view(Canvas& canvas) const;
  // cells as rectangles (white outlines) filled by `color_of(i, j)`
  template <typename ColorOf>
  void renderCells(const GridInterface& grid, Canvas& canvas, ColorOf&& color_of);

  BBox<MyFloat> _area;  // world rectangle shown on the canvas
};

#endif
