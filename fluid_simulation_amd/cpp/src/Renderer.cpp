// Renderer.cpp -- draws labels, level sets, face velocities and particles
// (reference: src/Renderer.cpp:19-248).
#include "Renderer.h"

#include <cstdio>
#include <stdexcept>
#include <vector>

Renderer::Renderer(BBox<MyFloat> area) : _area(area) {}

Renderer::~Renderer() {}

void Renderer::clearCanvas(Canvas& canvas) { canvas.fill(Color(1, 1, 1)); }

// scale = pixels per world unit; the translation is half the area's minimum
// corner in pixels, as the reference forms it (src/Renderer.cpp:25-29)
Renderer::View Renderer::view(Canvas& canvas) const {
  View v;
  v.scale_x = canvas.width() / (_area.x_max - _area.x_min);
  v.scale_y = canvas.height() / (_area.y_max - _area.y_min);
  v.translate_x = 0.5 * (_area.x_min * canvas.width());
  v.translate_y = 0.5 * (_area.y_min * canvas.height());
  return v;
}

template <typename ColorOf>
void Renderer::renderCells(const GridInterface& grid, Canvas& canvas, ColorOf&& color_of) {
  const View v = view(canvas);
  const MyFloat cw = grid.deltaX() * v.scale_x, ch = grid.deltaY() * v.scale_y;
  canvas.setLineColor(Color(1, 1, 1));
  for (int j = 0; j < grid.sizeY(); ++j)
    for (int i = 0; i < grid.sizeX(); ++i) {
      canvas.setFillColor(color_of(i, j));
      canvas.fillRectangle({static_cast<int>(-v.translate_x + i * cw),
                            static_cast<int>(-v.translate_x + (i + 1) * cw),
                            static_cast<int>(-v.translate_y + j * ch),
                            static_cast<int>(-v.translate_y + (j + 1) * ch)});
    }
}

void Renderer::renderGridCellsToCanvas(const MacGrid& grid, Canvas& canvas) {
  renderCells(grid, canvas, [&](int i, int j) {
    switch (grid.cellType(i, j)) {
      case LIQUID: return Color(0.7, 0.7, 1);
      case AIR: return Color(1, 1, 1);
      default: return Color(0.5, 0.5, 0.5);
    }
  });
}

void Renderer::renderLevelSetFunctionValuesToCanvas(const LevelSet& level_set, Canvas& canvas) {
  renderCells(level_set, canvas, [&](int i, int j) {
    const MyFloat value = level_set.value(i, j);
    return Color(value, value, value);
  });
}

// A line from each cell centre (all but the last row and column) along the
// cell-centred velocity times 5 pixels, coloured blue -> green -> red with its
// length in units of canvas.width() / 40 pixels.
void Renderer::renderGridVelocitiesToCanvas(const MacGrid& grid, Canvas& canvas) {
  const View v = view(canvas);
  const MyFloat cw = grid.deltaX() * v.scale_x, ch = grid.deltaY() * v.scale_y;
  const MyFloat line_scale = 5;
  for (int j = 0; j < grid.sizeY() - 1; ++j)
    for (int i = 0; i < grid.sizeX() - 1; ++i) {
      const MyFloat vx = grid.velX(i, j), vy = grid.velY(i, j);
      const int x0 = -v.translate_x + (0.5 + i) * cw;
      const int y0 = -v.translate_y + (0.5 + j) * ch;
      const int x1 = x0 + vx * line_scale;
      const int y1 = y0 + vy * line_scale;
      const float len = sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1));
      const float t = len / (canvas.width() / 40);
      const MyFloat blue = CLAMP(1 - t, 0, 1);
      const MyFloat green = CLAMP(t, 0, 1) - CLAMP(t - 1, 0, 1);
      const MyFloat red = CLAMP(t - 1, 0, 1);
      canvas.setLineColor(Color(red, green, blue));
      canvas.drawLine(x0, y0, x1, y1);
    }
}

void Renderer::renderParticlesToCanvas(const MarkerParticleSet& particle_set, Canvas& canvas) {
  const View v = view(canvas);
  canvas.setLineColor(Color(0.3, 0.6, 0.9));
  canvas.setFillColor(Color(0.3, 0.6, 0.9));
  for (const MarkerParticle& p : particle_set) {
    const int x = -v.translate_x + v.scale_x * p.posX();
    const int y = -v.translate_y + v.scale_y * p.posY();
    canvas.drawPoint(x, y, 3);
  }
}

// Each particle adds gaussian(|offset| / 5 clamped to 1, 0.4, 0) over a 10 x 10
// pixel window (offsets -5 .. 4); pixels above 0.6 take the fill colour, above 0.2
// the line colour.
void Renderer::renderMetaBallsToCanvas(const MarkerParticleSet& particle_set, Canvas& canvas) {
  const int w = canvas.width(), h = canvas.height();
  Grid<MyFloat> field(w, h);
  const View v = view(canvas);
  canvas.setLineColor(Color(0.5, 0.9, 1.0));
  canvas.setFillColor(Color(0, 0.5, 0.7));
  const int radius = 5;
  for (const MarkerParticle& p : particle_set) {
    const int px = -v.translate_x + v.scale_x * p.posX();
    const int py = -v.translate_y + v.scale_y * p.posY();
    for (int dj = -radius; dj < radius; ++dj)
      for (int di = -radius; di < radius; ++di) {
        MyFloat t = CLAMP(sqrt((di * di + dj * dj)) / radius, 0, 1);
        t = gaussian(t, 0.4, 0);
        const int x = px + di, y = py + dj;
        if (x >= 0 && x < w && y >= 0 && y < h) field(x, y) += t;
      }
  }
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const MyFloat f = field(x, y);
      if (f > 0.6)
        canvas.setPixel(x, y, canvas.fillColor());
      else if (f > 0.2)
        canvas.setPixel(x, y, canvas.lineColor());
    }
}

// Binary PPM (P6), each channel CLAMP(c, 0, 1) * 255 truncated to a byte.
void Renderer::writeCanvasToPpm(const char* file_path, Canvas& canvas) {
  const int w = canvas.width(), h = canvas.height();
  std::vector<unsigned char> bytes((size_t)w * h * 3);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const Color c = canvas.pixel(x, y);
      unsigned char* px = &bytes[((size_t)x + (size_t)y * w) * 3];
      px[0] = static_cast<unsigned char>(CLAMP(c.r, 0, 1) * 255);
      px[1] = static_cast<unsigned char>(CLAMP(c.g, 0, 1) * 255);
      px[2] = static_cast<unsigned char>(CLAMP(c.b, 0, 1) * 255);
    }
  FILE* fp = fopen(file_path, "wb");
  if (!fp) throw std::runtime_error(std::string("cannot open ") + file_path);
  fprintf(fp, "P6\n%d %d\n255\n", w, h);
  fwrite(bytes.data(), 1, bytes.size(), fp);
  fclose(fp);
}
