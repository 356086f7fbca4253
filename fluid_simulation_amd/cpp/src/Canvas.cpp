// Canvas.cpp -- clipped drawing on a row-major RGB float canvas
// (reference: src/Canvas.cpp).
#include "Canvas.h"

Canvas::Canvas(int width, int height)
    : _WIDTH(width),
      _HEIGHT(height),
      _line_color(0, 0, 0),
      _fill_color(1, 1, 1),
      _pixel_data(width, height) {}

Canvas::~Canvas() {}

// Unit steps k = 0, 1, ... while k < |d|, each pixel at from + (d / |d|) * k with
// truncating casts; the first pixel outside the canvas ends the line.
void Canvas::drawLine(int from_x, int from_y, int to_x, int to_y) {
  const MyFloat dx = to_x - from_x;
  const MyFloat dy = to_y - from_y;
  const MyFloat len = sqrt(dx * dx + dy * dy);
  for (int k = 0; k < len; ++k) {
    const int x = from_x + (dx / len) * k;
    const int y = from_y + (dy / len) * k;
    if (!inside(x, y)) return;
    _pixel_data(x, y) = _line_color;
  }
}

void Canvas::drawPoint(int pos_x, int pos_y, int size) {
  const int half = size / 2;
  fillRectangle({pos_x - half, pos_x + half, pos_y - half, pos_y + half});
}

// Bounds clamped through CLAMP (a MyFloat function); edge pixels take the line
// colour, interior pixels the fill colour.
void Canvas::fillRectangle(BBox<int> rect) {
  const int x0 = CLAMP(rect.x_min, 0, _WIDTH - 1), x1 = CLAMP(rect.x_max, 0, _WIDTH - 1);
  const int y0 = CLAMP(rect.y_min, 0, _HEIGHT - 1), y1 = CLAMP(rect.y_max, 0, _HEIGHT - 1);
  for (int y = y0; y <= y1; ++y)
    for (int x = x0; x <= x1; ++x) {
      const bool edge = x == x0 || x == x1 || y == y0 || y == y1;
      _pixel_data(x, y) = edge ? _line_color : _fill_color;
    }
}

void Canvas::setPixel(int i, int j, Color c) {
  if (inside(i, j)) _pixel_data(i, j) = c;
}

void Canvas::addToPixel(int i, int j, Color c) {
  if (inside(i, j)) _pixel_data(i, j) += c;
}

void Canvas::fill(Color c) {
  Color* p = _pixel_data.raw();
  for (size_t q = 0; q < (size_t)_WIDTH * _HEIGHT; ++q) p[q] = c;
}

Color Canvas::lineColor() { return _line_color; }
Color Canvas::fillColor() { return _fill_color; }
int Canvas::width() { return _WIDTH; }
int Canvas::height() { return _HEIGHT; }
Color Canvas::pixel(int i, int j) { return _pixel_data.value(i, j); }
void Canvas::setLineColor(Color c) { _line_color = c; }
void Canvas::setFillColor(Color c) { _fill_color = c; }
