// Canvas.h -- RGB float canvas of the drop-in C++ API.
//
// Public interface of the reference's include/Canvas.h:8-66 (Color, Canvas).  Pixels
// are a row-major Grid<Color> (x = i, y = j); drawing clips to the canvas.
#ifndef FSG_CANVAS_H
#define FSG_CANVAS_H

#include "Grid.h"
#include "MathDefinitions.h"

class Color {
 public:
  Color() : r(1), g(1), b(1) {}
  Color(MyFloat red, MyFloat green, MyFloat blue) : r(red), g(green), b(blue) {}
  ~Color() {}
  MyFloat r, g, b;
  inline Color& operator+=(Color& c) {
    r += c.r;
    g += c.g;
    b += c.b;
    return *this;
  }
  inline Color operator*(MyFloat& s) { return Color(r * s, g * s, b * s); }
};

class Canvas {
 public:
  Canvas(int width, int height);
  ~Canvas();

  // pixel steps from (from_x, from_y) towards (to_x, to_y), stopping at the edge
  void drawLine(int from_x, int from_y, int to_x, int to_y);
  // a size x size square (outline in the line colour, inside in the fill colour)
  void drawPoint(int pos_x, int pos_y, int size);
  // rectangle with inclusive bounds, clamped to the canvas
  void fillRectangle(BBox<int> rect);
  void setPixel(int i, int j, Color c);
  void addToPixel(int i, int j, Color c);
  void fill(Color c);

  Color lineColor();
  Color fillColor();
  int width();
  int height();
  Color pixel(int i, int j);

  void setLineColor(Color c);
  void setFillColor(Color c);

 private:
  bool inside(int i, int j) const { return i >= 0 && i < _WIDTH && j >= 0 && j < _HEIGHT; }

  const int _WIDTH;
  const int _HEIGHT;
  Color _line_color;
  Color _fill_color;
  Grid<Color> _pixel_data;
};

#endif
