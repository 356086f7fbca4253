// MacGrid.cpp -- device-backed MAC grid (drop-in for the reference's MacGrid).
#include "MacGrid.h"

MacGrid::MacGrid(int size_x, int size_y, MyFloat length_x, MyFloat length_y)
    : MacGrid(std::make_shared<FsgContext>(size_x, size_y, length_x, length_y, 1, 0), size_x,
              size_y, length_x, length_y) {}

// spacing = length / size in MyFloat (the reference's MacGrid constructor)
MacGrid::MacGrid(std::shared_ptr<FsgContext> ctx, int size_x, int size_y, MyFloat length_x,
                 MyFloat length_y)
    : GridInterface(size_x, size_y, length_x / size_x, length_y / size_y), _ctx(std::move(ctx)) {
  clearCellTypeBuffer();
}

MacGrid::~MacGrid() {}

// border SOLID, interior AIR (the state MacGrid's constructor leaves)
void MacGrid::clearCellTypeBuffer() {
  std::vector<uint8_t>& lab = _ctx->gridsForWrite().labels;
  for (int j = 0; j < _SIZE_Y; ++j)
    for (int i = 0; i < _SIZE_X; ++i) {
      const bool border = i == 0 || j == 0 || i == _SIZE_X - 1 || j == _SIZE_Y - 1;
      lab[(size_t)i + (size_t)j * _SIZE_X] = border ? SOLID : AIR;
    }
}

void MacGrid::updatePreviousVelocityBuffer() {
  _ctx->beforeDevice();
  _ctx->check(fsg_run_phase(_ctx->handle(), FSG_PH_SAVE_PREV, 0.f), "fsg_run_phase(SAVE_PREV)");
  _ctx->afterDevice();
}

void MacGrid::updateVelocityDiffBuffer() {
  _ctx->beforeDevice();
  _ctx->check(fsg_run_phase(_ctx->handle(), FSG_PH_DIFF, 0.f), "fsg_run_phase(DIFF)");
  _ctx->afterDevice();
}

void MacGrid::swapVelocityBuffers() { _ctx->swapVelocity(); }
