// projection.hpp - the pressure projection (src/fluid.cu:229-295) on gfx950.
#pragma once

#include "fluid_device.hpp"
#include "kernels.hpp"

namespace osy {

// Projection kernels selectable per handle.
enum ProjectionKernel : int {
  PROJ_AUTO = -1,   // pick by grid size (tiled when it pays)
  PROJ_COLOUR = 0,  // one launch per colour phase, 2n launches, in place
  PROJ_TILED = 1,   // LDS tiles, several sweeps per launch, ping-pong
};

// Enqueues the n sweeps of the projection.  `*cur` is the u/v buffer index
// holding the input; on return it names the buffer holding the result.
template <typename T>
void enqueue_projection(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                        int kernel, hipStream_t s);

}  // namespace osy
