// projection.hpp - the pressure projection (src/fluid.cu:229-295) on gfx950.
#pragma once

#include "fluid_device.hpp"
#include "kernels.hpp"

namespace osy {

// Lets the tiled kernels use up to 160 KiB of dynamic LDS on the current
// device; call once per device before the first launch.
hipError_t init_projection_kernels();

// Workgroups of the resident projection (PROJ_RESIDENT) with `rows` rows per
// wavefront that the current device holds at once: a shape fits when its
// tile count is at most this.
hipError_t resident_capacity(int rows, int* slots);

// Enqueues `sp.n` sweeps of the projection.  `*cur` is the u/v buffer index
// holding the input; on return it names the buffer holding the result.
template <typename T>
void enqueue_projection(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                        const ProjTuning& t, hipStream_t s);

}  // namespace osy
