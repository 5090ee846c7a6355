// fluid_internal.hpp - the solver handle behind the C ABI (replaces the
// reference's class Fluid, inc/fluid.cuh:15-116), shared by the single-grid
// entry points (fluid_host.cpp) and the batch (batch_host.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "kernels.hpp"
#include "osy_internal.hpp"

#define OSY_HIP_CHECK(call)                                                              \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::osy::set_error(OSY_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

using osy::FieldPtrs;
using osy::ProjTuning;
using osy::StepParams;

struct osy_fluid {
  osy_config cfg{};
  int W = 0, H = 0, P = 0, cs = 1, storage = 0, elem = 4;
  FieldPtrs f{};
  int uv = 0, sm = 0;
  hipStream_t stream = nullptr;
  int device = 0;
  void* block = nullptr;     // one device allocation holding every buffer
  size_t block_bytes = 0;
  float* staging = nullptr;  // W*H*4 bytes, reference-order copies
  void* trace_buf = nullptr;  // path-line output, grown on demand
  size_t trace_cap = 0;
  ProjTuning tuning;  // projection kernel shape, chosen at create
  long long spin_ticks = 0;  // bound of a resident projection's neighbour wait
  int wall_khz = 100000;     // the device's wall-clock rate
  bool resident_pending = false;  // a resident projection ran since the last
                                  // read of the error word
  osy_batch* batch = nullptr;  // the batch stepping this handle, if any
  struct Graph {
    hipGraphExec_t exec = nullptr;
    float d_t = 0.0f;
    int uv_out = 0, sm_out = 0;
  } graphs[2][2];

  ~osy_fluid() {
    for (auto& row : graphs)
      for (auto& g : row)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (block) (void)hipFree(block);
    if (trace_buf) (void)hipFree(trace_buf);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // StepParams for this configuration and step size; `src` may be null.
  StepParams params(float d_t, const osy_source* src) const {
    const osy_config& c = cfg;
    StepParams sp{};
    sp.W = W; sp.H = H; sp.P = P; sp.cs = cs;
    sp.n = c.projection_n;
    sp.o = c.projection_o;
    sp.density = c.density;
    sp.d_t = d_t;
    sp.smoke_length = c.smoke_length;
    sp.band_lo = H / 2 - c.pipe_height / 2;
    sp.band_hi = H / 2 + c.pipe_height / 2;
    sp.smoke_count = c.smoke_count;
    sp.smoke_height = c.smoke_height;
    if (c.smoke_count != 1) {
      const int spacing = (c.pipe_height - c.smoke_count * c.smoke_height) / (c.smoke_count - 1);
      sp.smoke_mod = spacing + c.smoke_height;
    } else {
      sp.smoke_mod = 1;
    }
    sp.smoke_top = H / 2 + c.pipe_height / 2;
    sp.sband_lo = H / 2 - c.smoke_height / 2;
    sp.sband_hi = H / 2 + c.smoke_height / 2;
    sp.speed = c.wind_speed;
    sp.wind_smoke = c.wind_smoke;
    sp.drag_on = c.drag_coeff != 0.0f;
    // expf(-drag_coeff * d_t) (src/fluid.cu:332) is one value per step;
    // it is evaluated once here with the host's expf.
    sp.damping = sp.drag_on ? expf(-c.drag_coeff * d_t) : 1.0f;
    sp.gdt = c.g * d_t;
    sp.src_active = src && src->active;
    if (sp.src_active) {
      sp.src_x = src->position_x;
      sp.src_y = src->position_y;
      sp.src_smoke = src->smoke;
      sp.src_velocity = src->velocity;
    }
    sp.pressure_on = c.enable_pressure != 0;
    sp.smoke_on = c.enable_smoke && c.wind_smoke != 0.0f;
    sp.decay_on = c.enable_decay != 0;
    sp.decay_rd = c.decay_rate * d_t;
    sp.visc_on = c.viscosity != 0.0f;
    sp.visc_a = c.viscosity * d_t / (float)(cs * cs);
    sp.visc_den = 1.0f + 4.0f * sp.visc_a;
    sp.pdt_recip = (d_t != 0.0f && std::fabs(d_t) < 2.0f) ? 1.0 / (double)d_t : 0.0;
    sp.spin_ticks = spin_ticks;
    return sp;
  }
};

namespace osy {

// Bytes of the single device allocation of a handle of this shape.
size_t fluid_block_bytes(int W, int H, int storage);

// Whether the handle can run projection shape t: valid for its pressure
// mode and, for the resident kernel, every tile's workgroup resident at
// once on its device.
bool tuning_fits(const osy_fluid* fl, const ProjTuning& t);

// osy_fluid_autotune without the batch-membership check: the batch
// autotuner tunes the projection of members that step on their own graphs.
int fluid_autotune(osy_fluid* fl, osy_tuning* chosen, double* ms);

// Reads and clears the device error word (a resident projection whose
// neighbour wait timed out) when a resident projection ran since the last
// read; call with the stream synchronized.
int fluid_check_error(osy_fluid* fl);

// Batch membership hooks (batch_host.cpp): wait for the batch's queued
// steps, and leave the batch (when the handle is destroyed).
int batch_wait(osy_fluid* fl);
void batch_forget(osy_fluid* fl);

// `steps` steps of a handle through its own per-grid step graphs, enqueued
// on `s` (fluid_host.cpp); a batch runs its large members this way.
int fluid_enqueue_graph_steps(osy_fluid* fl, int32_t steps, hipStream_t s);

}  // namespace osy
