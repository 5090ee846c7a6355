// kernels.hpp - launch interface of the solver's HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/opensayal.h"

namespace osy {

// Geometry of the solid mask (src/fluid.cu:111-126).
struct MaskParams {
  int W, H, P;
  int enable_drain;
  int obstacle_enable;
  double cx, cy, radius;  // radius as the float config value, widened
  int pipe_length, pipe_height;
};

// Everything a step needs besides the field pointers.
struct StepParams {
  int W, H, P, cs;
  int n;                 // projection sweeps
  float o, density, d_t;
  // apply_external_forces_at (src/fluid.cu:308-349)
  int smoke_length;
  int band_lo, band_hi;  // u = speed for band_lo <= j <= band_hi
  int smoke_count, smoke_height, smoke_mod, smoke_top;  // stripe rule
  int sband_lo, sband_hi;                                // single-band rule
  float speed, wind_smoke;
  int drag_on;
  float damping;          // expf(-drag_coeff * d_t), computed once on the host
  float gdt;              // g * d_t
  int src_active, src_x, src_y;
  float src_smoke, src_velocity;
  int pressure_on;
  int smoke_on;           // enable_smoke && wind_tunnel.smoke != 0
  int decay_on;
  float decay_rd;         // decay_rate * d_t
  // apply_diffusion_at (src/fluid.cu:176-183): a = viscosity * d_t / cs^2
  // and the divisor 1 + 4a, both in float as the reference computes them
  int visc_on;
  float visc_a, visc_den;
  // set by the step: the first projection launch applies the velocity
  // forces to what it loads and starts the pressure sums at 0
  int forces_in_projection;
  // 1.0 / (double)d_t when 0 < |d_t| < 2, else 0: the pressure sums'
  // division by d_t as a double multiply (PressureDivisor)
  double pdt_recip;
  // bound of one neighbour wait of the resident projection (wall-clock ticks)
  long long spin_ticks;
};

// Device buffers of one simulation; T is the storage type of u, v, smoke.
struct FieldPtrs {
  void* u[2];
  void* v[2];
  void* smoke[2];
  float* p;
  uint8_t* flags;
  uint8_t* gates;   // projection gate byte per cell (static: built with the flags)
  float* scratch;   // reductions: partial results
  float* prange;    // [min, max] of pressure
  double* dscratch; // divergence reduction partials
  unsigned* rflags; // resident projection: stages each tile has published,
                    // then the count of finished workgroups (kResidentMaxTiles + 1)
  int* error;       // set by a resident projection whose neighbour wait timed out
};

// Projection kernels selectable per handle.
enum ProjectionKernel : int {
  PROJ_COLOUR = 0,  // one launch per colour phase, 2n launches, in place
  PROJ_TILED = 1,   // LDS tiles, several sweeps per launch, ping-pong
  PROJ_REGS = 2,    // register-resident tiles, several sweeps per launch
  PROJ_RESIDENT = 3,  // register-resident tiles kept for all n sweeps, one launch
};

// Tiles a resident projection may have (its flag words); a shape also needs
// every tile's workgroup resident on the GPU at once.
constexpr int kResidentMaxTiles = 4096;

// Register-resident kernel: 8 wavefronts stacked vertically, each holding
// 128 cells x kRegRows rows in VGPRs; the region is 128 x 8*kRegRows cells.
constexpr int kRegWaves = 8;
inline bool reg_rows_supported(int rows) {
  return rows == 4 || rows == 8 || rows == 10 || rows == 12 || rows == 16;
}

// Workgroup of the tiled kernel: 8 wavefronts of 64 lanes.
constexpr int kTileWaves = 8;
constexpr int kTileThreads = 64 * kTileWaves;
constexpr size_t kLdsLimit = 160 * 1024;

// Shape of the tiled projection.  A workgroup owns a tile of tw x th cells;
// it loads the tile plus a 2k-cell halo on every side (the region) into LDS,
// runs up to k sweeps there and writes the tile back.  A lane handles one
// pair of horizontally adjacent cells; a region row of pw pairs takes
// ceil(pw / 64) wavefronts.
struct ProjTuning {
  int kernel = PROJ_TILED;
  int tw = 240;  // tile width (even)
  int th = 32;   // tile height
  int k = 4;     // sweeps per launch
  int region_w() const { return tw + 4 * k; }
  int region_h() const { return th + 4 * k; }
  int pw() const { return region_w() / 2; }
  int waves_per_row() const {
    const int n = (pw() + 63) / 64;
    return n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : 8;
  }
  // LDS: u and v (float), gate and flag bytes, and pressure sums (float)
  size_t lds_bytes(bool pressure) const {
    const size_t cells = (size_t)region_h() * 2 * (pw() + 1);
    const size_t flag_bytes = (size_t)(region_h() + 1) * (region_w() + 1);
    return cells * 4 * (pressure ? 3 : 2) + cells + flag_bytes;
  }
  // PROJ_REGS: rows each wavefront holds
  int reg_rows() const { return (th + 4 * k) / kRegWaves; }
  bool valid(bool pressure) const {
    if (kernel == PROJ_COLOUR) return true;
    if (kernel == PROJ_REGS || kernel == PROJ_RESIDENT) {
      // pressure sums need 2 more VGPRs per row: 16 rows spill
      const bool regs = k >= 1 && k <= 28 && tw == 128 - 4 * k && th >= 1 &&
                        (th + 4 * k) % kRegWaves == 0 && reg_rows_supported(reg_rows()) &&
                        !(pressure && reg_rows() > 8);
      // resident: a halo of 2k reaches only the adjacent tiles
      return regs && (kernel == PROJ_REGS || (tw >= 2 * k && th >= 2 * k));
    }
    return kernel == PROJ_TILED && k >= 1 && k <= 64 && tw >= 2 && th >= 1 && tw % 2 == 0 &&
           pw() <= 64 * kTileWaves && lds_bytes(pressure) <= kLdsLimit;
  }
  // projection launches for `sweeps` sweeps
  int launches(int sweeps) const {
    if (kernel == PROJ_RESIDENT) return sweeps > 0 ? 1 : 0;
    return kernel == PROJ_COLOUR ? 2 * sweeps : (sweeps + k - 1) / k;
  }
  // register kernels: tiles over a W x H grid
  int tiles(int W, int H) const { return ((W + tw - 1) / tw) * ((H + th - 1) / th); }
};

// Number of float slots `scratch` must hold, and doubles for `dscratch`.
constexpr int kReduceBlocks = 512;

// Cell flags, then the projection gate bytes derived from them.
void launch_build_flags(const MaskParams& m, uint8_t* flags, uint8_t* gates, hipStream_t s);

// One full step (src/fluid.cu:770-795).  `uv`/`sm` give the buffer index
// currently holding u,v / smoke and are advanced to the index holding the
// step's result.
void enqueue_step(int storage, const StepParams& sp, const FieldPtrs& f,
                  int* uv, int* sm, const ProjTuning& t, hipStream_t s);

// The projection alone (`sp.n` sweeps); advances *uv like enqueue_step.
void enqueue_projection_only(int storage, const StepParams& sp, const FieldPtrs& f, int* uv,
                             const ProjTuning& t, hipStream_t s);

// Divergence L2 / max-abs in FP64 over interior fluid cells; result in
// dscratch[0], dscratch[1] after the stream reaches this point.
void enqueue_divergence(int storage, const StepParams& sp, const FieldPtrs& f,
                        int uv, hipStream_t s);

// Layout conversion between the device layout and the reference's indx
// order, for a field of storage type `storage` (or float/int32 for the
// other fields).  `ref` is a device buffer of W*H 4-byte elements.
void enqueue_export(int field, int storage, const StepParams& sp,
                    const FieldPtrs& f, int uv, int sm, void* ref, hipStream_t s);
void enqueue_import(int field, int storage, const StepParams& sp,
                    const FieldPtrs& f, int uv, int sm, const void* ref, hipStream_t s);

// Frame export (update_fluid_pixels_kernel, src/graphics_handler.cu:258-285):
// W*H RGBA8888 pixels in the reference's indx order into the device buffer
// `out`, using the pressure range in f.prange.  pressure_on / smoke_on are the
// config's enable_pressure / enable_smoke.
void enqueue_render(int storage, const StepParams& sp, const FieldPtrs& f, int sm,
                    int pressure_on, int smoke_on, uint32_t* out, hipStream_t s);

// Path lines (update_traces_kernel, src/graphics_handler.cu:384-396) for the
// seed grid of spacing `distance`: (W/distance)*(H/distance)*length int32
// entries into each of the device buffers tx, ty.
void enqueue_traces(int storage, const StepParams& sp, const FieldPtrs& f, int uv, int distance,
                    int length, float d_t, int32_t* tx, int32_t* ty, hipStream_t s);

// Plain device-to-device copy of `bytes` (a multiple of 16) with 16-byte
// loads and stores, for calibrating hardware byte counters.
void enqueue_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

}  // namespace osy
