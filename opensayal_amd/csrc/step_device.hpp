// step_device.hpp - pointwise pieces of Fluid::update that other kernels
// apply while they load: the external forces (src/fluid.cu:308-349) and the
// edge extrapolation (src/fluid.cu:720-747).
#pragma once

#include "fluid_device.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace osy {

// apply_external_forces_at, velocity part: inflow u = speed, drag, mouse
// source, then v += g * d_t.  A pure function of (i, j, u, v), so a kernel
// may apply it to every cell it loads, halo included.
__device__ inline void force_uv(const StepParams& sp, int i, int j, float& u, float& v) {
  if (i <= sp.smoke_length && i != 0 && j >= sp.band_lo && j <= sp.band_hi) u = sp.speed;
  if (sp.drag_on) {
    u = u * sp.damping;
    v = v * sp.damping;
  }
  if (sp.src_active) {
    const int dx = i - sp.src_x, dy = j - sp.src_y;
    if (dx * dx + dy * dy < 40 * 40) {
      u = u + sp.src_velocity * (float)dx;
      v = v + sp.src_velocity * (float)dy;
    }
  }
  v = v + sp.gdt;
}

// apply_external_forces_at, smoke part: the inflow band (single band or
// stripes) and the mouse source set smoke to a constant.
__device__ inline float force_smoke(const StepParams& sp, int i, int j, float s) {
  if (i <= sp.smoke_length && i != 0 && j >= sp.band_lo && j <= sp.band_hi) {
    if ((sp.smoke_count == 1 && j >= sp.sband_lo && j <= sp.sband_hi) ||
        (sp.smoke_count != 1 && (sp.smoke_top - j) % sp.smoke_mod < sp.smoke_height))
      s = sp.wind_smoke;
  }
  if (sp.src_active && sp.src_smoke != 0.0f) {
    const int dx = i - sp.src_x, dy = j - sp.src_y;
    if (dx * dx + dy * dy < 40 * 40) s = sp.src_smoke;
  }
  return s;
}

// apply_extrapolation as a function of the state before it: reading cell
// (i, j) of the extrapolated field reads another cell (or 0) of the field
// before extrapolation.  Row phase u(i,0) = u(i,1), v(i,1) = 0,
// u(i,H-1) = u(i,H-2); then column phase v(0,j) = v(1,j), u(1,j) = 0,
// v(W-1,j) = v(W-2,j).  Every value the column phase copies is one the row
// phase did not touch except v(1,1) and v(W-2,1), which are 0, so:
//   u_ext(i,j) = 0 if i == 1, else u(i, 1) if j == 0, else u(i, H-2) if
//                j == H-1, else u(i,j)
//   v_ext(i,j) = 0 if j == 1, else v(1, j) if i == 0, else v(W-2, j) if
//                i == W-1, else v(i,j)
// (W, H >= 3.)
template <typename T>
__device__ inline float u_extrapolated(const T* u, const GridView& g, int i, int j) {
  if (i == 1) return 0.0f;
  const int jj = (j == 0) ? 1 : (j == g.H - 1 ? g.H - 2 : j);
  return ldf(u, g.at(i, jj));
}
template <typename T>
__device__ inline float v_extrapolated(const T* v, const GridView& g, int i, int j) {
  if (j == 1) return 0.0f;
  const int ii = (i == 0) ? 1 : (i == g.W - 1 ? g.W - 2 : i);
  return ldf(v, g.at(ii, j));
}

}  // namespace osy
