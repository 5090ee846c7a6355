// fluid_device.hpp - per-cell device functions of the solver for gfx950.
//
// Each function computes what the named function of the reference's
// src/fluid.cu computes, with the same IEEE operation order: every float
// operation is a rounded binary32 operation in source order (the library is
// built with -ffp-contract=off; nothing here is fused), the sub-expressions
// the reference writes in double stay double, and division and square root
// are the correctly rounded ones (the HIP default; no fast-math).  That makes
// the FP32 path bit-identical to a sequential IEEE evaluation of the
// reference's update rule.
//
// Device layout (ours, not the reference's): a field is H rows of `pitch`
// elements, row j holding cells (0..W-1, j), j = 0 the bottom row.  The
// boundary converts to the reference's indx order on copy.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>

#pragma clang fp contract(off)

namespace osy {

// Per-cell flag byte, built once from the solid mask.
enum : uint8_t {
  CF_FLUID = 1u << 0,  // !is_solid(i, j)
  CF_UPD = 1u << 1,    // projected cell: 1<=i<=W-2, 1<=j<=H-2 and fluid
  CF_FU = 1u << 2,     // u face (i, j): i >= 1, cells (i-1, j) and (i, j) fluid
  CF_FV = 1u << 3,     // v face (i, j): j >= 1, cells (i, j-1) and (i, j) fluid
  CF_S_SHIFT = 4,      // bits 4..6: total_s, in-bounds fluid 4-neighbours
  CF_S_MASK = 7u << 4,
};

__host__ __device__ inline int cf_total_s(uint8_t f) { return (f >> CF_S_SHIFT) & 7; }

// ---- storage types -------------------------------------------------------
template <typename T> struct Store;
template <> struct Store<float> {
  __device__ static float ld(float x) { return x; }
  __device__ static float st(float x) { return x; }
};
template <> struct Store<__half> {
  __device__ static float ld(__half x) { return __half2float(x); }
  __device__ static __half st(float x) { return __float2half_rn(x); }
};
template <> struct Store<__hip_bfloat16> {
  __device__ static float ld(__hip_bfloat16 x) { return __bfloat162float(x); }
  __device__ static __hip_bfloat16 st(float x) { return __float2bfloat16(x); }
};

template <typename T> __device__ inline float ldf(const T* p, long k) { return Store<T>::ld(p[k]); }
template <typename T> __device__ inline void stf(T* p, long k, float x) { p[k] = Store<T>::st(x); }

// Grid geometry and the flag array, passed by value to every kernel.
struct GridView {
  int W, H, P, cs;
  const uint8_t* flags;
  __device__ long at(int i, int j) const { return (long)j * P + i; }
  // Fluid::is_valid_fluid (src/fluid.cu:356-362)
  __device__ bool valid(int i, int j) const {
    return (unsigned)i < (unsigned)W && (unsigned)j < (unsigned)H &&
           (flags[at(i, j)] & CF_FLUID);
  }
};

// int i = x / cell_size (src/fluid.cu:419-420, 480-481, 645-646).  The
// reference's float->int conversion is undefined for NaN and out-of-range
// values; the quotient is clamped to +-2^24 first (NaN to -2^24) so such
// samples land outside the grid.  In range this is plain truncation.
__device__ inline int to_index(float q) {
  if (!(q >= -16777216.0f)) q = -16777216.0f;
  if (q > 16777216.0f) q = 16777216.0f;
  return (int)q;
}

// get_general_velocity_x (src/fluid.cu:479-539).
template <typename T>
__device__ inline float sample_u(const GridView& g, const T* u, float x, float y) {
  const int cs = g.cs;
  const int i = to_index(x / (float)cs);
  const int j = to_index(y / (float)cs);
  if (!g.valid(i, j)) return 0.0f;
  const float in_x = x - (float)(i * cs);
  const float in_y = y - (float)(j * cs);
  const double h = cs / 2.0;
  float acc = 0.0f;
  if ((double)in_y <= h) {
    const float d_y = (float)(h - (double)in_y);
    const float w_x = 1.0f - in_x / (float)cs;
    const float w_y = 1.0f - d_y / (float)cs;
    acc = acc + (w_y * w_x) * ldf(u, g.at(i, j));
    if (g.valid(i + 1, j)) acc = acc + (w_y * (1.0f - w_x)) * ldf(u, g.at(i + 1, j));
    if (g.valid(i, j - 1)) acc = acc + ((1.0f - w_y) * w_x) * ldf(u, g.at(i, j - 1));
    if (g.valid(i + 1, j - 1))
      acc = acc + ((1.0f - w_y) * (1.0f - w_x)) * ldf(u, g.at(i + 1, j - 1));
  } else {
    const float d_y = (float)((double)in_y - h);
    const float w_x = 1.0f - in_x / (float)cs;
    const float w_y = 1.0f - d_y / (float)cs;
    acc = acc + (w_y * w_x) * ldf(u, g.at(i, j));
    if (g.valid(i, j + 1)) acc = acc + ((1.0f - w_y) * w_x) * ldf(u, g.at(i, j + 1));
    if (g.valid(i + 1, j)) acc = acc + (w_y * (1.0f - w_x)) * ldf(u, g.at(i + 1, j));
    if (g.valid(i + 1, j + 1))
      acc = acc + ((1.0f - w_y) * (1.0f - w_x)) * ldf(u, g.at(i + 1, j + 1));
  }
  return acc;
}

// get_general_velocity_y (src/fluid.cu:418-477); strict `<` on x.
template <typename T>
__device__ inline float sample_v(const GridView& g, const T* v, float x, float y) {
  const int cs = g.cs;
  const int i = to_index(x / (float)cs);
  const int j = to_index(y / (float)cs);
  if (!g.valid(i, j)) return 0.0f;
  const float in_x = x - (float)(i * cs);
  const float in_y = y - (float)(j * cs);
  const double h = cs / 2.0;
  float acc = 0.0f;
  if ((double)in_x < h) {
    const float d_x = (float)(h - (double)in_x);
    const float w_x = 1.0f - d_x / (float)cs;
    const float w_y = 1.0f - in_y / (float)cs;
    acc = acc + (w_y * w_x) * ldf(v, g.at(i, j));
    if (g.valid(i - 1, j)) acc = acc + (w_y * (1.0f - w_x)) * ldf(v, g.at(i - 1, j));
    if (g.valid(i - 1, j + 1))
      acc = acc + ((1.0f - w_y) * (1.0f - w_x)) * ldf(v, g.at(i - 1, j + 1));
    if (g.valid(i, j + 1)) acc = acc + ((1.0f - w_y) * w_x) * ldf(v, g.at(i, j + 1));
  } else {
    const float d_x = (float)((double)in_x - h);
    const float w_x = 1.0f - d_x / (float)cs;
    const float w_y = 1.0f - in_y / (float)cs;
    acc = acc + (w_y * w_x) * ldf(v, g.at(i, j));
    if (g.valid(i, j + 1)) acc = acc + ((1.0f - w_y) * w_x) * ldf(v, g.at(i, j + 1));
    if (g.valid(i + 1, j + 1))
      acc = acc + ((1.0f - w_y) * (1.0f - w_x)) * ldf(v, g.at(i + 1, j + 1));
    if (g.valid(i + 1, j)) acc = acc + (w_y * (1.0f - w_x)) * ldf(v, g.at(i + 1, j));
  }
  return acc;
}

// get_center_position (src/fluid.cu:547-550).
__device__ inline float centre(int a, int cs) { return (float)(((double)a + 0.5) * (double)cs); }

// interpolate_smoke (src/fluid.cu:644-716) with get_distance
// (inc/helper.cuh:77-81).
template <typename T>
__device__ inline float sample_smoke(const GridView& g, const T* s, float x, float y) {
  const int cs = g.cs;
  const int i = to_index(x / (float)cs);
  const int j = to_index(y / (float)cs);
  const float in_x = x - (float)(i * cs);
  const float in_y = y - (float)(j * cs);
  const double h = cs / 2.0;
  const bool left = (double)in_x < h, low = (double)in_y < h;
  const int i2 = left ? i - 1 : i + 1;
  const int j3 = low ? j - 1 : j + 1;
  // cells: 1 = (i, j), 2 = (i2, j), 3 = (i, j3), 4 = (i2, j3)
  const float x1 = centre(i, cs), y1 = centre(j, cs);
  const float x2 = centre(i2, cs), y3 = centre(j3, cs);
  float dx, dy;
  dx = x - x1; dy = y - y1;
  const float inv1 = (float)(1.0 / ((double)sqrtf(dx * dx + dy * dy) + 1e-6));
  dx = x - x2; dy = y - y1;
  const float inv2 = (float)(1.0 / ((double)sqrtf(dx * dx + dy * dy) + 1e-6));
  dx = x - x1; dy = y - y3;
  const float inv3 = (float)(1.0 / ((double)sqrtf(dx * dx + dy * dy) + 1e-6));
  dx = x - x2; dy = y - y3;
  const float inv4 = (float)(1.0 / ((double)sqrtf(dx * dx + dy * dy) + 1e-6));
  const float sum_inv = ((inv1 + inv2) + inv3) + inv4;
  float acc = 0.0f;
  if (g.valid(i, j)) acc = acc + (inv1 / sum_inv) * ldf(s, g.at(i, j));
  if (g.valid(i2, j)) acc = acc + (inv2 / sum_inv) * ldf(s, g.at(i2, j));
  if (g.valid(i, j3)) acc = acc + (inv3 / sum_inv) * ldf(s, g.at(i, j3));
  if (g.valid(i2, j3)) acc = acc + (inv4 / sum_inv) * ldf(s, g.at(i2, j3));
  return acc;
}

}  // namespace osy
