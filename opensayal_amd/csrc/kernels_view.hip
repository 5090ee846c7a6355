// kernels_view.hip - what the reference's GraphicsHandler computes on the
// device from a Fluid each frame: the RGBA frame of
// update_fluid_pixels_kernel (src/graphics_handler.cu:214-285) and the path
// lines of update_traces_kernel / Fluid::trace (src/graphics_handler.cu:
// 367-396, src/fluid.cu:16-36).  Both read the solver's device layout
// directly and write the reference's host-side order, so the host copy is a
// single contiguous transfer.
//
// float -> integer conversions are written out (truncate, saturate, NaN to
// 0) because the reference's static_casts are undefined outside the
// destination range; the CPU restatement makes the same choice.
#include "fluid_device.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace osy {
namespace {

constexpr int kViewBX = 64, kViewBY = 4;

__device__ inline uint32_t sat_u8(float x) {
  return x > 0.0f ? (x >= 255.0f ? 255u : (uint32_t)x) : 0u;
}

__device__ inline int32_t sat_i32(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x < -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}

// map_rgba (src/helper.cu:45-47)
__device__ inline uint32_t rgba(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return r << 24 | g << 16 | b << 8 | a;
}

// hsv_to_rgb (src/helper.cu:3-43), hue in degrees; returns the opaque pixel.
__device__ inline uint32_t hsv_pixel(float h, float s, float v) {
  const float c = v * s;
  const float x = c * (1.0f - fabsf(fmodf(h / 60.0f, 2.0f) - 1.0f));
  const float m = v - c;
  float r_, g_, b_;
  if (h < 60.0f) {
    r_ = c; g_ = x; b_ = 0.0f;
  } else if (h < 120.0f) {
    r_ = x; g_ = c; b_ = 0.0f;
  } else if (h < 180.0f) {
    r_ = 0.0f; g_ = c; b_ = x;
  } else if (h < 240.0f) {
    r_ = 0.0f; g_ = x; b_ = c;
  } else if (h < 300.0f) {
    r_ = x; g_ = 0.0f; b_ = c;
  } else {
    r_ = c; g_ = 0.0f; b_ = x;
  }
  return rgba(sat_u8((r_ + m) * 255.0f), sat_u8((g_ + m) * 255.0f), sat_u8((b_ + m) * 255.0f),
              255u);
}

// clamp (src/helper.cu:49-51); a NaN passes through.
__device__ inline float clampf(float x, float lo, float hi) {
  return (x < lo) ? lo : (x > hi) ? hi : x;
}

// One frame pixel per cell.  Row j of the device layout lands in row
// H - 1 - j of the frame: GraphicsHandler::indx(i, H - 1 - j) equals
// Fluid::indx(i, j).  A cell with neither smoke nor pressure shown gets 0
// (the reference leaves it unwritten).
template <typename T>
__global__ void __launch_bounds__(kViewBX * kViewBY)
k_render(int W, int H, int P, const uint8_t* __restrict__ flags, const T* __restrict__ smoke,
         const float* __restrict__ p, const float* __restrict__ prange, int pressure_on,
         int smoke_on, uint32_t* __restrict__ out) {
  // pressure_on / smoke_on are the config's enable_pressure / enable_smoke,
  // the flags update_fluid_pixels_kernel tests (:268-284)
  const int i = blockIdx.x * kViewBX + threadIdx.x;
  const int j = blockIdx.y * kViewBY + threadIdx.y;
  if (i >= W || j >= H) return;
  const long k = (long)j * P + i;
  uint32_t px = 0;
  if (!(flags[k] & CF_FLUID)) {
    px = rgba(80, 80, 80, 255);
  } else if (pressure_on) {
    const float pmin = prange[0], pmax = prange[1];
    const float pr = p[k];
    float norm = 0.0f;
    if (smoke_on) {  // update_smoke_and_pressure (:221-238)
      if (pr < 0.0f && pmin != 0.0f)
        norm = -pr / pmin;
      else if (pmax != 0.0f)
        norm = pr / pmax;
    } else {  // update_pressure_pixel (:240-256)
      norm = pr < 0.0f ? -pr / pmin : pr / pmax;
    }
    norm = clampf(norm, -1.0f, 1.0f);
    px = hsv_pixel((1.0f - norm) * 120.0f, 1.0f, smoke_on ? ldf(smoke, k) : 1.0f);
  } else if (smoke_on) {  // update_smoke_pixels (:214-219)
    const uint32_t c = 255u - sat_u8(ldf(smoke, k) * 255.0f);
    px = rgba(255, c, c, 255);
  }
  out[(long)(H - 1 - j) * W + i] = px;
}

// Path lines: one thread per seed (ti * distance, tj * distance) for
// ti < tw = W / distance, tj < th = H / distance; the seed's `length` points
// start at indx_traces(ti, tj, 0) = ((th - 1 - tj) * tw + ti) * length
// (src/graphics_handler.cu:7-11).  A solid seed gets -1 in every entry (the
// reference marks entry 0, which its drawing loop tests).  Otherwise the
// position starts at the cell centre and moves by the general velocity
// (get_general_velocity_x / _y, src/fluid.cu:418-545, on the stored u and v)
// times d_t; each point is (round(x), H - 1 - round(y)), round half away
// from zero.
template <typename T, bool CS1>
__global__ void __launch_bounds__(256)
k_traces(GridView g, const T* __restrict__ u, const T* __restrict__ v, int distance, int tw,
         int th, int length, float d_t, int32_t* __restrict__ tx, int32_t* __restrict__ ty) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= tw * th) return;
  const int tj = s / tw, ti = s - tj * tw;
  const int i = ti * distance, j = tj * distance;
  const long o = ((long)(th - 1 - tj) * tw + ti) * length;
  if (!g.valid(i, j)) {
    for (int k = 0; k < length; k++) tx[o + k] = ty[o + k] = -1;
    return;
  }
  auto valid = [&](int a, int b) -> bool { return g.valid(a, b); };
  auto uval = [&](int a, int b) -> float { return g.in(a, b) ? ldf(u, g.at(a, b)) : 0.0f; };
  auto vval = [&](int a, int b) -> float { return g.in(a, b) ? ldf(v, g.at(a, b)) : 0.0f; };
  const int cs = g.cs;
  float x = centre<CS1>(i, cs), y = centre<CS1>(j, cs);
  const long top = (long)g.H - 1;
  for (int k = 0; k < length; k++) {
    if (k > 0) {
      const float vx = sample_u_bf<CS1>(cs, x, y, valid, uval);
      const float vy = sample_v_bf<CS1>(cs, x, y, valid, vval);
      x = x + vx * d_t;
      y = y + vy * d_t;
    }
    tx[o + k] = sat_i32(roundf(x));
    ty[o + k] = (int32_t)(uint32_t)(top - sat_i32(roundf(y)));
  }
}

template <typename T>
void render_t(const StepParams& sp, const FieldPtrs& f, int sm, int pressure_on, int smoke_on,
              uint32_t* out, hipStream_t s) {
  const dim3 grid((sp.W + kViewBX - 1) / kViewBX, (sp.H + kViewBY - 1) / kViewBY);
  hipLaunchKernelGGL(k_render<T>, grid, dim3(kViewBX, kViewBY), 0, s, sp.W, sp.H, sp.P,
                     f.flags, (const T*)f.smoke[sm], f.p, f.prange, pressure_on, smoke_on, out);
}

template <typename T>
void traces_t(const StepParams& sp, const FieldPtrs& f, int uv, int distance, int length,
              float d_t, int32_t* tx, int32_t* ty, hipStream_t s) {
  const int tw = sp.W / distance, th = sp.H / distance;
  if (tw * th == 0) return;
  const GridView g{sp.W, sp.H, sp.P, sp.cs, f.flags};
  const dim3 grid((tw * th + 255) / 256);
  if (sp.cs == 1)
    hipLaunchKernelGGL((k_traces<T, true>), grid, dim3(256), 0, s, g, (const T*)f.u[uv],
                       (const T*)f.v[uv], distance, tw, th, length, d_t, tx, ty);
  else
    hipLaunchKernelGGL((k_traces<T, false>), grid, dim3(256), 0, s, g, (const T*)f.u[uv],
                       (const T*)f.v[uv], distance, tw, th, length, d_t, tx, ty);
}

}  // namespace

void enqueue_render(int storage, const StepParams& sp, const FieldPtrs& f, int sm,
                    int pressure_on, int smoke_on, uint32_t* out, hipStream_t s) {
  switch (storage) {
    case OSY_STORAGE_F16: render_t<__half>(sp, f, sm, pressure_on, smoke_on, out, s); break;
    case OSY_STORAGE_BF16:
      render_t<__hip_bfloat16>(sp, f, sm, pressure_on, smoke_on, out, s);
      break;
    default: render_t<float>(sp, f, sm, pressure_on, smoke_on, out, s); break;
  }
}

void enqueue_traces(int storage, const StepParams& sp, const FieldPtrs& f, int uv, int distance,
                    int length, float d_t, int32_t* tx, int32_t* ty, hipStream_t s) {
  switch (storage) {
    case OSY_STORAGE_F16: traces_t<__half>(sp, f, uv, distance, length, d_t, tx, ty, s); break;
    case OSY_STORAGE_BF16:
      traces_t<__hip_bfloat16>(sp, f, uv, distance, length, d_t, tx, ty, s);
      break;
    default: traces_t<float>(sp, f, uv, distance, length, d_t, tx, ty, s); break;
  }
}

}  // namespace osy
