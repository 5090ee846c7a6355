/*
 * fluid_oracle.c - CPU oracle for the OpenSayal per-step fluid solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP solver
 * is compared against, and the host-core baseline the benchmark times beside
 * it.  The product library (opensayal_amd/csrc) neither links nor calls it.
 *
 * It is a plain-C restatement of the reference's update rule, src/fluid.cu of
 * gopmur/OpenSayal, with the operation order of that file.  The reference is
 * CUDA and cannot be built here, and it has no tests or golden data; each
 * function below is pinned instead by the checks in tests/test_oracle_*.py
 * (closed forms, invariants, brute force on tiny grids).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off; no fast-math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/opensayal.h"

typedef struct orc_grid {
  int W, H, cs;
  const int32_t* solid;   /* is_solid, reference layout */
  const int32_t* total_s; /* total_s, reference layout  */
} orc_grid;

/* Fluid::indx (src/fluid.cu:163-165). */
static inline int orc_idx(const orc_grid* g, int i, int j) {
  return (g->H - 1 - j) * g->W + i;
}

/* Fluid::is_valid_fluid = index_is_valid && !is_solid (src/fluid.cu:356-362). */
static inline int orc_valid(const orc_grid* g, int i, int j) {
  return i >= 0 && j >= 0 && i < g->W && j < g->H && !g->solid[orc_idx(g, i, j)];
}

/* ---- 16-bit storage (sim.storage f16 / bf16) ----------------------------
 * The state is held in the storage type: every phase's outputs are rounded
 * to it (round to nearest, ties to even), as if each of the reference's
 * kernels stored into 16-bit arrays; arithmetic stays single precision.
 * Only the single-precision oracle rounds; the double one is the FP64
 * reference run. */
static float orc_round16(float x, int storage) {
  uint32_t b;
  memcpy(&b, &x, 4);
  const uint32_t a = b & 0x7fffffffu;
  if (a > 0x7f800000u) return x; /* NaN */
  if (storage == OSY_STORAGE_BF16) {
    b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
    memcpy(&x, &b, 4);
    return x;
  }
  /* binary16 */
  if (a >= 0x477ff000u) return copysignf(INFINITY, x); /* |x| >= 65520 */
  if (a < 0x38800000u) {                               /* |x| < 2^-14: spacing 2^-24 */
    const float r = x >= 0.0f ? (x + 0.5f) - 0.5f : (x - 0.5f) + 0.5f;
    return copysignf(r, x);
  }
  b = (b + 0x0fffu + ((b >> 13) & 1u)) & 0xffffe000u;
  memcpy(&x, &b, 4);
  return x;
}

static void orc_round_f32(float* a, long n, int storage) {
  if (storage != OSY_STORAGE_F16 && storage != OSY_STORAGE_BF16) return;
  for (long k = 0; k < n; k++) a[k] = orc_round16(a[k], storage);
}

static void orc_round_f64(double* a, long n, int storage) {
  (void)a;
  (void)n;
  (void)storage;
}

/* Rounding of an array in place, for the tests of the rounding itself. */
void orc_round16_array(float* a, long n, int storage) { orc_round_f32(a, n, storage); }

/* ---- single precision: the reference's arithmetic ---------------------- */
#define REAL float
#define SFX _f32
#define ORC_SQRT sqrtf
#define ORC_EXP expf
#define ORC_FMIN fminf
#define ORC_FMAX fmaxf
#include "fluid_oracle_impl.h"
#undef REAL
#undef SFX
#undef ORC_SQRT
#undef ORC_EXP
#undef ORC_FMIN
#undef ORC_FMAX

/* ---- double precision: the FP64 run for precision-mode error reports --- */
#define REAL double
#define SFX _f64
#define ORC_SQRT sqrt
#define ORC_EXP exp
#define ORC_FMIN fmin
#define ORC_FMAX fmax
#include "fluid_oracle_impl.h"
#undef REAL
#undef SFX
#undef ORC_SQRT
#undef ORC_EXP
#undef ORC_FMIN
#undef ORC_FMAX

/* Fluid::init_device_memory mask build (src/fluid.cu:111-142).  is_solid is
 * the left column, bottom and top rows, the right column when the drain is
 * off, the obstacle disc (distance computed in double with std::pow/sqrt, a
 * strict `<` against the float radius), and the pipe walls (pipe_length is
 * fixed at 0 by the reference's parser, so they never appear).  total_s
 * counts in-bounds non-solid 4-neighbours of every cell; the reference
 * increments an uninitialised malloc buffer, the intent (and the value for
 * large grids, where glibc maps zeroed pages) is a count from 0. */
int orc_build_masks(const osy_config* c, int32_t* is_solid, int32_t* total_s) {
  const int W = c->width, H = c->height;
  if (W < 3 || H < 3) return -1;
  orc_grid g = {W, H, 1, is_solid, total_s};
  for (int i = 0; i < W; i++) {
    for (int j = 0; j < H; j++) {
      const double dx = (double)i - (double)c->obstacle_center_x;
      const double dy = (double)j - (double)c->obstacle_center_y;
      const int in_obstacle =
          c->obstacle_enable && sqrt(dx * dx + dy * dy) < (double)c->obstacle_radius;
      const int pipe_wall =
          i < c->pipe_length &&
          (j == H / 2 - c->pipe_height / 2 - 1 || j == H / 2 + c->pipe_height / 2 + 1);
      is_solid[orc_idx(&g, i, j)] =
          (i == 0 || j == 0 || j == H - 1 || (!c->enable_drain && i == W - 1) ||
           in_obstacle || pipe_wall);
    }
  }
  for (int i = 0; i < W; i++) {
    for (int j = 0; j < H; j++) {
      int s = 0;
      if (i - 1 >= 0 && !is_solid[orc_idx(&g, i - 1, j)]) s++;
      if (i + 1 < W && !is_solid[orc_idx(&g, i + 1, j)]) s++;
      if (j - 1 >= 0 && !is_solid[orc_idx(&g, i, j - 1)]) s++;
      if (j + 1 < H && !is_solid[orc_idx(&g, i, j + 1)]) s++;
      total_s[orc_idx(&g, i, j)] = s;
    }
  }
  return 0;
}

/* ---- flat entry points for ctypes -------------------------------------- */

#define ORC_GRID(c, solid, ts) \
  orc_grid g = {(c)->width, (c)->height, (int)(c)->cell_size, (solid), (ts)}

#define ORC_EXPORTS(R, S)                                                      \
  void orc_step##S(const osy_config* c, const int32_t* solid,                  \
                   const int32_t* ts, R* u, R* v, R* smoke, R* p, R* scratch,  \
                   R* prange, const osy_source* src, R d_t, int steps,         \
                   int nthreads) {                                             \
    ORC_GRID(c, solid, ts);                                                    \
    for (int s = 0; s < steps; s++)                                            \
      orc_update##S(c, &g, u, v, smoke, p, scratch, prange, src, d_t,          \
                    nthreads);                                                 \
  }                                                                            \
  void orc_phase_forces##S(const osy_config* c, const int32_t* solid,          \
                           const int32_t* ts, R* u, R* v, R* smoke, R* p,      \
                           const osy_source* src, R d_t) {                     \
    ORC_GRID(c, solid, ts);                                                    \
    orc_forces##S(c, &g, u, v, smoke, p, src, d_t, 1);                         \
  }                                                                            \
  void orc_phase_projection##S(const osy_config* c, const int32_t* solid,      \
                               const int32_t* ts, R* u, R* v, R* p, R d_t,     \
                               int sweeps, int nthreads) {                     \
    ORC_GRID(c, solid, ts);                                                    \
    orc_projection##S(c, &g, u, v, p, d_t, sweeps, nthreads);                  \
  }                                                                            \
  void orc_phase_colour##S(const osy_config* c, const int32_t* solid,          \
                           const int32_t* ts, R* u, R* v, R* p, R d_t,         \
                           int colour) {                                       \
    ORC_GRID(c, solid, ts);                                                    \
    orc_projection_phase##S(c, &g, u, v, p, d_t, colour, 1);                   \
  }                                                                            \
  void orc_phase_extrapolation##S(const osy_config* c, const int32_t* solid,   \
                                  const int32_t* ts, R* u, R* v) {             \
    ORC_GRID(c, solid, ts);                                                    \
    orc_extrapolation##S(&g, u, v);                                            \
  }                                                                            \
  void orc_phase_advect_velocity##S(const osy_config* c, const int32_t* solid, \
                                    const int32_t* ts, const R* u,             \
                                    const R* v, R* u_out, R* v_out, R d_t) {   \
    ORC_GRID(c, solid, ts);                                                    \
    orc_advect_velocity##S(&g, u, v, u_out, v_out, d_t, 1);                    \
  }                                                                            \
  void orc_phase_advect_smoke##S(const osy_config* c, const int32_t* solid,    \
                                 const int32_t* ts, const R* u, const R* v,    \
                                 const R* smoke, R* smoke_out, R d_t) {        \
    ORC_GRID(c, solid, ts);                                                    \
    orc_advect_smoke##S(&g, u, v, smoke, smoke_out, d_t, 1);                   \
  }                                                                            \
  void orc_phase_decay##S(const osy_config* c, const int32_t* solid,           \
                          const int32_t* ts, R* smoke, R rate, R d_t) {        \
    ORC_GRID(c, solid, ts);                                                    \
    orc_decay##S(&g, smoke, rate, d_t, 1);                                     \
  }                                                                            \
  void orc_phase_diffusion##S(const osy_config* c, const int32_t* solid,       \
                              const int32_t* ts, R* u, R* tmp, R d_t,          \
                              int sweeps) {                                    \
    ORC_GRID(c, solid, ts);                                                    \
    orc_diffusion##S(c, &g, u, tmp, d_t, sweeps, 1);                           \
  }                                                                            \
  R orc_point_gx##S(const osy_config* c, const int32_t* solid,                 \
                    const int32_t* ts, const R* u, R x, R y) {                 \
    ORC_GRID(c, solid, ts);                                                    \
    return orc_gx##S(&g, u, x, y);                                             \
  }                                                                            \
  R orc_point_gy##S(const osy_config* c, const int32_t* solid,                 \
                    const int32_t* ts, const R* v, R x, R y) {                 \
    ORC_GRID(c, solid, ts);                                                    \
    return orc_gy##S(&g, v, x, y);                                             \
  }                                                                            \
  R orc_point_idw##S(const osy_config* c, const int32_t* solid,                \
                     const int32_t* ts, const R* smoke, R x, R y) {            \
    ORC_GRID(c, solid, ts);                                                    \
    return orc_idw##S(&g, smoke, x, y);                                        \
  }                                                                            \
  void orc_divergence##S(const osy_config* c, const int32_t* solid,            \
                         const int32_t* ts, const R* u, const R* v,            \
                         double* out) {                                        \
    ORC_GRID(c, solid, ts);                                                    \
    orc_divergence_norm##S(&g, u, v, out);                                     \
  }

ORC_EXPORTS(float, _f32)
ORC_EXPORTS(double, _f64)

/* ---- frame export: update_fluid_pixels_kernel (src/graphics_handler.cu:
 * 214-285) with hsv_to_rgb and map_rgba (src/helper.cu:3-47) -----------------
 * float -> integer conversions follow the GPU's convert instructions: round
 * toward zero, saturate to the destination range, NaN to 0 (the reference's
 * static_cast is undefined outside the range). */
static uint32_t orc_sat_u8(float x) {
  if (!(x > 0.0f)) return 0; /* negatives, -0 and NaN */
  if (x >= 255.0f) return 255;
  return (uint32_t)x;
}

static void orc_hsv_to_rgb(float h, float s, float v, uint32_t* r, uint32_t* g, uint32_t* b) {
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
  *r = orc_sat_u8((r_ + m) * 255.0f);
  *g = orc_sat_u8((g_ + m) * 255.0f);
  *b = orc_sat_u8((b_ + m) * 255.0f);
}

static uint32_t orc_rgba(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return r << 24 | g << 16 | b << 8 | a;
}

static float orc_clampf(float x, float lo, float hi) { return (x < lo) ? lo : (x > hi) ? hi : x; }

/* One pixel; the image is in the reference's indx order (row 0 = top),
 * which is the screen order of the reference's pixel buffer. */
uint32_t orc_pixel(int solid, int pressure_on, int smoke_on, float smoke, float p, float pmin,
                   float pmax) {
  uint32_t r, g, b;
  if (solid) return orc_rgba(80, 80, 80, 255);
  if (pressure_on && smoke_on) {
    float norm = 0.0f;
    if (p < 0.0f && pmin != 0.0f)
      norm = -p / pmin;
    else if (pmax != 0.0f)
      norm = p / pmax;
    norm = orc_clampf(norm, -1.0f, 1.0f);
    orc_hsv_to_rgb((1.0f - norm) * 120.0f, 1.0f, smoke, &r, &g, &b);
    return orc_rgba(r, g, b, 255);
  }
  if (pressure_on) {
    float norm = p < 0.0f ? -p / pmin : p / pmax;
    norm = orc_clampf(norm, -1.0f, 1.0f);
    orc_hsv_to_rgb((1.0f - norm) * 120.0f, 1.0f, 1.0f, &r, &g, &b);
    return orc_rgba(r, g, b, 255);
  }
  if (smoke_on) {
    const uint32_t c = (255u - orc_sat_u8(smoke * 255.0f)) & 255u;
    return orc_rgba(255, c, c, 255);
  }
  return 0; /* the reference leaves such pixels unwritten */
}

void orc_render_rgba(const osy_config* c, const int32_t* solid, const float* smoke,
                     const float* p, float pmin, float pmax, uint32_t* out) {
  const long n = (long)c->width * c->height;
  for (long k = 0; k < n; k++)
    out[k] = orc_pixel(solid[k], c->enable_pressure, c->enable_smoke, smoke[k], p[k], pmin, pmax);
}

/* ---- path lines: Fluid::trace (src/fluid.cu:16-36) driven by
 * update_traces_kernel / update_trace_at (src/graphics_handler.cu:367-396) --
 * Seeds every `distance` cells, (ti * distance, tj * distance) for
 * ti < W / distance, tj < H / distance; seed (ti, tj) owns the `length`
 * entries from indx_traces(ti, tj, 0) = ((th - 1 - tj) * tw + ti) * length
 * (src/graphics_handler.cu:7-11).  A seed on a solid cell gets -1 in every
 * entry (the reference writes -1 to entry 0, which is what its drawing loop
 * tests, and leaves the rest unwritten).  Otherwise the path starts at the
 * cell centre and moves by (GX, GY) * d_t per point; each point is stored as
 * (round(x), H - 1 - round(y)), round half away from zero, with the float ->
 * int conversion of the GPU's convert instruction (saturate, NaN to 0; the
 * reference's static_cast is undefined outside the int range). */
static int32_t orc_sat_i32(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x < -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}

static void orc_trace_one(const orc_grid* g, const float* u, const float* v, int i, int j,
                          int length, float d_t, int32_t* tx, int32_t* ty) {
  float x = orc_center_f32(i, g->cs);
  float y = orc_center_f32(j, g->cs);
  for (int k = 0; k < length; k++) {
    if (k > 0) {
      const float vx = orc_gx_f32(g, u, x, y);
      const float vy = orc_gy_f32(g, v, x, y);
      x = x + vx * d_t;
      y = y + vy * d_t;
    }
    tx[k] = orc_sat_i32(roundf(x));
    ty[k] = (int32_t)(uint32_t)((long)g->H - 1 - orc_sat_i32(roundf(y)));
  }
}

int orc_trace_grid(const osy_config* c, const int32_t* solid, const int32_t* ts, const float* u,
                   const float* v, int distance, int length, float d_t, int32_t* tx,
                   int32_t* ty) {
  if (distance < 1 || length < 1) return -1;
  ORC_GRID(c, solid, ts);
  const int tw = g.W / distance, th = g.H / distance;
  for (int tj = 0; tj < th; tj++) {
    for (int ti = 0; ti < tw; ti++) {
      const long o = ((long)(th - 1 - tj) * tw + ti) * length;
      const int i = ti * distance, j = tj * distance;
      if (solid[orc_idx(&g, i, j)]) {
        for (int k = 0; k < length; k++) tx[o + k] = ty[o + k] = -1;
      } else {
        orc_trace_one(&g, u, v, i, j, length, d_t, tx + o, ty + o);
      }
    }
  }
  return 0;
}

/* Lets a loader check that it opened this library and not another. */
int orc_abi_marker(void) { return 0x05a7a1; }
