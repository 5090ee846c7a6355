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
// With cell_size 1 (template flag CS1) a few steps are written in a cheaper
// form that gives the same bits: x / 1 == x for every float x; the double
// comparisons in_y <= 1/2.0 and in_x < 1/2.0 equal the float comparisons
// with 0.5f; and (float)(0.5 - (double)t) == 0.5f - t for every float t in
// [0, 1) (the double difference is exact for t >= 2^-29, and for smaller t
// both sides round to 0.5f), likewise (float)((double)t - 0.5) == t - 0.5f
// for t in [0.5, 1) (exact in both).  The sample position has 0 <= t < 1
// whenever its cell is in the grid; outside the grid the result is 0 either
// way.
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
// A value as the storage type holds it: the identity for float, round to
// nearest even for the 16-bit types.  Every phase rounds the faces it
// updates (as a launch per phase storing into 16-bit arrays would), so
// 16-bit results do not depend on how sweeps are grouped into launches.
template <typename T> __device__ inline float rnd(float x) { return Store<T>::ld(Store<T>::st(x)); }

// rnd of two values with one packed conversion (v_cvt_pk_f16_f32 /
// v_cvt_pk_bf16_f32, round to nearest even like the scalar conversions;
// exhaustive check on the device: osy_arith_check, modes 3 and 4).
typedef float Float2V __attribute__((ext_vector_type(2)));
template <typename T> struct Round2 {
  __device__ static void run(float& a, float& b) {
    a = rnd<T>(a);
    b = rnd<T>(b);
  }
};
template <> struct Round2<float> {
  __device__ static void run(float&, float&) {}
};
template <> struct Round2<__half> {
  __device__ static void run(float& a, float& b) {
    typedef _Float16 H2 __attribute__((ext_vector_type(2)));
    const H2 h = __builtin_convertvector((Float2V){a, b}, H2);
    a = (float)h.x;
    b = (float)h.y;
  }
};
template <> struct Round2<__hip_bfloat16> {
  __device__ static void run(float& a, float& b) {
    typedef __bf16 B2 __attribute__((ext_vector_type(2)));
    const B2 h = __builtin_convertvector((Float2V){a, b}, B2);
    a = (float)h.x;
    b = (float)h.y;
  }
};
template <typename T> __device__ inline void rnd2(float& a, float& b) { Round2<T>::run(a, b); }
template <typename T> __device__ inline void stf(T* p, long k, float x) { p[k] = Store<T>::st(x); }

// Workgroup b of a launch of n runs on XCD b % 8 (workgroups are dispatched
// round-robin over the 8 XCDs).  The tile it takes: XCD x gets a contiguous
// run of the row-major tile order, so neighbouring tiles, whose halos
// overlap, share that XCD's L2.  A bijection of [0, n) for every n.
constexpr int kXcds = 8;
__device__ inline int xcd_tile(int b, int n) {
  const int x = b % kXcds, k = b / kXcds;
  const int q = n / kXcds, r = n % kXcds;
  return x * q + (x < r ? x : r) + k;
}

// Grid geometry and the flag array, passed by value to every kernel.
struct GridView {
  int W, H, P, cs;
  const uint8_t* flags;
  __device__ long at(int i, int j) const { return (long)j * P + i; }
  __device__ bool in(int i, int j) const {
    return (unsigned)i < (unsigned)W && (unsigned)j < (unsigned)H;
  }
  // Fluid::is_valid_fluid (src/fluid.cu:356-362)
  __device__ bool valid(int i, int j) const { return in(i, j) && (flags[at(i, j)] & CF_FLUID); }
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

// x / cs, with cs == 1 known at compile time when CS1.
template <bool CS1> __device__ inline float div_cs(float x, int cs) {
  if constexpr (CS1) return x;
  else return x / (float)cs;
}

// Quotient x / n for a neighbour count n in 1..4, as IEEE division: n = 1,
// 2, 4 multiply by the exact reciprocal, n = 3 divides.
__device__ inline float div_count(float x, int n) {
  if (n == 3) return x / 3.0f;
  return x * __uint_as_float(0x3F800000u - ((uint32_t)(n >> 1) << 23));
}

// x / 3 exactly: q0 = RN(x * RN(1/3)) corrected once with fused multiply-adds
// (q0 itself when it is zero or infinite).  Equal to IEEE x / 3.0f for every
// binary32 x (exhaustive check in tests/csrc/quotient_check.c).
__device__ inline float third_quotient(float x) {
  const float y = 1.0f / 3.0f;
  const float q0 = x * y;
  const float q1 = __builtin_fmaf(__builtin_fmaf(-q0, 3.0f, x), y, q0);
  return (q0 == 0.0f || __builtin_isinf(q0)) ? q0 : q1;
}

// div_count without a branch: the n = 3 quotient through third_quotient.
__device__ inline float div_count_bf(float x, int n) {
  const float p2 = x * __uint_as_float(0x3F800000u - ((uint32_t)(n >> 1) << 23));
  return n == 3 ? third_quotient(x) : p2;
}

// acc + t when `on`, else acc + -0 (which is acc, bit for bit, for every
// acc): the branch-free form of `if (on) acc = acc + t`.
__device__ inline float add_if(float acc, bool on, float t) { return acc + (on ? t : -0.0f); }

// get_general_velocity_x (src/fluid.cu:479-539).  valid(i, j) is
// Fluid::is_valid_fluid, val(i, j) reads u.
template <bool CS1, class Valid, class Val>
__device__ inline float sample_u(int cs, float x, float y, const Valid& valid, const Val& val) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  if (!valid(i, j)) return 0.0f;
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  float acc = 0.0f;
  const bool low = CS1 ? (in_y <= 0.5f) : ((double)in_y <= h);
  if (low) {
    const float d_y = CS1 ? 0.5f - in_y : (float)(h - (double)in_y);
    const float w_x = 1.0f - div_cs<CS1>(in_x, cs);
    const float w_y = 1.0f - div_cs<CS1>(d_y, cs);
    acc = acc + (w_y * w_x) * val(i, j);
    if (valid(i + 1, j)) acc = acc + (w_y * (1.0f - w_x)) * val(i + 1, j);
    if (valid(i, j - 1)) acc = acc + ((1.0f - w_y) * w_x) * val(i, j - 1);
    if (valid(i + 1, j - 1)) acc = acc + ((1.0f - w_y) * (1.0f - w_x)) * val(i + 1, j - 1);
  } else {
    const float d_y = CS1 ? in_y - 0.5f : (float)((double)in_y - h);
    const float w_x = 1.0f - div_cs<CS1>(in_x, cs);
    const float w_y = 1.0f - div_cs<CS1>(d_y, cs);
    acc = acc + (w_y * w_x) * val(i, j);
    if (valid(i, j + 1)) acc = acc + ((1.0f - w_y) * w_x) * val(i, j + 1);
    if (valid(i + 1, j)) acc = acc + (w_y * (1.0f - w_x)) * val(i + 1, j);
    if (valid(i + 1, j + 1)) acc = acc + ((1.0f - w_y) * (1.0f - w_x)) * val(i + 1, j + 1);
  }
  return acc;
}

// get_general_velocity_y (src/fluid.cu:418-477); strict `<` on x.
template <bool CS1, class Valid, class Val>
__device__ inline float sample_v(int cs, float x, float y, const Valid& valid, const Val& val) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  if (!valid(i, j)) return 0.0f;
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  float acc = 0.0f;
  const bool left = CS1 ? (in_x < 0.5f) : ((double)in_x < h);
  if (left) {
    const float d_x = CS1 ? 0.5f - in_x : (float)(h - (double)in_x);
    const float w_x = 1.0f - div_cs<CS1>(d_x, cs);
    const float w_y = 1.0f - div_cs<CS1>(in_y, cs);
    acc = acc + (w_y * w_x) * val(i, j);
    if (valid(i - 1, j)) acc = acc + (w_y * (1.0f - w_x)) * val(i - 1, j);
    if (valid(i - 1, j + 1)) acc = acc + ((1.0f - w_y) * (1.0f - w_x)) * val(i - 1, j + 1);
    if (valid(i, j + 1)) acc = acc + ((1.0f - w_y) * w_x) * val(i, j + 1);
  } else {
    const float d_x = CS1 ? in_x - 0.5f : (float)((double)in_x - h);
    const float w_x = 1.0f - div_cs<CS1>(d_x, cs);
    const float w_y = 1.0f - div_cs<CS1>(in_y, cs);
    acc = acc + (w_y * w_x) * val(i, j);
    if (valid(i, j + 1)) acc = acc + ((1.0f - w_y) * w_x) * val(i, j + 1);
    if (valid(i + 1, j + 1)) acc = acc + ((1.0f - w_y) * (1.0f - w_x)) * val(i + 1, j + 1);
    if (valid(i + 1, j)) acc = acc + (w_y * (1.0f - w_x)) * val(i + 1, j);
  }
  return acc;
}

// sample_u, sample_v and sample_smoke without branches, for accessors that
// may be called for any of the sampled cells (staged cells in LDS).  The
// terms are added in the same order as the branching forms, skipped terms
// become + -0, and the two orders of the four terms are selected, so the
// results are the same bits.
template <bool CS1, class Valid, class Val>
__device__ inline float sample_u_bf(int cs, float x, float y, const Valid& valid, const Val& val) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  const bool low = CS1 ? (in_y <= 0.5f) : ((double)in_y <= h);
  const float d_lo = CS1 ? 0.5f - in_y : (float)(h - (double)in_y);
  const float d_hi = CS1 ? in_y - 0.5f : (float)((double)in_y - h);
  const float w_x = 1.0f - div_cs<CS1>(in_x, cs);
  const float w_y = 1.0f - div_cs<CS1>(low ? d_lo : d_hi, cs);
  const int jo = low ? j - 1 : j + 1;
  const float tA = (w_y * w_x) * val(i, j);
  const float tB = (w_y * (1.0f - w_x)) * val(i + 1, j);
  const float tC = ((1.0f - w_y) * w_x) * val(i, jo);
  const float tD = ((1.0f - w_y) * (1.0f - w_x)) * val(i + 1, jo);
  const bool vB = valid(i + 1, j), vC = valid(i, jo), vD = valid(i + 1, jo);
  float acc = 0.0f + tA;
  acc = low ? add_if(acc, vB, tB) : add_if(acc, vC, tC);
  acc = low ? add_if(acc, vC, tC) : add_if(acc, vB, tB);
  acc = add_if(acc, vD, tD);
  return valid(i, j) ? acc : 0.0f;
}

template <bool CS1, class Valid, class Val>
__device__ inline float sample_v_bf(int cs, float x, float y, const Valid& valid, const Val& val) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  const bool left = CS1 ? (in_x < 0.5f) : ((double)in_x < h);
  const float d_l = CS1 ? 0.5f - in_x : (float)(h - (double)in_x);
  const float d_r = CS1 ? in_x - 0.5f : (float)((double)in_x - h);
  const float w_x = 1.0f - div_cs<CS1>(left ? d_l : d_r, cs);
  const float w_y = 1.0f - div_cs<CS1>(in_y, cs);
  const int io = left ? i - 1 : i + 1;
  const float tA = (w_y * w_x) * val(i, j);
  const float tB = (w_y * (1.0f - w_x)) * val(io, j);
  const float tC = ((1.0f - w_y) * (1.0f - w_x)) * val(io, j + 1);
  const float tD = ((1.0f - w_y) * w_x) * val(i, j + 1);
  const bool vB = valid(io, j), vC = valid(io, j + 1), vD = valid(i, j + 1);
  float acc = 0.0f + tA;
  acc = left ? add_if(acc, vB, tB) : add_if(acc, vD, tD);
  acc = add_if(acc, vC, tC);
  acc = left ? add_if(acc, vD, tD) : add_if(acc, vB, tB);
  return valid(i, j) ? acc : 0.0f;
}

// get_center_position (src/fluid.cu:547-550): ((a + 0.5) * cs) in double,
// rounded to float.  With cs == 1, (float)a + 0.5f rounds the same exact value.
template <bool CS1> __device__ inline float centre(int a, int cs) {
  if constexpr (CS1) return (float)a + 0.5f;
  else return (float)(((double)a + 0.5) * (double)cs);
}

// 1 / x in double: the hardware reciprocal refined by two Newton-Raphson
// steps r <- r + r (1 - x r), each a pair of fused multiply-adds.  For x
// converted from a finite nonzero binary32 value the relative error is at
// most 2^-50 (exhaustive check on the device: osy_arith_check, mode 1).
__device__ inline double recip_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// 1.0 / (distance + 1e-6) in double, rounded to float (src/fluid.cu:690-693),
// with get_distance (inc/helper.cuh:77-81).  The quotient is recip_nr's, an
// infinite distance giving a zero of its sign; rounded to float that equals
// the rounded IEEE quotient for every binary32 distance, NaN and infinities
// included (exhaustive check on the device: osy_arith_check, mode 0).
__device__ inline float inv_from_distance(float s) {
  const double x = (double)s + 1e-6;
  const float r = (float)recip_nr(x);
  return __builtin_isinf(x) ? __builtin_copysignf(0.0f, s) : r;
}

__device__ inline float inv_distance(float dx, float dy) {
  return inv_from_distance(sqrtf(dx * dx + dy * dy));
}

// sqrtf for t = 0 and every t >= 2^-96: the hardware square root (within
// an ulp) corrected from the fused residuals t - s' * s of its two
// neighbours s'.  Below 2^-96 the residuals are subnormal and lose the bits
// the test needs, so smaller t are not in its domain.  Equal to sqrtf on
// its domain (osy_arith_check mode 7).
__device__ inline float sqrt_staged(float t) {
  const float s = __builtin_amdgcn_sqrtf(t);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float rm = __builtin_fmaf(-sm, s, t);
  const float rp = __builtin_fmaf(-sp, s, t);
  float r = rm <= 0.0f ? sm : s;
  r = rp > 0.0f ? sp : r;
  return r;
}

// inv_distance for a sample inside the staged region.  Its distances to the
// four cell centres are a few cells at most, so the distance is finite and
// the infinite case of inv_from_distance never arises; and the squared
// distance is 0 or at least 2^-50 (a nonzero coordinate difference of two
// positions of magnitude 1/4 or more is at least 2^-25), so never
// subnormal.
__device__ inline float inv_distance_staged(float dx, float dy) {
  return (float)recip_nr((double)sqrt_staged(dx * dx + dy * dy) + 1e-6);
}

// inv / sum rounded to float (the weights of src/fluid.cu:700-712), given
// rs = recip_nr(sum): (float)(inv * rs) in double.  Exact: rs is within
// 2^-50 of 1 / sum and the product rounds once more (2^-53), so the double
// is within 2^-49 of inv / sum; a quotient of two binary32 numbers A 2^a /
// (B 2^b) (significands below 2^24) differs from every binary32 rounding
// midpoint (2M + 1) 2^(e-1) by at least 2^(e-1) / B, more than 2^-49 of
// itself, so both round to the same float.  (Mode 2 of osy_arith_check
// compares the two on 2^32 pairs.)  A zero sum (every weight 0) gives NaN,
// as the division does.
__device__ inline float idw_ratio(float inv, double rs) { return (float)((double)inv * rs); }

// interpolate_smoke (src/fluid.cu:644-716).
template <bool CS1, class Valid, class Val>
__device__ inline float sample_smoke(int cs, float x, float y, const Valid& valid,
                                     const Val& val) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  const bool left = CS1 ? (in_x < 0.5f) : ((double)in_x < h);
  const bool low = CS1 ? (in_y < 0.5f) : ((double)in_y < h);
  const int i2 = left ? i - 1 : i + 1;
  const int j3 = low ? j - 1 : j + 1;
  // cells: 1 = (i, j), 2 = (i2, j), 3 = (i, j3), 4 = (i2, j3)
  const float x1 = centre<CS1>(i, cs), y1 = centre<CS1>(j, cs);
  const float x2 = centre<CS1>(i2, cs), y3 = centre<CS1>(j3, cs);
  const float inv1 = inv_distance(x - x1, y - y1);
  const float inv2 = inv_distance(x - x2, y - y1);
  const float inv3 = inv_distance(x - x1, y - y3);
  const float inv4 = inv_distance(x - x2, y - y3);
  const float sum_inv = ((inv1 + inv2) + inv3) + inv4;
  const double rs = recip_nr((double)sum_inv);
  float acc = 0.0f;
  if (valid(i, j)) acc = acc + idw_ratio(inv1, rs) * val(i, j);
  if (valid(i2, j)) acc = acc + idw_ratio(inv2, rs) * val(i2, j);
  if (valid(i, j3)) acc = acc + idw_ratio(inv3, rs) * val(i, j3);
  if (valid(i2, j3)) acc = acc + idw_ratio(inv4, rs) * val(i2, j3);
  return acc;
}

// The branch-free samplers over a row-major LDS stage of `RX` columns whose
// cell (0, 0) is grid cell (x0, y0): one index per sample, neighbours at
// constant offsets.  `a` holds the sampled field, `f` the cell flags.  Same
// terms, order and results as sample_u_bf / sample_v_bf / sample_smoke_bf.
template <bool CS1, int RX>
__device__ inline float lds_sample_u(int cs, float x, float y, const float* a,
                                     const uint8_t* f, int x0, int y0) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  const bool low = CS1 ? (in_y <= 0.5f) : ((double)in_y <= h);
  const float d_lo = CS1 ? 0.5f - in_y : (float)(h - (double)in_y);
  const float d_hi = CS1 ? in_y - 0.5f : (float)((double)in_y - h);
  const float w_x = 1.0f - div_cs<CS1>(in_x, cs);
  const float w_y = 1.0f - div_cs<CS1>(low ? d_lo : d_hi, cs);
  const int b = (j - y0) * RX + (i - x0);
  const int bo = b + (low ? -RX : RX);
  const float tA = (w_y * w_x) * a[b];
  const float tB = (w_y * (1.0f - w_x)) * a[b + 1];
  const float tC = ((1.0f - w_y) * w_x) * a[bo];
  const float tD = ((1.0f - w_y) * (1.0f - w_x)) * a[bo + 1];
  const bool vB = f[b + 1] & CF_FLUID, vC = f[bo] & CF_FLUID, vD = f[bo + 1] & CF_FLUID;
  float acc = 0.0f + tA;
  acc = low ? add_if(acc, vB, tB) : add_if(acc, vC, tC);
  acc = low ? add_if(acc, vC, tC) : add_if(acc, vB, tB);
  acc = add_if(acc, vD, tD);
  return (f[b] & CF_FLUID) ? acc : 0.0f;
}

template <bool CS1, int RX>
__device__ inline float lds_sample_v(int cs, float x, float y, const float* a,
                                     const uint8_t* f, int x0, int y0) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  const bool left = CS1 ? (in_x < 0.5f) : ((double)in_x < h);
  const float d_l = CS1 ? 0.5f - in_x : (float)(h - (double)in_x);
  const float d_r = CS1 ? in_x - 0.5f : (float)((double)in_x - h);
  const float w_x = 1.0f - div_cs<CS1>(left ? d_l : d_r, cs);
  const float w_y = 1.0f - div_cs<CS1>(in_y, cs);
  const int b = (j - y0) * RX + (i - x0);
  const int bo = b + (left ? -1 : 1);
  const float tA = (w_y * w_x) * a[b];
  const float tB = (w_y * (1.0f - w_x)) * a[bo];
  const float tC = ((1.0f - w_y) * (1.0f - w_x)) * a[bo + RX];
  const float tD = ((1.0f - w_y) * w_x) * a[b + RX];
  const bool vB = f[bo] & CF_FLUID, vC = f[bo + RX] & CF_FLUID, vD = f[b + RX] & CF_FLUID;
  float acc = 0.0f + tA;
  acc = left ? add_if(acc, vB, tB) : add_if(acc, vD, tD);
  acc = add_if(acc, vC, tC);
  acc = left ? add_if(acc, vD, tD) : add_if(acc, vB, tB);
  return (f[b] & CF_FLUID) ? acc : 0.0f;
}

// lds_sample_u / lds_sample_v at the centre of staged cell b with cell size
// 1.  There in_x = in_y = 0.5: for u the weights are w_x = 1/2, w_y = 1 (the
// lower row pair), for v w_x = 1 (the right column pair), w_y = 1/2, so two
// terms are 1/2 * value and two are 0 * value (+-0, or NaN for an infinite
// value); they are added in the samplers' order, which gives the samplers'
// results bit for bit without computing the weights.
template <int RX>
__device__ inline float lds_centre_u(const float* a, const uint8_t* f, int b) {
  const int bo = b - RX;
  float acc = 0.0f + 0.5f * a[b];
  acc = add_if(acc, f[b + 1] & CF_FLUID, 0.5f * a[b + 1]);
  acc = add_if(acc, f[bo] & CF_FLUID, 0.0f * a[bo]);
  acc = add_if(acc, f[bo + 1] & CF_FLUID, 0.0f * a[bo + 1]);
  return (f[b] & CF_FLUID) ? acc : 0.0f;
}

template <int RX>
__device__ inline float lds_centre_v(const float* a, const uint8_t* f, int b) {
  float acc = 0.0f + 0.5f * a[b];
  acc = add_if(acc, f[b + RX] & CF_FLUID, 0.5f * a[b + RX]);
  acc = add_if(acc, f[b + 1 + RX] & CF_FLUID, 0.0f * a[b + 1 + RX]);
  acc = add_if(acc, f[b + 1] & CF_FLUID, 0.0f * a[b + 1]);
  return (f[b] & CF_FLUID) ? acc : 0.0f;
}

template <bool CS1, class Valid, class Val>
__device__ inline float sample_smoke_bf(int cs, float x, float y, const Valid& valid,
                                        const Val& val) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  const bool left = CS1 ? (in_x < 0.5f) : ((double)in_x < h);
  const bool low = CS1 ? (in_y < 0.5f) : ((double)in_y < h);
  const int i2 = left ? i - 1 : i + 1;
  const int j3 = low ? j - 1 : j + 1;
  const float x1 = centre<CS1>(i, cs), y1 = centre<CS1>(j, cs);
  const float x2 = centre<CS1>(i2, cs), y3 = centre<CS1>(j3, cs);
  const float inv1 = inv_distance(x - x1, y - y1);
  const float inv2 = inv_distance(x - x2, y - y1);
  const float inv3 = inv_distance(x - x1, y - y3);
  const float inv4 = inv_distance(x - x2, y - y3);
  const float sum_inv = ((inv1 + inv2) + inv3) + inv4;
  const double rs = recip_nr((double)sum_inv);
  float acc = 0.0f;
  acc = add_if(acc, valid(i, j), idw_ratio(inv1, rs) * val(i, j));
  acc = add_if(acc, valid(i2, j), idw_ratio(inv2, rs) * val(i2, j));
  acc = add_if(acc, valid(i, j3), idw_ratio(inv3, rs) * val(i, j3));
  acc = add_if(acc, valid(i2, j3), idw_ratio(inv4, rs) * val(i2, j3));
  return acc;
}

template <bool CS1, int RX>
__device__ inline float lds_sample_smoke(int cs, float x, float y, const float* a,
                                         const uint8_t* f, int x0, int y0) {
  const int i = to_index(div_cs<CS1>(x, cs));
  const int j = to_index(div_cs<CS1>(y, cs));
  const float in_x = x - (float)(CS1 ? i : i * cs);
  const float in_y = y - (float)(CS1 ? j : j * cs);
  const double h = cs / 2.0;
  const bool left = CS1 ? (in_x < 0.5f) : ((double)in_x < h);
  const bool low = CS1 ? (in_y < 0.5f) : ((double)in_y < h);
  const int i2 = left ? i - 1 : i + 1;
  const int j3 = low ? j - 1 : j + 1;
  const float x1 = centre<CS1>(i, cs), y1 = centre<CS1>(j, cs);
  const float x2 = centre<CS1>(i2, cs), y3 = centre<CS1>(j3, cs);
  const float inv1 = inv_distance_staged(x - x1, y - y1);
  const float inv2 = inv_distance_staged(x - x2, y - y1);
  const float inv3 = inv_distance_staged(x - x1, y - y3);
  const float inv4 = inv_distance_staged(x - x2, y - y3);
  const float sum_inv = ((inv1 + inv2) + inv3) + inv4;
  const double rs = recip_nr((double)sum_inv);
  const int b1 = (j - y0) * RX + (i - x0);
  const int b2 = b1 + (left ? -1 : 1);
  const int dj = low ? -RX : RX;
  float acc = 0.0f;
  acc = add_if(acc, f[b1] & CF_FLUID, idw_ratio(inv1, rs) * a[b1]);
  acc = add_if(acc, f[b2] & CF_FLUID, idw_ratio(inv2, rs) * a[b2]);
  acc = add_if(acc, f[b1 + dj] & CF_FLUID, idw_ratio(inv3, rs) * a[b1 + dj]);
  acc = add_if(acc, f[b2 + dj] & CF_FLUID, idw_ratio(inv4, rs) * a[b2 + dj]);
  return acc;
}

}  // namespace osy
