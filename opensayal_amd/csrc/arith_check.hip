// arith_check.hip - exhaustive on-device checks of the arithmetic shortcuts
// the smoke advection's inverse-distance weights and square roots, the
// 16-bit storage modes' paired roundings and the pressure sums take
// (fluid_device.hpp, projection_device.hpp):
// each mode runs every one of the 2^32 binary32 bit patterns through the
// shortcut and through the IEEE operations it replaces, on the GPU that
// runs the solver, and counts the inputs where they differ.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "fluid_device.hpp"
#include "fluid_internal.hpp"

#pragma clang fp contract(off)

namespace osy {
namespace {

constexpr unsigned kCheckThreads = 256;
constexpr unsigned kCheckBlocks = 4096;

// The reference's form: 1.0 / (s + 1e-6) in double, rounded to float.
__device__ __attribute__((noinline)) float inv_ieee(float s) {
  return (float)(1.0 / ((double)s + 1e-6));
}

__device__ __attribute__((noinline)) float div_ieee(float a, float b) { return a / b; }

__device__ __attribute__((noinline)) float sqrt_ieee(float a) { return sqrtf(a); }

// A numerator for denominator pattern b: b scaled by a factor in (0, 1]
// from a hash of the pattern (an inverse-distance weight never exceeds the
// weights' sum), or a special value for a few patterns.
__device__ inline float numerator(uint32_t bits, float b) {
  uint32_t h = bits * 0x9E3779B1u;
  h ^= h >> 15;
  h *= 0x85EBCA77u;
  h ^= h >> 13;
  switch (h & 1023u) {
    case 0: return 0.0f;
    case 1: return b;
    case 2: return __builtin_nanf("");
    default: break;
  }
  const float u = (float)((h >> 8) | 1u) * 0x1p-24f;  // (0, 1)
  return b * u;
}

__device__ inline void record(unsigned long long* count, uint32_t* first, uint32_t bits) {
  atomicAdd(count, 1ULL);
  atomicMin(first, bits);
}

__global__ void __launch_bounds__(kCheckThreads)
    k_arith_check(int mode, unsigned long long* count, uint32_t* first) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < (1ULL << 32);
       v += stride) {
    const uint32_t bits = (uint32_t)v;
    const float x = __uint_as_float(bits);
    bool bad = false;
    if (mode == 0) {
      bad = __float_as_uint(inv_from_distance(x)) != __float_as_uint(inv_ieee(x));
    } else if (mode == 1) {
      // relative error of recip_nr for finite nonzero x: the residual
      // 1 - x r, computed with one rounding, within 2^-50
      if (x != 0.0f && !__builtin_isinf(x) && !__builtin_isnan(x)) {
        const double d = (double)x;
        const double e = __builtin_fma(-d, recip_nr(d), 1.0);
        bad = !(__builtin_fabs(e) <= 0x1p-50);
      }
    } else if (mode == 7) {
      // sqrt_staged against sqrtf on 0 and every finite input >= 2^-96
      if (x == 0.0f || (x >= 0x1p-96f && x <= 0x1.fffffep127f))
        bad = __float_as_uint(sqrt_staged(x)) != __float_as_uint(sqrt_ieee(x));
    } else if (mode == 5 || mode == 6) {
      // the pressure sums' quotient x / d_t as a double multiply by the
      // double nearest 1 / d_t, for the step sizes 0.05 and 1/60
      const float b = mode == 5 ? 0.05f : 1.0f / 60.0f;
      const float q1 = (float)((double)x * (1.0 / (double)b));
      bad = __float_as_uint(q1) != __float_as_uint(div_ieee(x, b));
    } else if (mode == 3 || mode == 4) {
      // rnd2 (one packed conversion) against rnd (one scalar conversion per
      // value) for the storage type, each pattern in both halves of a pair
      const float y = __uint_as_float(bits ^ 0x9E3779B9u);
      float a = x, b = y, c = y, d = x;
      float ra, rb, rc, rd;
      if (mode == 3) {
        rnd2<__half>(a, b);
        rnd2<__half>(c, d);
        ra = rnd<__half>(x), rb = rnd<__half>(y), rc = rb, rd = ra;
      } else {
        rnd2<__hip_bfloat16>(a, b);
        rnd2<__hip_bfloat16>(c, d);
        ra = rnd<__hip_bfloat16>(x), rb = rnd<__hip_bfloat16>(y), rc = rb, rd = ra;
      }
      bad = __float_as_uint(a) != __float_as_uint(ra) || __float_as_uint(b) != __float_as_uint(rb) ||
            __float_as_uint(c) != __float_as_uint(rc) || __float_as_uint(d) != __float_as_uint(rd);
    } else {
      // the weights' quotient: positive finite sums (and 0) with a
      // numerator no larger; NaN compared as NaN
      if (!(x >= 0.0f) || __builtin_isinf(x)) continue;
      const float a = numerator(bits, x);
      const float q0 = div_ieee(a, x);
      const float q1 = idw_ratio(a, recip_nr((double)x));
      bad = __builtin_isnan(q0) ? !__builtin_isnan(q1)
                                : __float_as_uint(q0) != __float_as_uint(q1);
    }
    if (bad) record(count, first, bits);
  }
}

}  // namespace
}  // namespace osy

using namespace osy;

extern "C" int osy_arith_check(int32_t mode, uint64_t* mismatches, uint32_t* first_input) {
  if (!mismatches || !first_input) return set_error(OSY_E_INVALID, "null argument");
  if (mode < 0 || mode > 7) return set_error(OSY_E_INVALID, "mode must be 0 to 7");
  void* buf = nullptr;
  OSY_HIP_CHECK(hipMalloc(&buf, 16));
  unsigned long long* count = (unsigned long long*)buf;
  uint32_t* first = (uint32_t*)((char*)buf + 8);
  const uint32_t init[4] = {0u, 0u, 0xFFFFFFFFu, 0u};
  hipError_t e = hipMemcpy(buf, init, 16, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_arith_check, dim3(kCheckBlocks), dim3(kCheckThreads), 0, nullptr,
                       (int)mode, count, first);
    e = hipGetLastError();
  }
  uint32_t out[4] = {0, 0, 0, 0};
  if (e == hipSuccess) e = hipMemcpy(out, buf, 16, hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess)
    return set_error(OSY_E_HIP, std::string("arith check: ") + hipGetErrorString(e));
  *mismatches = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
  *first_input = out[2];
  return OSY_OK;
}
