/*
 * Exhaustive check of the quotient the tiled projection kernel computes.
 *
 * The kernel needs q = RN(div / s) for s = total_s in {0, 1, 2, 3, 4}.
 *  - s in {1, 2, 4}: q = RN(div * 2^-l2) with l2 = s >> 1, y built from its
 *    bit pattern 0x3F800000 - (l2 << 23).
 *  - s in {0, 3}: the kernel divides (a rarely taken branch).
 * It also checks the fused one-step correction q3 = fma(fma(-q0, 3, div),
 * RN(1/3), q0) with q0 = RN(div * RN(1/3)), guarded to return q0 when q0 is
 * zero or infinite: a cheaper exact s = 3 path the kernel may use.
 *
 * Every one of the 2^32 binary32 values of div is run through each path and
 * compared bit for bit with div / s (any NaN equals any NaN).  Prints
 * "mismatches <path> <count>"; exit status 0 when all counts are 0.
 * Built and run by tests/test_quotient.py (gcc -O2 -mfma -ffp-contract=off
 * -fopenmp; fmaf is the fused IEEE operation).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static inline float from_bits(uint32_t b) {
  float f;
  memcpy(&f, &b, 4);
  return f;
}
static inline uint32_t to_bits(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  return b;
}

static inline float pow2_quotient(float div, int s) {
  const uint32_t l2 = (uint32_t)s >> 1;
  const float y = from_bits(0x3F800000u - (l2 << 23));
  return div * y;
}

static inline float third_quotient(float div) {
  const float y = 1.0f / 3.0f;
  const float q0 = div * y;
  if (q0 == 0.0f || isinf(q0)) return q0;
  return fmaf(fmaf(-q0, 3.0f, div), y, q0);
}

static int same(float a, float b) {
  if (isnan(a) && isnan(b)) return 1;
  return to_bits(a) == to_bits(b);
}

int main(void) {
  const int paths[4] = {1, 2, 4, 3};
  int fail = 0;
  for (int pi = 0; pi < 4; pi++) {
    const int s = paths[pi];
    const float sf = (float)s;
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long k = 0; k < (1LL << 32); k++) {
      const float div = from_bits((uint32_t)k);
      const float got = (s == 3) ? third_quotient(div) : pow2_quotient(div, s);
      if (!same(div / sf, got)) bad++;
    }
    printf("mismatches s=%d %lld\n", s, bad);
    if (bad) fail = 1;
  }
  return fail;
}
