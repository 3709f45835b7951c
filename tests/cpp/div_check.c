/* Host twin of the tracer's division-free quotient and branch-free fmod
 * (pars_amd/csrc/pa_device.hpp: div_corrected, exact_fmod), checked against
 * IEEE division and fmodf.  Compiled without contraction so every operation
 * is the single rounding the device performs. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static float from_bits(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t to_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

static float div_corrected(float a, float h, float y) {
  const float q0 = a * y;
  const float q1 = fmaf(fmaf(-q0, h, a), y, q0);
  const float q2 = fmaf(fmaf(-q1, h, a), y, q1);
  return copysignf(q2, a);
}

static float fmod_branch_free(float d, float h, float q) {
  const float ax = fabsf(d);
  float n = floorf(fabsf(q));
  const float r = fmaf(-n, h, ax);
  n += (float)((int)(r >= h) - (int)(r < 0.0f));
  return copysignf(fmaf(-n, h, ax), d);
}

/* Mismatches of the corrected quotient over a = from_bits(lo), lo + step, ...
 * < hi (both signs). */
uint64_t quotient_mismatches(float h, uint32_t lo, uint32_t hi, uint32_t step) {
  const float y = 1.0f / h;
  uint64_t bad = 0;
  for (uint64_t u = lo; u < hi; u += step)
    for (int s = 0; s < 2; ++s) {
      const float a = from_bits((uint32_t)u | (s ? 0x80000000u : 0u));
      if (to_bits(div_corrected(a, h, y)) != to_bits(a / h)) ++bad;
    }
  return bad;
}

/* Mismatches of the branch-free fmod for d = from_bits(lo) ... < hi with
 * d / h < 2^22. */
uint64_t fmod_mismatches(float h, uint32_t lo, uint32_t hi, uint32_t step) {
  uint64_t bad = 0;
  for (uint64_t u = lo; u < hi; u += step) {
    const float d = from_bits((uint32_t)u);
    const float q = d / h;
    if (!(q < 4194304.0f)) continue;
    if (to_bits(fmod_branch_free(d, h, q)) != to_bits(fmodf(d, h))) ++bad;
  }
  return bad;
}
