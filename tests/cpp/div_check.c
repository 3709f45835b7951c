/* Host twin of the tracer's division-free quotient, remainder and weight
 * (pars_amd/csrc/pa_device.hpp: div_unsigned, div_corrected, rem_located,
 * rem_key, weight_located, exact_fmod), checked against IEEE division and
 * fmodf.  Compiled without contraction so every operation is the single
 * rounding the device performs. */
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

/* 1/h as y + yl, as the field set-up computes it. */
static void recip(float h, float* y, float* yl) {
  *y = 1.0f / h;
  *yl = (float)(1.0 / (double)h - (double)*y);
}

static float div_unsigned(float a, float h, float y, float yl) {
  const float q0 = fmaf(a, y, a * yl);
  return fmaf(fmaf(-q0, h, a), y, q0);
}

static float div_corrected(float a, float h, float y, float yl) { return copysignf(div_unsigned(a, h, y, yl), a); }

static float rem_located(float d, float h, float m) {
  const float r0 = fmaf(-m, h, d);
  return r0 < 0.0f ? r0 + h : r0;
}

static int rem_range(float r) { return to_bits(r) - 1u >= 0x2B800000u - 1u; }

static float exact_fmod(float d, float h, float q) {
  const float ax = fabsf(d);
  float n = floorf(fabsf(q));
  const float r = fmaf(-n, h, ax);
  n += (float)((int)(r >= h) - (int)(r < 0.0f));
  return copysignf(fmaf(-n, h, ax), d);
}

/* Mismatches of the corrected quotient over a = from_bits(lo), lo + step, ...
 * < hi (both signs); zero quotients must keep their sign. */
uint64_t quotient_mismatches(float h, uint32_t lo, uint32_t hi, uint32_t step) {
  float y, yl;
  recip(h, &y, &yl);
  uint64_t bad = 0;
  for (uint64_t u = lo; u < hi; u += step)
    for (int s = 0; s < 2; ++s) {
      const float a = from_bits((uint32_t)u | (s ? 0x80000000u : 0u));
      if (to_bits(div_corrected(a, h, y, yl)) != to_bits(a / h)) ++bad;
      /* the unsigned form differs at most in the sign of a zero */
      const float qu = div_unsigned(a, h, y, yl);
      if (!(qu == a / h)) ++bad;
    }
  return bad;
}

/* Mismatches of the weight path for d = from_bits(lo) ... < hi, d >= 0 with
 * d / h < 2^22 (a located point): the remainder from m = floor(q) against
 * fmodf (up to the sign of a zero), the range key against its definition,
 * and the weight against fmodf(d, h) / h bit for bit where the key admits
 * it. */
uint64_t weight_mismatches(float h, uint32_t lo, uint32_t hi, uint32_t step) {
  float y, yl;
  recip(h, &y, &yl);
  uint64_t bad = 0;
  for (uint64_t u = lo; u < hi; u += step) {
    const float d = from_bits((uint32_t)u);
    const float q = div_unsigned(d, h, y, yl);
    if (!(q < 4194304.0f)) continue;
    const float r = rem_located(d, h, floorf(q));
    const float want = fmodf(d, h);
    if (!(r == want)) ++bad;
    const int in_range = want == 0.0f || fabsf(want) >= 0x1p-40f;
    if (rem_range(r) != in_range) ++bad;
    if (in_range && to_bits(copysignf(div_unsigned(r, h, y, yl), d)) != to_bits(want / h)) ++bad;
  }
  return bad;
}

/* Mismatches of the fallback fmod (exact_fmod) for d = from_bits(lo) ... < hi
 * with d / h < 2^22. */
uint64_t fmod_mismatches(float h, uint32_t lo, uint32_t hi, uint32_t step) {
  uint64_t bad = 0;
  for (uint64_t u = lo; u < hi; u += step) {
    const float d = from_bits((uint32_t)u);
    const float q = d / h;
    if (!(q < 4194304.0f)) continue;
    if (to_bits(exact_fmod(d, h, q)) != to_bits(fmodf(d, h))) ++bad;
  }
  return bad;
}

/* Signed zeros and the range key at its edges: 0 when all hold. */
uint64_t edge_mismatches(float h) {
  float y, yl;
  recip(h, &y, &yl);
  uint64_t bad = 0;
  const float zs[2] = {0.0f, -0.0f};
  for (int i = 0; i < 2; ++i) {
    const float d = zs[i];
    const float q = div_unsigned(d, h, y, yl);
    const float r = rem_located(d, h, floorf(q));
    if (!rem_range(r)) ++bad;
    if (to_bits(copysignf(div_unsigned(r, h, y, yl), d)) != to_bits(fmodf(d, h) / h)) ++bad;
    if (to_bits(div_corrected(d, h, y, yl)) != to_bits(d / h)) ++bad;
  }
  if (rem_range(0x1p-41f) || !rem_range(0x1p-40f) || rem_range(from_bits(1u)) || !rem_range(0.5f)) ++bad;
  return bad;
}
