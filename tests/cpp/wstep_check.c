/* Host twin of the w-component shortcut of the frozen-derivative steps
 * (pars_amd/csrc/trace.hip: zero_terms, used as w + w_rk4 and w + w_euler in
 * pa_device.hpp step_frozen): w + a_0 0 + a_1 0 + ... (left to right) against
 * w + zero_terms(a).  Compiled without contraction. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static uint32_t to_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

static float zero_terms(const float* a, int n) {
  int all_neg = 1;
  for (int i = 0; i < n; ++i) {
    const float t = a[i] * 0.0f;
    if (t != t) return NAN;
    if (!signbit(t)) all_neg = 0;
  }
  return all_neg ? -0.0f : 0.0f;
}

/* Mismatches over w in ws[] and the RK4 (b = 1/6, 1/3, 1/3, 1/6) and Euler
 * coefficients of each dt in dts[]: bit for bit, or both NaN. */
uint64_t wstep_mismatches(const float* ws, int nw, const float* dts, int nd) {
  uint64_t bad = 0;
  for (int j = 0; j < nd; ++j) {
    const float dt = dts[j];
    const float a[4] = {(1.0f / 6.0f) * dt, (1.0f / 3.0f) * dt, (1.0f / 3.0f) * dt, (1.0f / 6.0f) * dt};
    const float z4 = zero_terms(a, 4), z1 = zero_terms(&dt, 1);
    for (int i = 0; i < nw; ++i) {
      const float w = ws[i];
      const float full = (((w + a[0] * 0.0f) + a[1] * 0.0f) + a[2] * 0.0f) + a[3] * 0.0f;
      const float fast = w + z4;
      if (!(isnan(full) && isnan(fast)) && to_bits(full) != to_bits(fast)) ++bad;
      const float e_full = w + dt * 0.0f, e_fast = w + z1;
      if (!(isnan(e_full) && isnan(e_fast)) && to_bits(e_full) != to_bits(e_fast)) ++bad;
    }
  }
  return bad;
}
