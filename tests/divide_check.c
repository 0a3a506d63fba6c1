/* Counts the floats a of a range whose quotient a/b differs, in any bit,
   from the two-FMA form q0 = a*y, r = fma(-q0,b,a), q1 = fma(r,y,q0) with
   y = 1/b rounded to float (the form the fine and coarse kick kernels use
   for b = 6 and b = pi in real(4)).  Built with gcc -O2 -mfma
   -ffp-contract=off by tests/test_divide_const.py. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* bit patterns [lo, hi) of positive floats, both signs; skip_fix drops the
   correction step (q1 = q0), for the test that the count can see a fault */
int64_t count_bad(float b, uint32_t lo, uint32_t hi, int skip_fix) {
  volatile float bv = b;
  const float y = 1.0f / bv;
  int64_t bad = 0;
  for (uint32_t u = lo; u < hi; ++u) {
    for (int s = 0; s < 2; ++s) {
      const float a = s ? -f_of(u) : f_of(u);
      const float q = a / bv;
      const float q0 = a * y;
      const float r = fmaf(-q0, bv, a);
      const float q1 = skip_fix ? q0 : fmaf(r, y, q0);
      bad += u_of(q1) != u_of(q);
    }
  }
  return bad;
}
