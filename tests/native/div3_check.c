/* Checks the divisions by w = 3 and 6 that event detection takes for the
 * t-statistics of tame reads (abea_getevents.hip, div3k and div3f):
 *   q = fma(fma(-q0, d, x), y, q0),  q0 = x * y,  y = RN(1/d)
 * (1) in double (div3k), against IEEE double division, on doubles that
 *     are sums of floats: all 2^24 signed significands at several exponents
 *     and random ones over the whole float range;
 * (2) a float taken to double, divided as in (1), rounded to float, against
 *     IEEE float division: all signed significands at the smallest normal, a
 *     middle and the largest exponent, and all subnormals.  (The same
 *     comparison over all 2^32 floats differs only for -0, whose quotient
 *     comes out +0.)
 * (3) in float (div3f), against IEEE float division: for d = 3 all signed
 *     significands at the subnormal and the lowest, middle and highest
 *     exponents; for d = 6 the same at the exponents of magnitudes from
 *     2^-125 up.  (Over all 2^32 floats, d = 3 differs only for -0; d = 6
 *     also differs for some magnitudes below 2^-125, whose quotients are
 *     subnormal, and those are not taken this way.)
 * Zero quotients compare equal whatever their sign.  Exit status 0 = no
 * mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float bf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t fb(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint64_t db(double f) { uint64_t u; memcpy(&u, &f, 8); return u; }
static uint64_t rng = 0x2545F4914F6CDD1Dull;
static uint64_t next(void) { rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17; return rng; }

static double div3k(double x, double d, double y) {
  double q0 = x * y;
  return fma(fma(-q0, d, x), y, q0);
}

static float div3f(float x, float d, float y) {
  float q0 = x * y;
  return fmaf(fmaf(-q0, d, x), y, q0);
}

int main(void) {
  const double ds[2] = {3.0, 6.0};
  long bad = 0, tested = 0;
  for (int di = 0; di < 2; ++di) {
    const double d = ds[di];
    volatile double yv = 1.0 / d;
    const double y = yv;
    const float df = (float)d;
    /* (2) float route */
    const uint32_t exps[] = {0, 1, 127, 254};
    for (unsigned e = 0; e < sizeof exps / sizeof exps[0]; ++e)
      for (uint32_t m = 0; m < (1u << 23); ++m)
        for (uint32_t s = 0; s < 2; ++s) {
          const float x = bf((s << 31) | (exps[e] << 23) | m);
          volatile float ref = x / df;
          const float q = (float)div3k((double)x, d, y);
          ++tested;
          if (fb(q) != fb(ref) && !(q == 0.0f && ref == 0.0f)) {
            if (bad < 5) printf("float route: d=%g x=%a got %a want %a\n", d, x, q, ref);
            ++bad;
          }
        }
    /* (3) float route: every exponent listed, both signs, all significands;
     * for d = 6 only magnitudes from 2^-125 up (biased exponent 2) */
    volatile float yfv = 1.0f / df;
    const float yf = yfv;
    const uint32_t fexps3[] = {0, 1, 2, 3, 64, 127, 128, 200, 253, 254};
    const uint32_t fexps6[] = {2, 3, 4, 64, 127, 128, 200, 253, 254};
    const uint32_t *fe = di == 0 ? fexps3 : fexps6;
    const unsigned nfe = di == 0 ? sizeof fexps3 / sizeof fexps3[0] : sizeof fexps6 / sizeof fexps6[0];
    for (unsigned e = 0; e < nfe; ++e)
      for (uint32_t m = 0; m < (1u << 23); ++m)
        for (uint32_t s = 0; s < 2; ++s) {
          const float x = bf((s << 31) | (fe[e] << 23) | m);
          volatile float ref = x / df;
          const float q = div3f(x, df, yf);
          ++tested;
          if (fb(q) != fb(ref) && !(q == 0.0f && ref == 0.0f)) {
            if (bad < 5) printf("float route (div3f): d=%g x=%a got %a want %a\n", d, x, q, ref);
            ++bad;
          }
        }
    /* (1) double route on doubles that are sums of a few floats */
    for (long i = 0; i < 20000000; ++i) {
      const uint64_t r = next();
      const float a = bf((uint32_t)(r & 0xffffffffu) & 0xff7fffffu);  /* finite */
      const float b = bf((uint32_t)(r >> 32) & 0xff7fffffu);
      const double x = (i & 1) ? (double)a + (double)b : (double)a * 3.0 + (double)b;
      if (!isfinite(x)) continue;
      volatile double ref = x / d;
      const double q = div3k(x, d, y);
      ++tested;
      if (db(q) != db(ref) && !(q == 0.0 && ref == 0.0)) {
        if (bad < 5) printf("double route: d=%g x=%a got %a want %a\n", d, x, q, ref);
        ++bad;
      }
    }
    for (int e = -149; e < 128; e += 7)
      for (uint32_t m = 0; m < (1u << 24); m += 3) {
        const double x = ldexp((double)((1u << 24) | m), e - 24);
        volatile double ref = x / d;
        const double q = div3k(x, d, y);
        ++tested;
        if (db(q) != db(ref)) {
          if (bad < 5) printf("double route: d=%g x=%a got %a want %a\n", d, x, q, ref);
          ++bad;
        }
      }
  }
  printf("%ld tested, %ld mismatches\n", tested, bad);
  return bad != 0;
}
