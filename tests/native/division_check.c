/* Checks the quotient the core kernel uses for the emission,
 *   q0 = n * rs,  r = fma(-q0, s, n),  q = fma(r, rs, q0),  rs = RN(1/s),
 * against the IEEE quotient RN(n / s) the reference computes
 * (GPU/align_cuda.cl:527).  Exhaustive over the 2^23 significands of n at the
 * given exponents for a set of divisors (worst-case significands included),
 * plus random numerators for the divisors read from a model file.
 * Exit status 0 = no mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) { rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17; return rng; }

static long check(float s, const int *exps, int n_exps, long n_random, long *tested) {
  volatile float rs = 1.0f / s;
  long bad = 0;
  for (int x = 0; x < n_exps; ++x)
    for (uint32_t m = 0; m < (1u << 23); ++m)
      for (int sg = 0; sg < 2; ++sg) {
        float n = bits_f(((uint32_t)sg << 31) | ((uint32_t)exps[x] << 23) | m);
        float q0 = n * rs, q = fmaf(fmaf(-q0, s, n), rs, q0);
        volatile float ref = n / s;
        if (f_bits(q) != f_bits(ref)) { if (bad < 3) printf("mismatch s=%a n=%a\n", s, n); ++bad; }
        ++*tested;
      }
  for (long i = 0; i < n_random; ++i) {
    uint64_t r = next();
    float n = bits_f((uint32_t)((r >> 63) << 31) | ((107u + (uint32_t)((r >> 40) % 38)) << 23) |
                     (uint32_t)(r & 0x7fffff));
    float q0 = n * rs, q = fmaf(fmaf(-q0, s, n), rs, q0);
    volatile float ref = n / s;
    if (f_bits(q) != f_bits(ref)) { if (bad < 3) printf("mismatch s=%a n=%a\n", s, n); ++bad; }
    ++*tested;
  }
  return bad;
}

int main(int argc, char **argv) {
  long tested = 0, bad = 0;
  const int exps[] = {127, 140};  /* n in [1,2) and [8192,16384) */
  const float divisors[] = {bits_f(0x3fffffff), bits_f(0x3f7fffff), bits_f(0x407fffff), 2.0f,
                            bits_f(0x3f800001), 1.5f, 2.5f, 5.74f, 1.22f};
  for (unsigned i = 0; i < sizeof divisors / sizeof divisors[0]; ++i)
    bad += check(divisors[i], exps, 2, 0, &tested);
  if (argc > 1) {
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    static float model[4096 * 3];
    size_t k = fread(model, 4, 4096 * 3, f);
    fclose(f);
    for (size_t i = 0; i < k / 3; ++i) bad += check(model[3 * i + 1], exps, 0, 2000, &tested);
  }
  /* The emission tail: the kernel computes fma(RN(a*a), -0.5, c) for the
   * reference's c + RN(RN(-0.5a) * a) (GPU/align_cuda.cl:528).  Equal
   * whenever 0.5a^2 is a normal float (halving commutes with rounding), and
   * for smaller |a| unless c is exactly 0. */
  long tail_tested = 0, tail_bad = 0;
  for (long i = 0; i < 20000000; ++i) {
    uint64_t r = next();
    /* |a| in [2^-60, 2^60), either sign */
    float a = bits_f((uint32_t)((r >> 63) << 31) | ((67u + (uint32_t)((r >> 40) % 120)) << 23) |
                     (uint32_t)(r & 0x7fffff));
    uint64_t r2 = next();
    /* c in [-2^5, 2^5) excluding the exact zero, plus values near 0 */
    float c = bits_f((uint32_t)((r2 >> 63) << 31) | ((100u + (uint32_t)((r2 >> 40) % 32)) << 23) |
                     (uint32_t)(r2 & 0x7fffff));
    volatile float h = -0.5f * a;
    volatile float t = h * a;
    volatile float ref = c + t;
    float got = fmaf(a * a, -0.5f, c);
    if (f_bits(got) != f_bits(ref)) {
      if (tail_bad < 3) printf("tail mismatch a=%a c=%a\n", a, c);
      ++tail_bad;
    }
    ++tail_tested;
  }
  const float specials[] = {0.0f, -0.0f, INFINITY, -INFINITY, 1e-30f, 3.5f};
  for (unsigned i = 0; i < 6; ++i)
    for (unsigned j = 0; j < 6; ++j) {
      float a = specials[i], c = specials[j] == 0.0f ? -1.75f : -specials[j];
      volatile float h = -0.5f * a;
      volatile float t = h * a;
      volatile float ref = c + t;
      float got = fmaf(a * a, -0.5f, c);
      if (f_bits(got) != f_bits(ref) && !(isnan(got) && isnan(ref))) ++tail_bad;
      ++tail_tested;
    }
  printf("tested %ld emission tails, %ld tail mismatches\n", tail_tested, tail_bad);
  bad += tail_bad;
  printf("tested %ld quotients, %ld mismatches\n", tested, bad);
  return bad != 0;
}
