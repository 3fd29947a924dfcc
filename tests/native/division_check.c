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
  printf("tested %ld quotients, %ld mismatches\n", tested, bad);
  return bad != 0;
}
