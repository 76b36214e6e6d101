/*
 * svs_math.h — the two elementary functions of the anisotropic weight
 * s = expf(-alpha * powf(|grad I1|, beta)) (anisotropicDifferentialOps.cu:123),
 * written once and compiled both for the GPU kernels and for the CPU oracle so
 * that both round identically.
 *
 * Method (textbook, evaluated in IEEE double with no fused multiply-add):
 *   log x: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(z),
 *          z = (m - 1)/(m + 1), |z| <= 0.1716, atanh by its Taylor series to
 *          z^25; e ln 2 with ln 2 split into a 32-bit head and a tail
 *          (Cody & Waite, "Software Manual for the Elementary Functions", 1980).
 *   exp x: k = nearest integer to x / ln 2, r = (x - k ln2_hi) - k ln2_lo,
 *          |r| <= 0.35, e^r by its Taylor series to r^16, times 2^k.
 * The float results round the double value once, so they are faithful (at
 * most one float ulp from the true value; tested against glibc and libm in
 * double).  Only x > 0 reaches svs_powf (the caller guards |grad| > 0).
 */
#ifndef SVS_MATH_H
#define SVS_MATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define SVS_HD __host__ __device__ inline
#else
#define SVS_HD static inline
#endif

#define SVS_LN2_HI 6.93147180369123816490e-01 /* 32 leading bits of ln 2 */
#define SVS_LN2_LO 1.90821492927058770002e-10 /* ln 2 - SVS_LN2_HI */
#define SVS_INV_LN2 1.44269504088896338700e+00

/* 2^k for integer k in [-1074, 1023] built from the bit pattern (exact). */
SVS_HD double svs_exp2i(int k) {
  uint64_t bits;
  if (k >= -1022) {
    bits = (uint64_t)(k + 1023) << 52;
  } else {
    bits = (uint64_t)1 << (k + 1074); /* subnormal */
  }
  double r;
  memcpy(&r, &bits, sizeof(r));
  return r;
}

/* e^x in double for the range the weight needs; 0 below -745, +inf above 709.78. */
SVS_HD double svs_exp_d(double x) {
  if (x != x) return x;
  if (x > 709.78) return __builtin_huge_val();
  if (x < -745.2) return 0.0;
  const double kd = x * SVS_INV_LN2;
  const int k = (int)(kd < 0.0 ? kd - 0.5 : kd + 0.5);
  const double r = (x - (double)k * SVS_LN2_HI) - (double)k * SVS_LN2_LO;
  /* Taylor series of e^r, Horner form, degree 16. */
  double p = 1.0 / 20922789888000.0;              /* 1/16! */
  p = p * r + 1.0 / 1307674368000.0;               /* 1/15! */
  p = p * r + 1.0 / 87178291200.0;
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  /* Scale in two steps so that k near the ends of the range stays exact. */
  const int k1 = k / 2;
  return (p * svs_exp2i(k1)) * svs_exp2i(k - k1);
}

/* ln x in double for finite x > 0. */
SVS_HD double svs_log_d(double x) {
  uint64_t bits;
  memcpy(&bits, &x, sizeof(bits));
  int e = (int)((bits >> 52) & 0x7ff);
  if (e == 0) { /* subnormal: normalise by 2^54 */
    x = x * 18014398509481984.0;
    memcpy(&bits, &x, sizeof(bits));
    e = (int)((bits >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  bits = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL; /* m in [1, 2) */
  double m;
  memcpy(&m, &bits, sizeof(m));
  if (m > 1.41421356237309504880) {
    m = m * 0.5;
    e += 1;
  }
  const double z = (m - 1.0) / (m + 1.0);
  const double z2 = z * z;
  /* 2 atanh(z) = 2 z (1 + z^2/3 + z^4/5 + ... + z^24/25), Horner in z^2. */
  double s = 1.0 / 25.0;
  s = s * z2 + 1.0 / 23.0;
  s = s * z2 + 1.0 / 21.0;
  s = s * z2 + 1.0 / 19.0;
  s = s * z2 + 1.0 / 17.0;
  s = s * z2 + 1.0 / 15.0;
  s = s * z2 + 1.0 / 13.0;
  s = s * z2 + 1.0 / 11.0;
  s = s * z2 + 1.0 / 9.0;
  s = s * z2 + 1.0 / 7.0;
  s = s * z2 + 1.0 / 5.0;
  s = s * z2 + 1.0 / 3.0;
  s = s * z2 + 1.0;
  const double logm = 2.0 * z * s;
  return ((double)e * SVS_LN2_HI + logm) + (double)e * SVS_LN2_LO;
}

/* Faithful float e^x. */
SVS_HD float svs_expf(float x) { return (float)svs_exp_d((double)x); }

/* Faithful float x^y for x > 0. */
SVS_HD float svs_powf(float x, float y) { return (float)svs_exp_d((double)y * svs_log_d((double)x)); }

#endif /* SVS_MATH_H */
