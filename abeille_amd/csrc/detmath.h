// detmath.h - deterministic double-precision log, sin and cos.
//
// The transport path needs log (flight distances, exponential sampling) and
// sin/cos (azimuthal angles) to give the same bits on the x86-64 host and on
// gfx950.  Vendor libms differ by an ulp here and there, so both sides use
// these routines, built only from IEEE +,-,*,/ and integer bit operations.
// Every translation unit that includes this header must be compiled with
// -ffp-contract=off so that no multiply-add is fused.
//
// Methods (published, re-derived here; coefficients are plain Taylor terms):
//  * exp: Cody-Waite reduction by ln 2 and the rational form
//    1 + 2r / (2 - c) of fdlibm with the Bernoulli series of r coth(r/2).
//  * log: argument reduction x = 2^k * m, m in [sqrt(2)/2, sqrt(2)); with
//    f = m - 1 and s = f / (2 + f), log(m) = 2 atanh(s) evaluated as
//    f - hfsq + s (hfsq + R(s^2)), hfsq = f^2/2, R(z) = sum_{n>=1} 2 z^n/(2n+1)
//    (the Sun fdlibm / Tang formulation of the atanh series), the sum in
//    Estrin's form.  ln 2 is split
//    into a 32-bit head and a tail so that k * ln2_hi is exact.
//  * sin/cos: Cody-Waite reduction by pi/2 with a three-term split of pi/2,
//    then Taylor polynomials on |r| <= pi/4 (through r^19 for sin and r^20
//    for cos, in Estrin's form) with the compensated 1 - r^2/2 step for cos.
// Accuracy is checked against glibc by tests/test_detmath.py (<= 1 ulp over
// the ranges the transport path uses).
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define DM_HD __host__ __device__ __forceinline__
#if defined(ABX_OUTLINE) && ABX_OUTLINE
#define DM_NOINL __host__ __device__ __attribute__((noinline))
#else
#define DM_NOINL __host__ __device__ __forceinline__
#endif
#else
#define DM_HD static inline
#define DM_NOINL static
#endif

DM_HD uint64_t dm_bits(double x) {
  union {
    double d;
    uint64_t u;
  } v;
  v.d = x;
  return v.u;
}

DM_HD double dm_from_bits(uint64_t u) {
  union {
    double d;
    uint64_t u;
  } v;
  v.u = u;
  return v.d;
}

// Natural logarithm.  Returns -inf for +-0, NaN for negative or NaN input and
// +inf for +inf.
DM_HD double dm_log(double x) {
  const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
  const double ln2_lo = 1.90821492927058770002e-10; /* 0x3dea39ef35793c76 */
  const double two54 = 18014398509481984.0;         /* 2^54 */

  uint64_t b = dm_bits(x);
  if (x != x) return x;
  if (x <= 0.0) {
    if (x == 0.0) return -1.0 / 0.0;
    return (x - x) / 0.0;
  }
  if (b == 0x7ff0000000000000ULL) return x;

  int k = 0;
  if ((b >> 52) == 0) {  // subnormal: scale into the normal range
    x *= two54;
    b = dm_bits(x);
    k -= 54;
  }
  k += (int)((b >> 52) & 0x7ff) - 1023;
  // mantissa in [1, 2)
  uint64_t mb = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  // sqrt(2) = 0x3ff6a09e667f3bcd
  if (mb > 0x3ff6a09e667f3bcdULL) {
    mb -= 0x0010000000000000ULL;  // m /= 2
    k += 1;
  }
  const double m = dm_from_bits(mb);
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  // R(z) = 2z/3 + 2z^2/5 + ... + 2z^12/25, in Estrin's form (pairs of
  // terms joined by powers z^2, z^4, z^8): eight dependent steps instead of
  // Horner's twenty-four.
  const double z2 = z * z, z4 = z2 * z2;
  const double p01 = 2.0 / 3.0 + z * (2.0 / 5.0), p23 = 2.0 / 7.0 + z * (2.0 / 9.0);
  const double p45 = 2.0 / 11.0 + z * (2.0 / 13.0), p67 = 2.0 / 15.0 + z * (2.0 / 17.0);
  const double p89 = 2.0 / 19.0 + z * (2.0 / 21.0), pab = 2.0 / 23.0 + z * (2.0 / 25.0);
  const double q0 = p01 + z2 * p23, q1 = p45 + z2 * p67, q2 = p89 + z2 * pab;
  const double R = z * (q0 + z4 * (q1 + z4 * q2));
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}

// Exponential.  Reduction x = k ln2 + r, |r| <= ln2 / 2 (Cody-Waite, ln2 in
// the same head / tail split as dm_log), then the rational form
// exp(r) = 1 + 2r / (2 - c) with c = r - r^2 P(r^2), where r^2 P is the
// series of r coth(r/2) - 2 (2 B_2n / (2n)!: 1/6, -1/360, 1/15120,
// -1/604800, 1/23950080, -691/653837184000, 1/37362124800; the next term is
// below 1e-19 for |r| <= ln2 / 2), kept in the head / tail split of r so that the
// rounding of r itself does not enter (the Sun fdlibm formulation); 2^k is
// applied through the exponent bits, in two steps near the underflow.
// Returns +inf above the overflow threshold and +0 below the underflow one.
DM_HD double dm_exp(double x) {
  const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
  const double ln2_lo = 1.90821492927058770002e-10; /* 0x3dea39ef35793c76 */
  const double inv_ln2 = 1.44269504088896338700e+00;
  if (x != x) return x;
  if (x > 709.782712893383973096) return 1.0 / 0.0;
  if (x < -745.13321910194110842) return 0.0;
  const double t = x * inv_ln2;
  const double fk = t >= 0.0 ? (double)(int64_t)(t + 0.5) : -(double)(int64_t)(0.5 - t);
  const int k = (int)fk;
  const double hi = x - fk * ln2_hi;
  const double lo = fk * ln2_lo;
  const double r = hi - lo;
  const double z = r * r;
  const double c =
      r - z * (1.0 / 6.0 +
               z * (-1.0 / 360.0 +
                    z * (1.0 / 15120.0 +
                         z * (-1.0 / 604800.0 +
                              z * (1.0 / 23950080.0 +
                                   z * (-691.0 / 653837184000.0 + z * (1.0 / 37362124800.0)))))));
  const double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
  if (k > 1023) return (y * dm_from_bits((uint64_t)(1023 + 1023) << 52)) * 2.0;
  if (k >= -1021) return y * dm_from_bits((uint64_t)(k + 1023) << 52);
  // subnormal result: scale by 2^(k + 1000), then by 2^-1000
  return (y * dm_from_bits((uint64_t)(k + 1000 + 1023) << 52)) *
         dm_from_bits((uint64_t)(-1000 + 1023) << 52);
}

// sin and cos of r, |r| <= pi/4 (Taylor kernels).
// Both polynomials are evaluated in Estrin's form (pairs of terms joined by
// z^2, z^4, z^8) for a short dependency chain.
DM_HD double dm_ksin(double r) {
  const double z = r * r;
  const double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
  // P(z) = -1/3! + z/5! - ... - z^8/19!
  const double p01 = -1.0 / 6.0 + z * (1.0 / 120.0);
  const double p23 = -1.0 / 5040.0 + z * (1.0 / 362880.0);
  const double p45 = -1.0 / 39916800.0 + z * (1.0 / 6227020800.0);
  const double p67 = -1.0 / 1307674368000.0 + z * (1.0 / 355687428096000.0);
  const double p8 = -1.0 / 121645100408832000.0;
  const double p = (p01 + z2 * p23) + z4 * (p45 + z2 * p67) + z8 * p8;
  return r + (r * z) * p;
}

DM_HD double dm_kcos(double r) {
  const double z = r * r;
  // C(z) = 1/4! - z/6! + ... + z^8/20!
  const double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
  const double p01 = 1.0 / 24.0 + z * (-1.0 / 720.0);
  const double p23 = 1.0 / 40320.0 + z * (-1.0 / 3628800.0);
  const double p45 = 1.0 / 479001600.0 + z * (-1.0 / 87178291200.0);
  const double p67 = 1.0 / 20922789888000.0 + z * (-1.0 / 6402373705728000.0);
  const double p8 = 1.0 / 2432902008176640000.0;
  const double p = (p01 + z2 * p23) + z4 * (p45 + z2 * p67) + z8 * p8;
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + (z * z) * p);
}

// sin and cos together.  Accurate for |x| up to about 1e5 (the reduction
// uses a three-part Cody-Waite split of pi/2); the transport path only
// calls it on [0, 2 pi].
struct DmSinCos {
  double s, c;
};

// Returned by value (registers) so that the out-of-line device version
// needs no stack.
DM_NOINL DmSinCos dm_sincos2(double x) {
  double sv, cv;
  double* sn = &sv;
  double* cs = &cv;
  const double invpio2 = 6.36619772367581382433e-01; /* 2/pi */
  const double pio2_1 = 1.57079632673412561417e+00;  /* first 33 bits */
  const double pio2_2 = 6.07710050630396597660e-11;  /* next 33 bits */
  const double pio2_3 = 2.02226624879595063154e-21;  /* next 33 bits */
  const double pio4 = 7.85398163397448278999e-01;
  double ax = x < 0.0 ? -x : x;
  if (!(ax == ax) || ax == 1.0 / 0.0) {  // NaN or inf
    return DmSinCos{x - x, x - x};
  }
  if (ax <= pio4) {
    return DmSinCos{dm_ksin(x), dm_kcos(x)};
  }
  // n = nearest integer to x * 2/pi (ties away from zero)
  double t = x * invpio2;
  double fn = t >= 0.0 ? (double)(int64_t)(t + 0.5) : -(double)(int64_t)(0.5 - t);
  const double r = ((x - fn * pio2_1) - fn * pio2_2) - fn * pio2_3;
  const int64_t n = (int64_t)fn;
  const double s = dm_ksin(r);
  const double c = dm_kcos(r);
  switch ((int)(n & 3)) {
    case 0:
      *sn = s;
      *cs = c;
      break;
    case 1:
      *sn = c;
      *cs = -s;
      break;
    case 2:
      *sn = -s;
      *cs = -c;
      break;
    default:
      *sn = -c;
      *cs = s;
      break;
  }
  return DmSinCos{sv, cv};
}

DM_HD void dm_sincos(double x, double* sn, double* cs) {
  const DmSinCos r = dm_sincos2(x);
  *sn = r.s;
  *cs = r.c;
}

DM_HD double dm_sin(double x) {
  double s, c;
  dm_sincos(x, &s, &c);
  return s;
}

DM_HD double dm_cos(double x) {
  double s, c;
  dm_sincos(x, &s, &c);
  return c;
}

// Single-precision helpers of the split-precision path.  On the device they
// map to the hardware's fast FP32 instructions (log2 / sin / cos / rcp /
// sqrt with ~1 ulp error); on the host to libm's float functions.  Results
// differ between the two in the last bits, which the split path tolerates.
#if defined(__HIP_DEVICE_COMPILE__)
DM_HD float fp_log(float x) { return __logf(x); }
DM_HD void fp_sincos(float x, float* s, float* c) { __sincosf(x, s, c); }
DM_HD float fp_div(float a, float b) { return __fdividef(a, b); }
DM_HD float fp_rcp(float x) { return __frcp_rn(x); }
DM_HD float fp_sqrt(float x) { return __fsqrt_rn(x); }
DM_HD float fp_rsqrt(float x) { return __frsqrt_rn(x); }
#else
DM_HD float fp_log(float x) { return logf(x); }
DM_HD void fp_sincos(float x, float* s, float* c) {
  *s = sinf(x);
  *c = cosf(x);
}
DM_HD float fp_div(float a, float b) { return a / b; }
DM_HD float fp_rcp(float x) { return 1.0f / x; }
DM_HD float fp_sqrt(float x) { return sqrtf(x); }
DM_HD float fp_rsqrt(float x) { return 1.0f / sqrtf(x); }
#endif
