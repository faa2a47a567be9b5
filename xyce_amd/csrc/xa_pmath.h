/* xa_pmath.h -- portable exp, log, pow, tan and sin for the parity build.
 *
 * The device kernels normally call the GPU math library (ocml) and the CPU
 * loads the C library (glibc); the two differ in the last bits of exp, log,
 * pow and tan, so a GPU load and a CPU load of the same instance agree only
 * to the conditioning of the model around each call.  Compiled into both
 * sides instead of the system functions (the library's parity build defines
 * XA_PORTABLE_MATH, the CPU restatement's defines OR_PORTABLE_MATH), these
 * functions give both the same result for the same argument: each is a fixed
 * sequence of IEEE-754 double operations (+ - * /, floor, frexp, ldexp, all
 * correctly rounded or exact on both sides), compiled with floating-point
 * contraction off.  Then every output of a load must match bit for bit.
 *
 * Method (Cody & Waite, "Software Manual for the Elementary Functions",
 * Prentice-Hall 1980): additive range reduction by a constant split into a
 * high part with trailing zero bits (so that k * high is exact) and a low
 * part, then a truncated series on the reduced argument.
 *   exp(x) = 2^k exp(r), k = nearest integer to x / ln 2, r = x - k ln 2,
 *            |r| <= ln(2)/2, exp(r) by its Taylor series to r^13 (Horner).
 *   log(x) = e ln 2 + log(m), x = m 2^e with sqrt(1/2) <= m < sqrt(2);
 *            log(m) = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716,
 *            written log(1 + f) = f - s (f - z P(z)), z = s^2, with the odd
 *            series of atanh to s^23.
 *   pow(x, y) = exp(y log x) with the special values of C99 F.9.4.4.
 *   sin(x) and tan(x): sin(r), cos(r) (signed by k mod 4) and
 *            sin(r) / cos(r) or -cos(r) / sin(r) for k even or odd,
 *            k = nearest integer to x / (pi/2), r = x - k pi/2 (pi/2 in three
 *            parts), |r| <= pi/4, sine to r^19 and cosine to r^20 by Taylor
 *            series.
 * The split constants and series coefficients were computed for this file
 * to 60 digits (ln 2 and pi from their series) and rounded to double.
 * Accuracy against the C library, measured over the ranges the loads use:
 * exp and log within 1 ulp, tan and sin within 3 ulps for |x| <= 20 (sin
 * absolute 1e-16 near its zeros), pow within
 * 2 + 2 |y log x| ulps (the rounding of log x and of y log x, carried
 * through exp) (tests/test_pmath.py).  They are a parity tool, not
 * a replacement for the system functions in production builds.
 */
#ifndef XA_PMATH_H
#define XA_PMATH_H

#include <math.h>

#if defined(__HIPCC__)
#define XA_PM_FN static __host__ __device__ inline
#else
#define XA_PM_FN static inline
#endif

/* ln 2 = XA_PM_LN2_HI + XA_PM_LN2_LO; the high part has 32 significant bits */
#define XA_PM_LN2_HI 0.6931471803691238
#define XA_PM_LN2_LO 1.9082149292705877e-10
#define XA_PM_INV_LN2 1.4426950408889634
/* pi/2 = P1 + P2 + P3; P1 and P2 have 33 significant bits */
#define XA_PM_PIO2_1 1.5707963267341256
#define XA_PM_PIO2_2 6.077100506303966e-11
#define XA_PM_PIO2_3 2.0222662487959506e-21
#define XA_PM_TWO_OVER_PI 0.6366197723675814
#define XA_PM_SQRT_HALF 0.7071067811865476

XA_PM_FN double xa_pm_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.1332191019412) return 0.0;
  const double kd = floor(x * XA_PM_INV_LN2 + 0.5);
  const int k = (int)kd;
  const double r = (x - kd * XA_PM_LN2_HI) - kd * XA_PM_LN2_LO;
  double p = 1.6059043836821613e-10;        /* 1/13! */
  p = p * r + 2.08767569878681e-09;         /* 1/12! */
  p = p * r + 2.505210838544172e-08;
  p = p * r + 2.755731922398589e-07;
  p = p * r + 2.7557319223985893e-06;
  p = p * r + 2.48015873015873e-05;
  p = p * r + 0.0001984126984126984;
  p = p * r + 0.001388888888888889;
  p = p * r + 0.008333333333333333;
  p = p * r + 0.041666666666666664;
  p = p * r + 0.16666666666666666;
  p = p * r + 0.5;
  p = p * r;                                /* exp(r) - 1 - r = r^2 (1/2 + ...) */
  p = p * r + r;
  return ldexp(1.0 + p, k);
}

XA_PM_FN double xa_pm_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return (x - x) / 0.0;        /* NaN */
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int e = 0;
  double m = frexp(x, &e);                  /* x = m 2^e, 0.5 <= m < 1 (exact) */
  if (m < XA_PM_SQRT_HALF) {
    m = m + m;
    e = e - 1;
  }
  const double f = m - 1.0;                 /* exact: 0.7 < m < 1.42 */
  const double s = f / (2.0 + f);
  const double z = s * s;
  double P = 0.08695652173913043;           /* 2/23 */
  P = P * z + 0.09523809523809523;          /* 2/21 */
  P = P * z + 0.10526315789473684;          /* 2/19 */
  P = P * z + 0.11764705882352941;          /* 2/17 */
  P = P * z + 0.13333333333333333;          /* 2/15 */
  P = P * z + 0.15384615384615385;          /* 2/13 */
  P = P * z + 0.18181818181818182;          /* 2/11 */
  P = P * z + 0.2222222222222222;           /* 2/9 */
  P = P * z + 0.2857142857142857;           /* 2/7 */
  P = P * z + 0.4;                          /* 2/5 */
  P = P * z + 0.6666666666666666;           /* 2/3 */
  P = P * z;
  const double lm = f - s * (f - P);        /* 2 atanh(s) = f - s f + s P */
  const double ed = (double)e;
  return ed * XA_PM_LN2_HI + (ed * XA_PM_LN2_LO + lm);
}

/* y is an odd integer */
XA_PM_FN int xa_pm_odd(double y) {
  if (floor(y) != y) return 0;
  const double h = y * 0.5;
  return floor(h) != h;
}

XA_PM_FN double xa_pm_pow(double x, double y) {
  if (y == 0.0 || x == 1.0) return 1.0;
  if (x != x || y != y) return x + y;
  if (x == 0.0) {
    if (y > 0.0) return xa_pm_odd(y) ? x : 0.0;
    return xa_pm_odd(y) ? (x < 0.0 || 1.0 / x < 0.0 ? -HUGE_VAL : HUGE_VAL) : HUGE_VAL;
  }
  double sign = 1.0;
  if (x < 0.0) {
    if (floor(y) != y) return (x - x) / 0.0;   /* NaN: negative base, non-integer power */
    if (xa_pm_odd(y)) sign = -1.0;
    x = -x;
  }
  if (x == HUGE_VAL) return sign * (y > 0.0 ? HUGE_VAL : 0.0);
  return sign * xa_pm_exp(y * xa_pm_log(x));
}

/* sin(r) and cos(r) for |r| <= pi/4 by their Taylor series (to r^19 and
   r^20): s = r + r z S(z), c = 1 - z/2 + z^2 C(z), z = r^2 */
XA_PM_FN void xa_pm_sincos_reduced(double r, double* sn, double* cs) {
  const double z = r * r;
  /* sin(r) = r + r z S(z), cos(r) = 1 - z/2 + z^2 C(z) */
  double S = -8.22063524662433e-18;         /* -1/19! */
  S = S * z + 2.8114572543455206e-15;       /* 1/17! */
  S = S * z - 7.647163731819816e-13;        /* -1/15! */
  S = S * z + 1.6059043836821613e-10;       /* 1/13! */
  S = S * z - 2.505210838544172e-08;        /* -1/11! */
  S = S * z + 2.7557319223985893e-06;       /* 1/9! */
  S = S * z - 0.0001984126984126984;        /* -1/7! */
  S = S * z + 0.008333333333333333;         /* 1/5! */
  S = S * z - 0.16666666666666666;          /* -1/3! */
  *sn = r + r * (z * S);
  double Cc = 4.110317623312165e-19;        /* 1/20! */
  Cc = Cc * z - 1.5619206968586225e-16;     /* -1/18! */
  Cc = Cc * z + 4.779477332387385e-14;      /* 1/16! */
  Cc = Cc * z - 1.1470745597729725e-11;     /* -1/14! */
  Cc = Cc * z + 2.08767569878681e-09;       /* 1/12! */
  Cc = Cc * z - 2.755731922398589e-07;      /* -1/10! */
  Cc = Cc * z + 2.48015873015873e-05;       /* 1/8! */
  Cc = Cc * z - 0.001388888888888889;       /* -1/6! */
  Cc = Cc * z + 0.041666666666666664;       /* 1/4! */
  *cs = (1.0 - 0.5 * z) + (z * z) * Cc;
}

XA_PM_FN double xa_pm_tan(double x) {
  if (x != x || x == HUGE_VAL || x == -HUGE_VAL) return (x - x) / 0.0;
  const double kd = floor(x * XA_PM_TWO_OVER_PI + 0.5);
  const double r = ((x - kd * XA_PM_PIO2_1) - kd * XA_PM_PIO2_2) - kd * XA_PM_PIO2_3;
  double sn, cs;
  xa_pm_sincos_reduced(r, &sn, &cs);
  const double kh = kd * 0.5;
  return floor(kh) == kh ? sn / cs : -cs / sn;
}

/* sin(x) = sin(r), cos(r), -sin(r), -cos(r) for k = 0, 1, 2, 3 (mod 4) */
XA_PM_FN double xa_pm_sin(double x) {
  if (x != x || x == HUGE_VAL || x == -HUGE_VAL) return (x - x) / 0.0;
  const double kd = floor(x * XA_PM_TWO_OVER_PI + 0.5);
  const double r = ((x - kd * XA_PM_PIO2_1) - kd * XA_PM_PIO2_2) - kd * XA_PM_PIO2_3;
  double sn, cs;
  xa_pm_sincos_reduced(r, &sn, &cs);
  const double q = kd - 4.0 * floor(kd * 0.25);   /* k mod 4, exact */
  if (q == 0.0) return sn;
  if (q == 1.0) return cs;
  if (q == 2.0) return -sn;
  return -cs;
}

#endif /* XA_PMATH_H */
