/*
 * wg_libm.h -- deterministic exp, log and pow in FP64 for host and device.
 *
 * The reference calls the C library's exp and pow (daily.cpp:251,423,859,1126;
 * routing.cpp:1940-1942, 2426, 2479, 2728, 3477-3502, 3561, 7300-7338); which C
 * library is not pinned.  glibc on the host and the ROCm device library differ in
 * the last bit now and then, and the day loop amplifies such bits where a cell's
 * land area fraction reaches zero (storages are rescaled by its reciprocal,
 * daily.cpp:837, 947, 1096).  Both the kernels and the CPU oracle therefore use
 * these functions, written from IEEE basic operations and fused multiply-add only
 * (each correctly rounded on every target), so the two produce identical bits.
 *
 * Method (Tang's table-driven exponential, Newton-refined logarithm):
 *   exp:  x = (64k + j) ln2/64 + r, |r| <= ln2/128, with ln2/64 in three parts;
 *         e^x = 2^k 2^(j/64) (1 + p(r)), 2^(j/64) tabulated as double-double,
 *         p(r) = r + r^2 (1/2 + r/6 + ... + r^6/8!); error about 2^-67 relative
 *         before the final rounding, so results are correctly rounded except in
 *         rare near-midpoint cases.
 *   log:  y0 from the atanh series of the mantissa, then one Newton step with the
 *         double-double exp: log x = y0 + t - t^2/2, t = x e^(-y0) - 1.
 *   pow:  exp(y log x) with the product carried as a double-double.
 *   sin/cos of 2 pi u (the ensemble's Box-Muller normals): quadrant split and
 *         Taylor series, see wg_sincos2pi.
 * Agreement with glibc is checked by tests/test_libm.py.
 */
#ifndef WG_LIBM_H
#define WG_LIBM_H

#include "wg_libm_tables.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define WG_FN __host__ __device__ static inline __attribute__((always_inline))
#else
#define WG_FN static inline
#endif

#ifdef __cplusplus
#include <cmath>
#define WG_FMA std::fma
#define WG_FLOOR std::floor
#define WG_FREXP std::frexp
#define WG_LDEXP std::ldexp
#else
#include <math.h>
#define WG_FMA fma
#define WG_FLOOR floor
#define WG_FREXP frexp
#define WG_LDEXP ldexp
#endif

static const double wg_exp2_tab_h[128] = {WG_EXP2_TAB_VALUES};
#if defined(__HIPCC__) || defined(__HIP__)
__constant__ static const double wg_exp2_tab_d[128] = {WG_EXP2_TAB_VALUES};
#endif
#if defined(__HIP_DEVICE_COMPILE__) && defined(WG_EXP2_TAB_LDS)
/* the including file declares the LDS copy, which every kernel that calls these
 * functions fills from wg_exp2_tab_d before its first call */
#define WG_EXP2_TAB WG_EXP2_TAB_LDS
#elif defined(__HIP_DEVICE_COMPILE__)
#define WG_EXP2_TAB wg_exp2_tab_d
#else
#define WG_EXP2_TAB wg_exp2_tab_h
#endif

typedef struct wg_dd {
  double hi, lo;
} wg_dd;

WG_FN wg_dd wg_two_sum(double a, double b) {
  wg_dd r;
  double s = a + b, bb = s - a;
  r.hi = s;
  r.lo = (a - (s - bb)) + (b - bb);
  return r;
}

WG_FN wg_dd wg_fast_two_sum(double a, double b) {
  wg_dd r;
  double s = a + b;
  r.hi = s;
  r.lo = b - (s - a);
  return r;
}

WG_FN wg_dd wg_two_prod(double a, double b) {
  wg_dd r;
  double p = a * b;
  r.hi = p;
  r.lo = WG_FMA(a, b, -p);
  return r;
}

/* e^(xh + xl) as a double-double; xh in [-745.2, 709.8], |xl| <= ulp(xh). */
WG_FN wg_dd wg_exp_dd(double xh, double xl) {
  const double kk = WG_FLOOR(xh * WG_INV_LN2_64 + 0.5);
  const int ik = (int)kk;
  const int j = ik & 63;
  const int k = (ik - j) / 64;
  const double r1 = xh - kk * WG_LN2_64_1; /* exact: kk has < 18 bits, the constant 36 */
  wg_dd a = wg_two_sum(r1, -(kk * WG_LN2_64_2));
  const double rl = a.lo + (xl - kk * WG_LN2_64_3);
  wg_dd r = wg_fast_two_sum(a.hi, rl);
  const double rh = r.hi;
  const double q =
      rh * rh *
      (0.5 + rh * (1.0 / 6.0 + rh * (1.0 / 24.0 + rh * (1.0 / 120.0 + rh * (1.0 / 720.0 +
                                                                             rh * (1.0 / 5040.0 + rh * (1.0 / 40320.0)))))));
  wg_dd p = wg_fast_two_sum(rh, r.lo + q);
  const double th = WG_EXP2_TAB[2 * j], tl = WG_EXP2_TAB[2 * j + 1];
  wg_dd tp = wg_two_prod(th, p.hi);
  const double lo = tl + (tp.lo + (th * p.lo + tl * p.hi));
  wg_dd s = wg_fast_two_sum(th, tp.hi);
  wg_dd res = wg_fast_two_sum(s.hi, s.lo + lo);
  res.hi = WG_LDEXP(res.hi, k);
  res.lo = WG_LDEXP(res.lo, k);
  return res;
}

WG_FN double wg_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return 1.0 / 0.0;
  if (x < -745.1332191019412) return 0.0;
  wg_dd r = wg_exp_dd(x, 0.0);
  return r.hi + r.lo;
}

/* log(x) as a double-double for finite x > 0. */
WG_FN wg_dd wg_log_dd(double x) {
  int e;
  double m = WG_FREXP(x, &e);
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  const double y0 =
      e * WG_LN2_HI +
      2.0 * s *
          (1.0 + s2 * (1.0 / 3.0 + s2 * (1.0 / 5.0 + s2 * (1.0 / 7.0 + s2 * (1.0 / 9.0 + s2 * (1.0 / 11.0 +
                                                                                              s2 * (1.0 / 13.0 + s2 * (1.0 / 15.0))))))));
  wg_dd E = wg_exp_dd(-y0, 0.0);
  const double t = WG_FMA(x, E.hi, -1.0) + x * E.lo;
  return wg_fast_two_sum(y0, t - 0.5 * t * t);
}

WG_FN double wg_log(double x) {
  if (x != x || x < 0.0) return 0.0 / 0.0;
  if (x == 0.0) return -1.0 / 0.0;
  if (x == 1.0 / 0.0) return x;
  if (x == 1.0) return 0.0;
  wg_dd l = wg_log_dd(x);
  return l.hi + l.lo;
}

/* x^y for the arguments the day loop uses (x >= 0); negative x with integral y
 * follows the sign rule, otherwise NaN. */
WG_FN double wg_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (x != x || y != y) return x + y;
  if (x == 1.0) return 1.0;
  double sign = 1.0;
  if (x < 0.0) {
    if (WG_FLOOR(y) != y) return 0.0 / 0.0;
    const double half = 0.5 * y;
    if (WG_FLOOR(half) != half) sign = -1.0;
    x = -x;
  }
  if (x == 0.0) return y > 0.0 ? 0.0 : sign / 0.0;
  if (x == 1.0 / 0.0) return y > 0.0 ? sign * x : 0.0;
  wg_dd l = wg_log_dd(x);
  wg_dd z = wg_two_prod(y, l.hi);
  z = wg_fast_two_sum(z.hi, z.lo + y * l.lo);
  if (z.hi > 709.782712893384) return sign / 0.0;
  if (z.hi < -745.1332191019412) return sign * 0.0;
  wg_dd r = wg_exp_dd(z.hi, z.lo);
  return sign * (r.hi + r.lo);
}

/* sin(2 pi u) and cos(2 pi u) for u in [0, 1), from IEEE additions and
 * multiplications only (no fused multiply-add), so that host and device produce
 * the same bits.  4u = q + r splits off the quadrant q exactly; the angle within
 * the quadrant is folded to x = (pi/2) min(r, 1 - r) <= pi/4, where the Taylor
 * series of sin and cos, to x^17 and x^18, are below 1e-19 of their sums.
 * Absolute error about 2e-16. */
WG_FN void wg_sincos2pi(double u, double *s, double *c) {
  const double t = 4.0 * u;
  const double qf = WG_FLOOR(t);
  const int q = ((int)qf) & 3;
  const double r = t - qf;
  const int flip = r > 0.5;
  const double rr = flip ? 1.0 - r : r;
  const double x = rr * 1.5707963267948966;
  const double x2 = x * x;
  const double sx =
      x * (1.0 + x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0 +
           x2 * (-1.0 / 39916800.0 + x2 * (1.0 / 6227020800.0 + x2 * (-1.0 / 1307674368000.0 +
           x2 * (1.0 / 355687428096000.0)))))))));
  const double cx =
      1.0 + x2 * (-0.5 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0 + x2 * (-1.0 / 3628800.0 +
            x2 * (1.0 / 479001600.0 + x2 * (-1.0 / 87178291200.0 + x2 * (1.0 / 20922789888000.0 +
            x2 * (-1.0 / 6402373705728000.0)))))))));
  const double sa = flip ? cx : sx, ca = flip ? sx : cx;
  switch (q) {
    case 0: *s = sa; *c = ca; break;
    case 1: *s = ca; *c = -sa; break;
    case 2: *s = -sa; *c = -ca; break;
    default: *s = -ca; *c = sa; break;
  }
}

#endif
