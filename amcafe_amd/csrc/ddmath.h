// ddmath.h -- correctly rounded FP64 atan2 / sin / cos for the capture geometry.
//
// The CA capture test (S/VoxelsCA.C:415-424, 548-551, 591) evaluates
//     th = atan2(|dy|,|dx|)  folded to [0, pi/4]
//     ph = atan2(sqrt(dx^2+dy^2), |dz|)  folded to [pi/4, pi/2]
//     f  = sqrt(cos(th) * sin(ph))
// Each function here computes its result in double-double arithmetic (relative
// error below 2^-100) and returns the nearest double, i.e. the correctly rounded
// value except when the exact result lies within 2^-100 of a rounding midpoint.
// sqrt and division are IEEE operations on the device and need no help.
//
// Arguments are restricted to what the capture geometry produces:
//   amc_atan2_q1(y, x):  y >= 0, x >= 0, finite
//   amc_sin_q1 / amc_cos_q1(x):  0 <= x <= pi/2 (fast path when the result is >= 1/2)
// Error-free transformations use explicit fma(); the file must be compiled with
// -ffp-contract=off so the compiler does not fuse anything else.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include "ddtables.h"

struct amc_dd {
  double hi, lo;
};

__host__ __device__ static inline amc_dd amc_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
__host__ __device__ static inline amc_dd amc_fast_two_sum(double a, double b) {  // |a| >= |b|
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}
__host__ __device__ static inline amc_dd amc_two_prod(double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  return {p, e};
}
__host__ __device__ static inline amc_dd amc_dd_add(amc_dd x, amc_dd y) {
  amc_dd s = amc_two_sum(x.hi, y.hi);
  amc_dd t = amc_two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = amc_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return amc_fast_two_sum(s.hi, s.lo);
}
__host__ __device__ static inline amc_dd amc_dd_neg(amc_dd x) { return {-x.hi, -x.lo}; }
__host__ __device__ static inline amc_dd amc_dd_mul(amc_dd x, amc_dd y) {
  amc_dd p = amc_two_prod(x.hi, y.hi);
  p.lo += x.hi * y.lo + x.lo * y.hi;
  return amc_fast_two_sum(p.hi, p.lo);
}
__host__ __device__ static inline amc_dd amc_dd_mul_d(amc_dd x, double y) {
  amc_dd p = amc_two_prod(x.hi, y);
  p.lo += x.lo * y;
  return amc_fast_two_sum(p.hi, p.lo);
}
__host__ __device__ static inline amc_dd amc_dd_div(amc_dd x, amc_dd y) {
  double q1 = x.hi / y.hi;
  amc_dd r = amc_dd_add(x, amc_dd_neg(amc_dd_mul_d(y, q1)));
  double q2 = r.hi / y.hi;
  r = amc_dd_add(r, amc_dd_neg(amc_dd_mul_d(y, q2)));
  double q3 = r.hi / y.hi;
  amc_dd q = amc_fast_two_sum(q1, q2);
  return amc_dd_add(q, amc_dd{q3, 0.0});
}
__host__ __device__ static inline amc_dd amc_tab(const double (*t)[2], int i) { return {t[i][0], t[i][1]}; }

// Horner evaluation of sum_k c[k] * z^k in double-double.
__host__ __device__ static inline amc_dd amc_poly(const double (*c)[2], int n, amc_dd z) {
  amc_dd p = amc_tab(c, n - 1);
  for (int k = n - 2; k >= 0; --k) p = amc_dd_add(amc_dd_mul(p, z), amc_tab(c, k));
  return p;
}

// atan(q) for a double-double q in [0, 1]: table point c = i/64 plus the series of
// atan(r), r = (q - c) / (1 + q c), |r| <= 1/128 (terms through r^23).
__host__ __device__ static inline amc_dd amc_atan_dd01(amc_dd q) {
  int i = (int)rint(q.hi * AMC_DD_TAB);
  double c = (double)i / AMC_DD_TAB;
  amc_dd num = amc_dd_add(q, amc_dd{-c, 0.0});
  amc_dd den = amc_dd_add(amc_dd{1.0, 0.0}, amc_dd_mul_d(q, c));
  amc_dd r = amc_dd_div(num, den);
  amc_dd r2 = amc_dd_mul(r, r);
  amc_dd s = amc_dd_mul(r, amc_poly(kAtanCoef, 12, r2));
  return amc_dd_add(amc_tab(kAtanTab, i), s);
}

// atan2(y, x) rounded to nearest, for y >= 0, x >= 0 (both finite).
__host__ __device__ static inline double amc_atan2_dd(double y, double x) {
  if (y == 0.0) return 0.0;
  if (x == 0.0) return kPi2[0][0];
  bool swap = y > x;
  double num = swap ? x : y, den = swap ? y : x;
  amc_dd q = amc_dd_div(amc_dd{num, 0.0}, amc_dd{den, 0.0});
  amc_dd a = amc_atan_dd01(q);
  if (swap) a = amc_dd_add(amc_tab(kPi2, 0), amc_dd_neg(a));
  return a.hi + a.lo;
}

// sin(x), cos(x) rounded to nearest for 0 <= x <= pi/2: x = k/64 + r, |r| <= 1/128,
// Taylor series of sin r, cos r in double-double, then the addition formulas.
__host__ __device__ static inline void amc_sincos_dd(double x, amc_dd* s, amc_dd* c) {
  int k = (int)rint(x * AMC_DD_TAB);
  double r = x - (double)k / AMC_DD_TAB;  // exact
  amc_dd rr = {r, 0.0};
  amc_dd r2 = amc_two_prod(r, r);
  amc_dd sr = amc_dd_mul(rr, amc_poly(kSinCoef, 10, r2));
  amc_dd cr = amc_poly(kCosCoef, 10, r2);
  amc_dd sk = amc_tab(kSinTab, k), ck = amc_tab(kCosTab, k);
  if (s) *s = amc_dd_add(amc_dd_mul(sk, cr), amc_dd_mul(ck, sr));
  if (c) *c = amc_dd_add(amc_dd_mul(ck, cr), amc_dd_neg(amc_dd_mul(sk, sr)));
}
__host__ __device__ static inline double amc_sin_dd(double x) {
  amc_dd s;
  amc_sincos_dd(x, &s, nullptr);
  return s.hi + s.lo;
}
__host__ __device__ static inline double amc_cos_dd(double x) {
  amc_dd c;
  amc_sincos_dd(x, nullptr, &c);
  return c.hi + c.lo;
}

// ---------------------------------------------------------------------------------
// Fast paths with a rounding test (Ziv's strategy): an approximation hi + lo with
// relative error below 2^-65 is rounded to hi when the exact value provably lies on
// the same side of both neighbouring midpoints; otherwise the double-double routine
// above decides.  The bound used in the test, 2^-62 relative, leaves a factor 8 of
// margin; about 0.4% of calls take the slow path.
__host__ __device__ static inline bool amc_round_ok(double hi, double lo) {  // hi > 0
  const double err = hi * 0x1p-62;
  const double up = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, hi) + 1ull) - hi;
  const double dn = hi - __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, hi) - 1ull);
  return (lo + err < 0.5 * up) && (lo - err > -0.5 * dn);
}

// Fast part of atan2(y, x) for y, x >= 0 finite: hi + lo within 2^-65 relative of the
// exact value; no rounding test and no special cases (the callers select those).
__host__ __device__ static inline void amc_atan2_fast(double y, double x, double* hi_out, double* lo_out) {
  const bool swap = y > x;
  const double num = swap ? x : y, den = swap ? y : x;
  const double qh = num / den;
  const double ql = fma(-qh, den, num) / den;  // num/den = qh + ql to 2^-106
  // 0/0 (a special case the callers replace) must still give an in-range table index
  const int i = (int)rint((((qh >= 0.0) & (qh <= 1.0)) ? qh : 0.0) * AMC_DD_TAB);
  double rh, rl;
  if (i == 0) {
    rh = qh;
    rl = ql;
  } else {
    const double c = (double)i / AMC_DD_TAB;
    const double nh = qh - c;  // exact
    amc_dd p = amc_two_prod(qh, c);
    amc_dd s = amc_two_sum(1.0, p.hi);
    const double dl = s.lo + p.lo + ql * c;
    rh = nh / s.hi;
    double e = fma(-rh, s.hi, nh);
    e = e + ql - rh * dl;
    rl = e / s.hi;
  }
  // atan(r) = r - r^3/3 + ... (|r| <= 1/128; the r^13 term is below 2^-84 r)
  const double r2 = rh * rh;
  const double P = -0x1.5555555555555p-2 +
                   r2 * (0x1.999999999999ap-3 + r2 * (-0x1.2492492492492p-3 + r2 * (0x1.c71c71c71c71cp-4 +
                                                                                        r2 * -0x1.745d1745d1746p-4)));
  const double tail = rh * r2 * P - rl * r2;
  amc_dd a = amc_two_sum(kAtanTab[i][0], rh);
  double lo = a.lo + kAtanTab[i][1] + rl + tail;
  amc_dd res = amc_fast_two_sum(a.hi, lo);
  if (swap) {
    amc_dd h = amc_two_sum(kPi2[0][0], -res.hi);
    double l = h.lo + kPi2[0][1] - res.lo;
    res = amc_fast_two_sum(h.hi, l);
  }
  *hi_out = res.hi;
  *lo_out = res.lo;
}
__host__ __device__ static inline double amc_atan2_q1(double y, double x) {
  if (y == 0.0) return 0.0;
  if (x == 0.0) return kPi2[0][0];
  double hi, lo;
  amc_atan2_fast(y, x, &hi, &lo);
  if (amc_round_ok(hi, lo)) return hi;
  return amc_atan2_dd(y, x);
}

// sin / cos for 0 <= x <= pi/2 whose result is at least 1/2 (cos on [0, pi/4], sin on
// [pi/4, pi/2] in the capture test): x = k/64 + r, Taylor terms of r in double around
// the exact leading products.
__host__ __device__ static inline void amc_sincos_fast(double x, double* sh_out, double* sl_out, double* ch_out,
                                                       double* cl_out) {
  // arguments outside [0, pi/2] (from special cases the callers replace) still index the table
  const int k = (int)rint((((x >= 0.0) & (x <= 1.58)) ? x : 0.0) * AMC_DD_TAB);  // k <= 101
  const double r = x - (double)k / AMC_DD_TAB;
  const double r2 = r * r;
  const double sm = r * r2 * (-0x1.5555555555555p-3 + r2 * (0x1.1111111111111p-7 + r2 * (-0x1.a01a01a01a01ap-13 +
                                                                                        r2 * 0x1.71de3a556c734p-19)));
  const double cm = r2 * (-0.5 + r2 * (0x1.5555555555555p-5 + r2 * (-0x1.6c16c16c16c17p-10 +
                                                                    r2 * (0x1.a01a01a01a01ap-16 +
                                                                          r2 * -0x1.27e4fb7789f5cp-22))));
  const double sh = kSinTab[k][0], sl = kSinTab[k][1], ch = kCosTab[k][0], cl = kCosTab[k][1];
  if (sh_out) {
    amc_dd p = amc_two_prod(ch, r);
    amc_dd s = amc_two_sum(sh, p.hi);
    double lo = s.lo + p.lo + sl + cl * r + (sh * cm + ch * sm);
    amc_dd res = amc_fast_two_sum(s.hi, lo);
    *sh_out = res.hi;
    *sl_out = res.lo;
  }
  if (ch_out) {
    amc_dd p = amc_two_prod(-sh, r);
    amc_dd s = amc_two_sum(ch, p.hi);
    double lo = s.lo + p.lo + cl - sl * r + (ch * cm - sh * sm);
    amc_dd res = amc_fast_two_sum(s.hi, lo);
    *ch_out = res.hi;
    *cl_out = res.lo;
  }
}
__host__ __device__ static inline double amc_sin_q1(double x) {
  double h, l;
  amc_sincos_fast(x, &h, &l, nullptr, nullptr);
  if (h >= 0.5 && amc_round_ok(h, l)) return h;
  return amc_sin_dd(x);
}
__host__ __device__ static inline double amc_cos_q1(double x) {
  double h, l;
  amc_sincos_fast(x, nullptr, nullptr, &h, &l);
  if (h >= 0.5 && amc_round_ok(h, l)) return h;
  return amc_cos_dd(x);
}

// Shape factor of the decentred-octahedron capture test, S/VoxelsCA.C:415-418, 424:
//   pow(cos(th) * sin(ph), .5) with th, ph from the centre-to-centroid vector.
// The operation order (products, sum, folds) follows the reference exactly; the
// libm calls are the correctly rounded versions above.
// All four fast parts run as one branch-free sequence (the two atan2 chains are
// independent and overlap); only when some rounding test fails, or a special case
// applies, is the whole value recomputed call by call with the slow paths.
__host__ __device__ static inline double amc_shape_factor_ref(double dx, double dy, double dz) {
  const double pi4 = 0x1.921fb54442d18p-1, pi2 = 0x1.921fb54442d18p+0;  // M_PI/4.0, M_PI/2.0
  double th = amc_atan2_q1(fabs(dy), fabs(dx));
  th > pi4 ? th = pi2 - th : th;
  double rho = sqrt(dx * dx + dy * dy);
  double ph = amc_atan2_q1(rho, fabs(dz));
  ph < pi4 ? ph = pi2 - ph : ph;
  return sqrt(amc_cos_q1(th) * amc_sin_q1(ph));
}
__host__ __device__ static inline __attribute__((always_inline)) double amc_shape_factor(double dx, double dy, double dz) {
  const double pi4 = 0x1.921fb54442d18p-1, pi2 = 0x1.921fb54442d18p+0;  // M_PI/4.0, M_PI/2.0
  const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
  const double rho = sqrt(dx * dx + dy * dy);
  double t1h, t1l, t2h, t2l;
  amc_atan2_fast(ay, ax, &t1h, &t1l);
  amc_atan2_fast(rho, az, &t2h, &t2l);
  double th = t1h, ph = t2h;
  th > pi4 ? th = pi2 - th : th;
  ph < pi4 ? ph = pi2 - ph : ph;
  double ch, cl, sh, sl;
  amc_sincos_fast(th, nullptr, nullptr, &ch, &cl);
  amc_sincos_fast(ph, &sh, &sl, nullptr, nullptr);
  const bool ok = (ay > 0.0) & (ax > 0.0) & (rho > 0.0) & (az > 0.0) & amc_round_ok(t1h, t1l) &
                  amc_round_ok(t2h, t2l) & (ch >= 0.5) & amc_round_ok(ch, cl) & (sh >= 0.5) & amc_round_ok(sh, sl);
  if (ok) return sqrt(ch * sh);
  return amc_shape_factor_ref(dx, dy, dz);
}

// ---------------------------------------------------------------------------------
// pow(x, y) rounded to nearest for x > 0 finite and a result in the normal range: the
// growth velocity Avel * pow(tL - T, nvel) of S/VoxelsCA.C:405 at any sub-liquidus T.
// exp(y * log x) in double-double (relative error about 2^-95), then the nearest double;
// the result is the correctly rounded one unless pow(x, y) lies within that error of a
// rounding midpoint.
//   log x = e ln2 + log c + 2 atanh(s),  x = 2^e m, m in [sqrt(1/2), sqrt 2), c = i/64
//     the table point nearest m, s = (m - c)/(m + c), |s| < 1/180 (terms through s^17);
//   exp z = 2^k exp(j/64) exp(r),  z = k ln2 + j/64 + r, |r| <= 1/128 (terms through r^12).
__host__ __device__ static inline amc_dd amc_log_dd(double x) {
  int e = 0;
  double m = frexp(x, &e);  // m in [1/2, 1)
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  const int i = (int)rint(m * AMC_DD_TAB);  // 45 .. 91
  const double c = (double)i / AMC_DD_TAB;
  const amc_dd num = {m - c, 0.0};           // exact (Sterbenz)
  const amc_dd den = amc_two_sum(m, c);
  const amc_dd s = amc_dd_div(num, den);
  const amc_dd s2 = amc_dd_mul(s, s);
  amc_dd t = amc_dd_mul(s, amc_poly(kAtanhCoef, 9, s2));
  t = amc_dd_add(t, t);                       // 2 atanh(s)
  const amc_dd el = amc_dd_mul_d(amc_tab(kLn2, 0), (double)e);
  return amc_dd_add(amc_dd_add(el, amc_tab(kLogTab, i - 45)), t);
}
__host__ __device__ static inline double amc_exp_dd_round(amc_dd z) {
  const double k = rint(z.hi * 1.4426950408889634);  // z / ln 2
  amc_dd r = amc_dd_add(z, amc_dd_neg(amc_dd_mul_d(amc_tab(kLn2, 0), k)));
  const int j = (int)rint(r.hi * AMC_DD_TAB);         // -23 .. 23
  r = amc_dd_add(r, amc_dd{-(double)j / AMC_DD_TAB, 0.0});
  const amc_dd er = amc_poly(kExpCoef, 13, r);
  const amc_dd v = amc_dd_mul(amc_tab(kExpTab, j + 23), er);
  return ldexp(v.hi + v.lo, (int)k);
}
__host__ __device__ static inline double amc_pow_cr(double x, double y) {
  if (y == 0.0) return 1.0;
  if (y == 1.0) return x;
  return amc_exp_dd_round(amc_dd_mul_d(amc_log_dd(x), y));
}
