/* -----------------------------------------------------------------------------
 * Correctly rounded x^y for x > 0, shared by the host integrator and the
 * device integrator so both take identical step-size decisions.
 *
 * The integrator's step-size ratios are eta = 1 / ((k*dsm)^(1/L) + 1e-6)
 * (reference: SUNRpowerR = pow, cvode.c:3394, 3569, 3620, 3657).  Host libm
 * and the device math library may round pow differently in the last place,
 * and the step controller of stiff problems turns such one-ulp differences
 * into different step sequences.  This routine evaluates
 * exp(y * log(x)) in double-double arithmetic (about 100 significant bits)
 * and rounds once, so it returns the correctly rounded result except when
 * the exact value lies within ~2^-100 relative of a rounding boundary.
 *
 * Method: x = 2^e * f with f in [sqrt(1/2), sqrt(2)); log f = 2 atanh(s),
 * s = (f-1)/(f+1), by the odd series in s^2 (|s| < 0.1716, 26 terms to 2^-106);
 * log x = e*ln2 + log f with ln2 as a double-double constant.  exp: z = y*log x,
 * k = round(z/ln2), r = z - k*ln2 (|r| <= ln2/2), exp(r) by Taylor series in
 * double-double (27 terms), then scaled by 2^k.  Products use fma-based
 * exact two-product; all steps are IEEE operations with explicit fma, which
 * round identically on x86-64 and gfx950.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_POW_H
#define SUNAMD_POW_H

#include <math.h>

#if defined(__HIPCC__)
#define SUNAMD_HD __host__ __device__ static inline
#else
#define SUNAMD_HD static inline
#endif

typedef struct
{
  double hi, lo;
} sunamd_dd;

SUNAMD_HD sunamd_dd dd_two_sum(double a, double b)
{
  double s  = a + b;
  double bb = s - a;
  double e  = (a - (s - bb)) + (b - bb);
  sunamd_dd r = {s, e};
  return r;
}

SUNAMD_HD sunamd_dd dd_fast_two_sum(double a, double b)
{
  double s = a + b;
  double e = b - (s - a);
  sunamd_dd r = {s, e};
  return r;
}

SUNAMD_HD sunamd_dd dd_add(sunamd_dd a, sunamd_dd b)
{
  sunamd_dd s = dd_two_sum(a.hi, b.hi);
  sunamd_dd t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}

SUNAMD_HD sunamd_dd dd_mul(sunamd_dd a, sunamd_dd b)
{
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e        = fma(a.hi, b.lo, e);
  e        = fma(a.lo, b.hi, e);
  return dd_fast_two_sum(p, e);
}

SUNAMD_HD sunamd_dd dd_mul_d(sunamd_dd a, double b)
{
  double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e        = fma(a.lo, b, e);
  return dd_fast_two_sum(p, e);
}

SUNAMD_HD sunamd_dd dd_div(sunamd_dd a, sunamd_dd b)
{
  double q1   = a.hi / b.hi;
  sunamd_dd r = dd_add(a, dd_mul_d(b, -q1));
  double q2   = r.hi / b.hi;
  r           = dd_add(r, dd_mul_d(b, -q2));
  double q3   = r.hi / b.hi;
  sunamd_dd q = dd_fast_two_sum(q1, q2);
  sunamd_dd t = {q3, 0.0};
  return dd_add(q, t);
}

/* 1/n as a double-double: hi = RN(1/n), lo = RN((1 - hi*n)/n), the residual
   1 - hi*n being exact through fma */
SUNAMD_HD sunamd_dd dd_recip_int(int n)
{
  double d    = (double)n;
  double hi   = 1.0 / d;
  double res  = fma(-hi, d, 1.0);
  sunamd_dd r = {hi, res / d};
  return r;
}

/* natural log of x > 0 (finite, normal or subnormal) in double-double */
SUNAMD_HD sunamd_dd dd_log(double x)
{
  int e    = 0;
  double f = frexp(x, &e); /* x = f * 2^e, f in [0.5, 1) */
  if (f < 0.70710678118654752440)
  {
    f = f * 2.0;
    e = e - 1;
  }
  /* s = (f - 1) / (f + 1) in double-double (f - 1 and f + 1 are exact) */
  sunamd_dd num = {f - 1.0, 0.0};
  sunamd_dd den = dd_two_sum(f, 1.0);
  sunamd_dd s   = dd_div(num, den);
  sunamd_dd s2  = dd_mul(s, s);
  /* 2 * sum_{k>=0} s^(2k+1) / (2k+1) */
  sunamd_dd acc = dd_recip_int(53);
  for (int k = 25; k >= 0; k--) { acc = dd_add(dd_mul(acc, s2), dd_recip_int(2 * k + 1)); }
  sunamd_dd lf = dd_mul(dd_mul_d(acc, 2.0), s);
  const sunamd_dd ln2 = {6.93147180559945286227e-01, 2.31904681384629955842e-17};
  return dd_add(dd_mul_d(ln2, (double)e), lf);
}

/* exp(z) for a double-double argument, rounded to double */
SUNAMD_HD double dd_exp_round(sunamd_dd z)
{
  if (z.hi > 709.8) { return INFINITY; }
  if (z.hi < -745.2) { return 0.0; }
  const sunamd_dd ln2 = {6.93147180559945286227e-01, 2.31904681384629955842e-17};
  double k        = nearbyint(z.hi / ln2.hi);
  sunamd_dd r     = dd_add(z, dd_mul_d(ln2, -k));
  /* Taylor series sum r^n / n!, Horner from the top */
  const sunamd_dd one = {1.0, 0.0};
  sunamd_dd acc       = one;
  for (int n = 27; n >= 1; n--) { acc = dd_add(dd_mul(dd_mul(acc, r), dd_recip_int(n)), one); }
  /* acc.hi is RN(acc.hi + acc.lo); scaling by 2^k is exact for normal results */
  return ldexp(acc.hi, (int)k);
}

/* correctly rounded pow for x > 0; other arguments fall back to pow() */
SUNAMD_HD double sunamd_pow(double x, double y)
{
  if (!(x > 0.0) || !isfinite(x) || !isfinite(y)) { return pow(x, y); }
  if (y == 0.0 || x == 1.0) { return 1.0; }
  if (y == 1.0) { return x; }
  sunamd_dd l = dd_log(x);
  sunamd_dd z = dd_mul_d(l, y);
  return dd_exp_round(z);
}

#endif
