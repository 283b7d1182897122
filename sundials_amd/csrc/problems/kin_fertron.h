/* -----------------------------------------------------------------------------
 * kin_fertron.h -- cell arithmetic of the Newton test problem of
 * include/sunamd/kin_newton.h, shared by the MI355X kernels and the CPU twin.
 *
 * One cell is the Ferraris-Tronconi system of examples/kinsol/serial/
 * kinFerTron_dns.c:300-345, with its bounds written as extra unknowns so that
 * KINSOL's inequality constraints enforce them: u = (x1, x2, l1, L1, l2, L2),
 *   f1 = 1/2 sin(x1 x2) - x2 / (4 pi) - x1 / 2
 *   f2 = (1 - 1/(4 pi)) (exp(2 x1) - e) + e x2 / pi - 2 e x1
 *   f3 = l1 - x1 + lb1,  f4 = L1 - x1 + ub1,  f5 = l2 - x2 + lb2,
 *   f6 = L2 - x2 + ub2,  with l >= 0 and L <= 0,
 * lb = (0.25, 1.5), ub = (1, 2 pi), and the example's pi = 3.1415926 and
 * e = 2.7182818.  Cells are independent: 6 unknowns per cell, cell by cell.
 *
 * Variant 0 uses the math library's sin and exp (the reference example);
 * variant 1 uses kft_sincos / kft_exp below, fixed sequences of IEEE +, -,
 * *, / and rint / ldexp, so a device evaluation equals the CPU one bit for
 * bit.  Their accuracy is a few ulp for arguments of modest size, which is
 * all the test problem needs.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_KIN_FERTRON_H
#define SUNAMD_KIN_FERTRON_H

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define KFT_HD __host__ __device__ inline
#else
#define KFT_HD static inline
#endif

#define KFT_PI 3.1415926 /* the example's constants */
#define KFT_E  2.7182818
#define KFT_NEQ 6

/* exp(x): x = n ln2 + r with ln2 split so n * KFT_LN2_HI is exact,
   |r| <= ln2 / 2, exp(r) by its Taylor series through r^16 / 16! (nested,
   truncation below 1e-19), then scaled by 2^n */
KFT_HD double kft_exp(double x)
{
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  const double n      = rint(x * 1.4426950408889634);
  const double r      = (x - n * LN2_HI) - n * LN2_LO;
  double s            = 1.0;
  for (int k = 16; k >= 1; k--) { s = 1.0 + (s * r) / (double)k; }
  return ldexp(s, (int)n);
}

/* sin t and cos t: t = q pi/2 + r with pi/2 split in two parts, |r| <= pi/4,
   Taylor series of sin r and cos r through r^19 and r^18, then the quadrant
   q mod 4 selects and signs them */
KFT_HD void kft_sincos(double t, double* sn, double* cs)
{
  const double PIO2_HI = 1.57079632673412561417e+00;
  const double PIO2_LO = 6.07710050650619224932e-11;
  const double q       = rint(t * 0.63661977236758134308);
  const double r       = (t - q * PIO2_HI) - q * PIO2_LO;
  const double r2      = r * r;
  double S = 1.0, C = 1.0;
  for (int k = 9; k >= 1; k--)
  {
    S = 1.0 - (S * r2) / (double)((2 * k) * (2 * k + 1));
    C = 1.0 - (C * r2) / (double)((2 * k - 1) * (2 * k));
  }
  const double s0 = r * S, c0 = C;
  switch (((long)q) & 3)
  {
  case 0: *sn = s0; *cs = c0; break;
  case 1: *sn = c0; *cs = -s0; break;
  case 2: *sn = -s0; *cs = -c0; break;
  default: *sn = -c0; *cs = s0; break;
  }
}

KFT_HD void kft_bounds(double* lb, double* ub)
{
  lb[0] = 0.25;
  ub[0] = 1.0;
  lb[1] = 1.5;
  ub[1] = 2.0 * KFT_PI;
}

KFT_HD void kft_f(int variant, const double* u, double* f)
{
  double lb[2], ub[2];
  kft_bounds(lb, ub);
  const double x1 = u[0], x2 = u[1];
  double sn, cs, ex;
  if (variant == 0)
  {
    sn = sin(x1 * x2);
    ex = exp(2.0 * x1);
  }
  else
  {
    kft_sincos(x1 * x2, &sn, &cs);
    ex = kft_exp(2.0 * x1);
  }
  f[0] = 0.5 * sn - 0.25 * x2 / KFT_PI - 0.5 * x1;
  f[1] = (1.0 - 0.25 / KFT_PI) * (ex - KFT_E) + KFT_E * x2 / KFT_PI - 2.0 * KFT_E * x1;
  f[2] = u[2] - x1 + lb[0];
  f[3] = u[3] - x1 + ub[0];
  f[4] = u[4] - x2 + lb[1];
  f[5] = u[5] - x2 + ub[1];
}

/* analytic Jacobian of one cell, column-major 6 x 6: J[j * 6 + i] = df_i/du_j */
KFT_HD void kft_jac(int variant, const double* u, double* J)
{
  const double x1 = u[0], x2 = u[1];
  double sn, cs, ex;
  if (variant == 0)
  {
    cs = cos(x1 * x2);
    ex = exp(2.0 * x1);
  }
  else
  {
    kft_sincos(x1 * x2, &sn, &cs);
    ex = kft_exp(2.0 * x1);
  }
  for (int i = 0; i < KFT_NEQ * KFT_NEQ; i++) { J[i] = 0.0; }
  J[0 * 6 + 0] = 0.5 * x2 * cs - 0.5;
  J[1 * 6 + 0] = 0.5 * x1 * cs - 0.25 / KFT_PI;
  J[0 * 6 + 1] = (1.0 - 0.25 / KFT_PI) * 2.0 * ex - 2.0 * KFT_E;
  J[1 * 6 + 1] = KFT_E / KFT_PI;
  J[0 * 6 + 2] = -1.0;
  J[2 * 6 + 2] = 1.0;
  J[0 * 6 + 3] = -1.0;
  J[3 * 6 + 3] = 1.0;
  J[1 * 6 + 4] = -1.0;
  J[4 * 6 + 4] = 1.0;
  J[1 * 6 + 5] = -1.0;
  J[5 * 6 + 5] = 1.0;
}

/* constraints of one cell: x free, l >= 0, L <= 0 */
KFT_HD void kft_constraints(double* c)
{
  c[0] = 0.0;
  c[1] = 0.0;
  c[2] = 1.0;
  c[3] = -1.0;
  c[4] = 1.0;
  c[5] = -1.0;
}

/* initial guess of cell c: guess 1 at the lower bounds, guess 2 in the middle
   of the box (the example's two guesses); with perturb, x moves up by a
   fraction in [0, 0.02) of the box width, from an integer hash of (c, s) */
KFT_HD void kft_u0(long c, int guess, int perturb, double* u)
{
  double lb[2], ub[2], x[2];
  kft_bounds(lb, ub);
  for (int s = 0; s < 2; s++)
  {
    if (!perturb) { x[s] = (guess == 1) ? lb[s] : 0.5 * (lb[s] + ub[s]); }
    else
    {
      uint64_t h = (uint64_t)(2 * c + s) * 0x9E3779B97F4A7C15ull;
      h ^= h >> 29;
      h *= 0xBF58476D1CE4E5B9ull;
      h ^= h >> 32;
      const double d = (double)(h >> 11) * (1.0 / 9007199254740992.0) * 0.02;
      x[s]           = lb[s] + (((guess == 1) ? 0.0 : 0.5) + d) * (ub[s] - lb[s]);
    }
  }
  u[0] = x[0];
  u[1] = x[1];
  u[2] = x[0] - lb[0];
  u[3] = x[0] - ub[0];
  u[4] = x[1] - lb[1];
  u[5] = x[1] - ub[1];
}

#endif
