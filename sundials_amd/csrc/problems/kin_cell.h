/* -----------------------------------------------------------------------------
 * kin_cell.h -- cell arithmetic of the fixed-point test problem of
 * include/sunamd/kin_fp.h, shared by the MI355X kernels and the CPU twin.
 *
 * Variant 0 is the reference example's map (examples/kinsol/serial/
 * kinAnalytic_fp.c:300-330):
 *   g1 = 1/3 cos((y-1) z) + 1/6
 *   g2 = 1/9 sqrt(x^2 + sin z + 1.06) + 0.9
 *   g3 = -1/20 exp(-x (y-1)) - (10 pi - 3)/60
 * Variant 1 replaces cos, sin and exp by their first Taylor terms, so every
 * operation is an IEEE-rounded +, -, *, / or sqrt and a device evaluation
 * equals the CPU one bit for bit:
 *   g1 = 1/3 (1 - ((y-1) z)^2 / 2) + 1/6
 *   g2 = 1/9 sqrt(x^2 + z + 1.06) + 0.9
 *   g3 = -1/20 (1 - x (y-1)) - (10 pi - 3)/60
 * Cells are independent: u = (x, y, z) per cell, stored cell by cell.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_KIN_CELL_H
#define SUNAMD_KIN_CELL_H

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define KIN_HD __host__ __device__ inline
#else
#define KIN_HD static inline
#endif

#define KIN_PI 3.1415926535898 /* the example's value */

KIN_HD void kin_cell_g(int variant, const double* u, double* g)
{
  const double x = u[0], y = u[1], z = u[2];
  const double one = 1.0;
  if (variant == 0)
  {
    g[0] = (one / 3.0) * cos((y - one) * z) + (one / 6.0);
    g[1] = (one / 9.0) * sqrt(x * x + sin(z) + 1.06) + 0.9;
    g[2] = -(one / 20.0) * exp(-x * (y - one)) - (10.0 * KIN_PI - 3.0) / 60.0;
  }
  else
  {
    const double w = (y - one) * z;
    g[0] = (one / 3.0) * (one - (w * w) / 2.0) + (one / 6.0);
    g[1] = (one / 9.0) * sqrt(x * x + z + 1.06) + 0.9;
    g[2] = -(one / 20.0) * (one - x * (y - one)) - (10.0 * KIN_PI - 3.0) / 60.0;
  }
}

/* initial guess of cell c: the example's (0.1, 0.1, -0.1), moved by up to
   +-0.01 per component from an integer hash of (c, component) when perturb */
KIN_HD void kin_cell_u0(long c, int perturb, double* u)
{
  const double base[3] = {0.1, 0.1, -0.1};
  for (int s = 0; s < 3; s++)
  {
    double d = 0.0;
    if (perturb)
    {
      uint64_t h = (uint64_t)(3 * c + s) * 0x9E3779B97F4A7C15ull;
      h ^= h >> 29;
      h *= 0xBF58476D1CE4E5B9ull;
      h ^= h >> 32;
      d = ((double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5) * 0.02;
    }
    u[s] = base[s] + d;
  }
}

#endif
