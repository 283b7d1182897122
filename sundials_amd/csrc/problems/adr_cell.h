/* -----------------------------------------------------------------------------
 * Per-cell arithmetic of the IMEX advection-diffusion-reaction problem
 * (imex_adr.h), written once for the device kernels and their CPU twin so
 * both round identically (every product and sum is a separate IEEE
 * operation; the files including this one are built without contraction).
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_ADR_CELL_H
#define SUNAMD_ADR_CELL_H

#if defined(__HIPCC__)
#define ADR_HD __host__ __device__ static inline
#else
#define ADR_HD static inline
#endif

/* explicit part for one species: upwind advection + 7-point diffusion;
   ax = cx/dx, ..., dd = d/dx^2; m = minus neighbour, p = plus neighbour */
ADR_HD double adr_fe1(double val, double xm, double xp, double ym, double yp, double zm,
                      double zp, double ax, double ay, double az, double dd)
{
  const double lap = (((xp - 2.0 * val) + xm) + ((yp - 2.0 * val) + ym)) + ((zp - 2.0 * val) + zm);
  const double adv = (ax * (val - xm) + ay * (val - ym)) + az * (val - zm);
  return dd * lap - adv;
}

/* implicit part: the cell's reaction */
ADR_HD void adr_fi(const double* y, double A, double B, double kap, double* f)
{
  const double u = y[0], v = y[1], w = y[2];
  const double uu = u * u;
  f[0] = (A - (w + 1.0) * u) + uu * v;
  f[1] = w * u - uu * v;
  f[2] = kap * (B - w) - w * u;
}

/* the reaction Jacobian, row-major */
ADR_HD void adr_jac(const double* y, double kap, double* J)
{
  const double u = y[0], v = y[1], w = y[2];
  J[0] = (2.0 * u) * v - (w + 1.0);
  J[1] = u * u;
  J[2] = -u;
  J[3] = w - (2.0 * u) * v;
  J[4] = -(u * u);
  J[5] = u;
  J[6] = -w;
  J[7] = 0.0;
  J[8] = -kap - u;
}

ADR_HD void adr_jv(const double* J, const double* v, double* r)
{
  r[0] = (J[0] * v[0] + J[1] * v[1]) + J[2] * v[2];
  r[1] = (J[3] * v[0] + J[4] * v[1]) + J[5] * v[2];
  r[2] = (J[6] * v[0] + J[7] * v[1]) + J[8] * v[2];
}

/* P = I - gamma J factored with partial pivoting (first largest |pivot|):
   LU[9] row-major, multipliers below the diagonal; perm[r] = original row of
   factored row r; returns 0, or k+1 for a zero pivot in column k */
ADR_HD int adr_plu(const double* J, double gamma, double* LU, int* perm)
{
  for (int a = 0; a < 3; a++)
  {
    for (int b = 0; b < 3; b++) { LU[3 * a + b] = ((a == b) ? 1.0 : 0.0) - gamma * J[3 * a + b]; }
    perm[a] = a;
  }
  for (int k = 0; k < 2; k++)
  {
    int pr = k;
    for (int r = k + 1; r < 3; r++)
    {
      if (fabs(LU[3 * r + k]) > fabs(LU[3 * pr + k])) { pr = r; }
    }
    if (pr != k)
    {
      for (int b = 0; b < 3; b++)
      {
        const double t  = LU[3 * k + b];
        LU[3 * k + b]   = LU[3 * pr + b];
        LU[3 * pr + b]  = t;
      }
      const int t = perm[k];
      perm[k]     = perm[pr];
      perm[pr]    = t;
    }
    if (LU[3 * k + k] == 0.0) { return k + 1; }
    for (int r = k + 1; r < 3; r++)
    {
      const double l = LU[3 * r + k] / LU[3 * k + k];
      LU[3 * r + k]  = l;
      for (int b = k + 1; b < 3; b++) { LU[3 * r + b] = LU[3 * r + b] - l * LU[3 * k + b]; }
    }
  }
  return (LU[8] == 0.0) ? 3 : 0;
}

/* z = P^{-1} r from the factors */
ADR_HD void adr_psolve(const double* LU, const int* perm, const double* r, double* z)
{
  double t0 = r[perm[0]], t1 = r[perm[1]], t2 = r[perm[2]];
  t1      = t1 - LU[3] * t0;
  t2      = (t2 - LU[6] * t0) - LU[7] * t1;
  t2      = t2 / LU[8];
  t1      = (t1 - LU[5] * t2) / LU[4];
  t0      = ((t0 - LU[1] * t1) - LU[2] * t2) / LU[0];
  z[0] = t0;
  z[1] = t1;
  z[2] = t2;
}

#endif
