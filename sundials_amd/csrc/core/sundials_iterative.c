/* -----------------------------------------------------------------------------
 * Gram-Schmidt orthogonalisation and Givens QR of the Arnoldi Hessenberg
 * matrix, as the Krylov solvers use them.  Restatement of
 *   SUNModifiedGS   src/sundials/sundials_iterative.c:45-105
 *   SUNClassicalGS  src/sundials/sundials_iterative.c:116-170
 *   SUNQRfact       src/sundials/sundials_iterative.c:181-286
 *   SUNQRsol        src/sundials/sundials_iterative.c:297-329
 *   SUNQRAdd_MGS / _ICWY / _CGS2 / _DCGS2
 *                   src/sundials/sundials_iterative.c:333-606
 * The vector work goes through the N_Vector ops (DotProd / LinearSum for
 * modified GS, DotProdMulti / LinearCombination for classical GS), so on a
 * device vector the orthogonalisation runs on the device.  The
 * reorthogonalisation tests (the factor 1000) and the index arithmetic of the
 * classical variant's second pass are kept as the reference has them.
 * ---------------------------------------------------------------------------*/
#include <math.h>

#include "sunamd/sundials_core.h"
#include "sunamd/sundials_iterative.h"

#define GS_FACTOR 1000.0

static double nonneg_sqrt(double x) { return (x <= 0.0) ? 0.0 : sqrt(x); }

SUNErrCode SUNModifiedGS(N_Vector* v, sunrealtype** h, int k, int p, sunrealtype* new_vk_norm)
{
  const int col = k - 1;
  const int i0  = (k - p > 0) ? k - p : 0;
  const double vk_norm = nonneg_sqrt(N_VDotProd(v[k], v[k]));

  for (int i = i0; i < k; i++)
  {
    h[i][col] = N_VDotProd(v[i], v[k]);
    N_VLinearSum(1.0, v[k], -h[i][col], v[i], v[k]);
  }
  *new_vk_norm = nonneg_sqrt(N_VDotProd(v[k], v[k]));

  /* a new norm that is negligible against 1000 * the old one means the
     projection removed almost everything: orthogonalise a second time */
  double big = GS_FACTOR * vk_norm;
  if ((big + (*new_vk_norm)) != big) { return SUN_SUCCESS; }

  double removed2 = 0.0;
  for (int i = i0; i < k; i++)
  {
    double d = N_VDotProd(v[i], v[k]);
    big      = GS_FACTOR * h[i][col];
    if ((big + d) == big) { continue; }
    h[i][col] += d;
    N_VLinearSum(1.0, v[k], -d, v[i], v[k]);
    removed2 += d * d;
  }
  if (removed2 != 0.0)
  {
    double r = (*new_vk_norm) * (*new_vk_norm) - removed2;
    *new_vk_norm = (r > 0.0) ? sqrt(r) : 0.0;
  }
  return SUN_SUCCESS;
}

SUNErrCode SUNClassicalGS(N_Vector* v, sunrealtype** h, int k, int p, sunrealtype* new_vk_norm,
                          sunrealtype* stemp, N_Vector* vtemp)
{
  const int col = k - 1;
  const int i0  = (k - p > 0) ? k - p : 0;
  const int nv  = k - i0; /* vectors projected out */
  SUNErrCode err;

  /* all projections at once; the last entry is v[k].v[k] */
  err = N_VDotProdMulti(nv + 1, v[k], v + i0, stemp);
  if (err != SUN_SUCCESS) { return err; }
  const double vk_norm = nonneg_sqrt(stemp[nv]);

  /* v[k] = 1*v[k] - sum_i h_i v[i0+i] as one linear combination */
  for (int i = nv - 1; i >= 0; i--)
  {
    h[i][col]    = stemp[i];
    stemp[i + 1] = -stemp[i];
    vtemp[i + 1] = v[i];
  }
  stemp[0] = 1.0;
  vtemp[0] = v[k];
  err      = N_VLinearCombination(nv + 1, stemp, vtemp, v[k]);
  if (err != SUN_SUCCESS) { return err; }
  *new_vk_norm = nonneg_sqrt(N_VDotProd(v[k], v[k]));

  if ((GS_FACTOR * (*new_vk_norm)) < vk_norm)
  {
    /* second pass; the coefficient and vector indices are those of the
       reference (identical to the first pass's when i0 == 0) */
    err = N_VDotProdMulti(nv, v[k], v + i0, stemp + 1);
    if (err != SUN_SUCCESS) { return err; }
    stemp[0] = 1.0;
    vtemp[0] = v[k];
    for (int i = i0; i < k; i++)
    {
      h[i][col] += stemp[i - i0 + 1];
      stemp[i - i0 + 1] = -stemp[i - i0 + 1];
      vtemp[i - i0 + 1] = v[i - i0];
    }
    err = N_VLinearCombination(k + 1, stemp, vtemp, v[k]);
    if (err != SUN_SUCCESS) { return err; }
    *new_vk_norm = nonneg_sqrt(N_VDotProd(v[k], v[k]));
  }
  return SUN_SUCCESS;
}

/* Givens rotation (c, s) that zeroes b in (a, b) */
static void givens(double a, double b, double* c, double* s)
{
  if (b == 0.0)
  {
    *c = 1.0;
    *s = 0.0;
  }
  else if (fabs(b) >= fabs(a))
  {
    double t = a / b;
    *s       = -1.0 / nonneg_sqrt(1.0 + t * t);
    *c       = -(*s) * t;
  }
  else
  {
    double t = b / a;
    *c       = 1.0 / nonneg_sqrt(1.0 + t * t);
    *s       = -(*c) * t;
  }
}

/* apply the stored rotations 0..nrot-1 to column j of h */
static void rotate_column(sunrealtype** h, const sunrealtype* q, int nrot, int j)
{
  for (int r = 0; r < nrot; r++)
  {
    double c = q[2 * r], s = q[2 * r + 1];
    double a = h[r][j], b = h[r + 1][j];
    h[r][j]     = c * a - s * b;
    h[r + 1][j] = s * a + c * b;
  }
}

int SUNQRfact(int n, sunrealtype** h, sunrealtype* q, int job)
{
  int code = 0;
  if (job == 0)
  {
    /* factor all n columns from scratch; column k sees the first k-1
       rotations before its own is formed */
    for (int k = 0; k < n; k++)
    {
      rotate_column(h, q, k - 1 > 0 ? k - 1 : 0, k);
      double a = h[k][k], b = h[k + 1][k], c, s;
      givens(a, b, &c, &s);
      q[2 * k]     = c;
      q[2 * k + 1] = s;
      if ((h[k][k] = c * a - s * b) == 0.0) { code = k + 1; }
    }
    return code;
  }
  /* one new column n-1: previous n-1 rotations, then a new one */
  const int j = n - 1;
  rotate_column(h, q, j, j);
  double a = h[j][j], b = h[n][j], c, s;
  givens(a, b, &c, &s);
  q[2 * j]     = c;
  q[2 * j + 1] = s;
  if ((h[j][j] = c * a - s * b) == 0.0) { code = n; }
  return code;
}

int SUNQRsol(int n, sunrealtype** h, sunrealtype* q, sunrealtype* b)
{
  for (int k = 0; k < n; k++)
  {
    double c = q[2 * k], s = q[2 * k + 1];
    double t1 = b[k], t2 = b[k + 1];
    b[k]     = c * t1 - s * t2;
    b[k + 1] = s * t1 + c * t2;
  }
  for (int k = n - 1; k >= 0; k--)
  {
    if (h[k][k] == 0.0) { return k + 1; }
    b[k] /= h[k][k];
    for (int i = 0; i < k; i++) { b[i] -= b[k] * h[i][k]; }
  }
  return 0;
}

/* ---- QR updates of an Anderson-acceleration space ---------------------------
 * Each appends df as column m of Q R (R column-major, leading dimension mMax):
 * the projections of df on Q_0..Q_{m-1} go to R(0:m-1, m), the remainder's
 * norm to R(m, m) and the normalised remainder to Q[m].  Vector work through
 * the N_Vector ops in the reference's order. */
#define RC(i, j) R[(j) * mMax + (i)]

SUNErrCode SUNQRAdd_MGS(N_Vector* Q, sunrealtype* R, N_Vector df, int m, int mMax, void* QRdata)
{
  SUNQRData d = (SUNQRData)QRdata;
  N_Vector r  = d->vtemp;
  N_VScale(1.0, df, r);
  for (int j = 0; j < m; j++)
  {
    RC(j, m) = N_VDotProd(Q[j], r);
    N_VLinearSum(1.0, r, -RC(j, m), Q[j], r);
  }
  RC(m, m) = nonneg_sqrt(N_VDotProd(r, r));
  N_VScale(1.0 / RC(m, m), r, Q[m]);
  return SUN_SUCCESS;
}

/* inverse compact WY: T (lower triangle, row k-1 = Q^T Q_{k-1}) corrects the
   classical projections for the loss of orthogonality of Q */
SUNErrCode SUNQRAdd_ICWY(N_Vector* Q, sunrealtype* R, N_Vector df, int m, int mMax,
                         void* QRdata)
{
  SUNQRData d    = (SUNQRData)QRdata;
  sunrealtype* T = d->temp_array;
  N_VScale(1.0, df, d->vtemp);
  if (m > 0)
  {
    SUNErrCode e = N_VDotProdMulti(m, Q[m - 1], Q, T + (m - 1) * mMax);
    if (e) { return e; }
    T[(m - 1) * mMax + (m - 1)] = 1.0;
    e = N_VDotProdMulti(m, d->vtemp, Q, &RC(0, m));
    if (e) { return e; }
    /* forward substitution with T^T (unit diagonal) */
    for (int k = 0; k < m; k++)
    {
      for (int j = k + 1; j < m; j++) { RC(j, m) -= RC(k, m) * T[j * mMax + k]; }
    }
    e = N_VLinearCombination(m, &RC(0, m), Q, d->vtemp2);
    if (e) { return e; }
    N_VLinearSum(1.0, d->vtemp, -1.0, d->vtemp2, d->vtemp);
  }
  RC(m, m) = nonneg_sqrt(N_VDotProd(d->vtemp, d->vtemp));
  N_VScale(1.0 / RC(m, m), d->vtemp, Q[m]);
  return SUN_SUCCESS;
}

/* classical Gram-Schmidt with one full reorthogonalisation pass */
SUNErrCode SUNQRAdd_CGS2(N_Vector* Q, sunrealtype* R, N_Vector df, int m, int mMax,
                         void* QRdata)
{
  SUNQRData d    = (SUNQRData)QRdata;
  sunrealtype* z = d->temp_array;
  N_VScale(1.0, df, d->vtemp);
  if (m > 0)
  {
    SUNErrCode e = N_VDotProdMulti(m, d->vtemp, Q, &RC(0, m));
    if (e) { return e; }
    e = N_VLinearCombination(m, &RC(0, m), Q, d->vtemp2);
    if (e) { return e; }
    N_VLinearSum(1.0, d->vtemp, -1.0, d->vtemp2, d->vtemp2);
    e = N_VDotProdMulti(m, d->vtemp2, Q, z);
    if (e) { return e; }
    /* Q[m] is a work vector until it receives the new column */
    e = N_VLinearCombination(m, z, Q, Q[m]);
    if (e) { return e; }
    N_VLinearSum(1.0, d->vtemp2, -1.0, Q[m], d->vtemp);
    for (int j = 0; j < m; j++) { RC(j, m) = RC(j, m) + z[j]; }
  }
  RC(m, m) = nonneg_sqrt(N_VDotProd(d->vtemp, d->vtemp));
  N_VScale(1.0 / RC(m, m), d->vtemp, Q[m]);
  return SUN_SUCCESS;
}

/* classical Gram-Schmidt with the reorthogonalisation of the previous column
   delayed to this call */
SUNErrCode SUNQRAdd_DCGS2(N_Vector* Q, sunrealtype* R, N_Vector df, int m, int mMax,
                          void* QRdata)
{
  SUNQRData d    = (SUNQRData)QRdata;
  sunrealtype* s = d->temp_array;
  N_VScale(1.0, df, d->vtemp);
  if (m > 0)
  {
    SUNErrCode e = N_VDotProdMulti(m, d->vtemp, Q, &RC(0, m));
    if (e) { return e; }
    if (m > 1)
    {
      e = N_VDotProdMulti(m - 1, Q[m - 1], Q, s);
      if (e) { return e; }
      e = N_VLinearCombination(m - 1, s, Q, d->vtemp2);
      if (e) { return e; }
      N_VLinearSum(1.0, Q[m - 1], -1.0, d->vtemp2, Q[m - 1]);
      for (int j = 0; j < m - 1; j++) { RC(j, m - 1) = RC(j, m - 1) + s[j]; }
    }
    e = N_VLinearCombination(m, &RC(0, m), Q, d->vtemp2);
    if (e) { return e; }
    N_VLinearSum(1.0, d->vtemp, -1.0, d->vtemp2, d->vtemp);
  }
  RC(m, m) = nonneg_sqrt(N_VDotProd(d->vtemp, d->vtemp));
  N_VScale(1.0 / RC(m, m), d->vtemp, Q[m]);
  return SUN_SUCCESS;
}
#undef RC
