/* -----------------------------------------------------------------------------
 * CPU ORACLE (test infrastructure): serial N_Vector.
 *
 * Restates src/nvector/serial/nvector_serial.c with the same formulas,
 * branch structure and left-to-right summation order:
 *   constructor / ops table, fused ops off by default   :80-166
 *   clone / destroy                                     :300-360
 *   N_VLinearSum and its nine special cases             :387-480, :1600-1775
 *   N_VScale (copy / negate / in-place special cases)   :519-541
 *   reductions (DotProd, MaxNorm, WrmsNorm[Mask], Min, WL2Norm, L1Norm)
 *                                                        :605-761
 *   Compare, InvTest, ConstrMask, MinQuotient            :763-865
 *   fused LinearCombination / ScaleAddMulti / DotProdMulti :871-1027
 *   vector-array ops                                     :1033-1592
 * Built with -ffp-contract=off so every product rounds separately, as the
 * reference's x86-64 build does.
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct
{
  sunindextype length;
  sunbooleantype own_data;
  sunrealtype* data;
} SerialContent;

#define SC(v)   ((SerialContent*)((v)->content))
#define LEN(v)  (SC(v)->length)
#define DATA(v) (SC(v)->data)
#define ONE     1.0
#define ZERO    0.0
#define HALF    0.5
#define ONEPT5  1.5

static N_Vector_ID sid(N_Vector v)
{
  (void)v;
  return SUNDIALS_NVEC_SERIAL;
}
static sunindextype slen(N_Vector v) { return LEN(v); }
static sunrealtype* sptr(N_Vector v) { return DATA(v); }
static void ssetptr(sunrealtype* d, N_Vector v)
{
  if (LEN(v) > 0) { DATA(v) = d; }
}
static void sspace(N_Vector v, sunindextype* lrw, sunindextype* liw)
{
  *lrw = LEN(v);
  *liw = 1;
}

static N_Vector sclone_empty(N_Vector w);
static N_Vector sclone(N_Vector w);
static void sdestroy(N_Vector v);

/* ---- special cases used by LinearSum / Scale -------------------------- */
static void vcopy(N_Vector x, N_Vector z)
{
  sunrealtype *xd = DATA(x), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = xd[i]; }
}
static void vsum(N_Vector x, N_Vector y, N_Vector z)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = xd[i] + yd[i]; }
}
static void vdiff(N_Vector x, N_Vector y, N_Vector z)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = xd[i] - yd[i]; }
}
static void vneg(N_Vector x, N_Vector z)
{
  sunrealtype *xd = DATA(x), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = -xd[i]; }
}
static void vscalesum(sunrealtype c, N_Vector x, N_Vector y, N_Vector z)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = c * (xd[i] + yd[i]); }
}
static void vscalediff(sunrealtype c, N_Vector x, N_Vector y, N_Vector z)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = c * (xd[i] - yd[i]); }
}
static void vlin1(sunrealtype a, N_Vector x, N_Vector y, N_Vector z)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = (a * xd[i]) + yd[i]; }
}
static void vlin2(sunrealtype a, N_Vector x, N_Vector y, N_Vector z)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = (a * xd[i]) - yd[i]; }
}
static void vaxpy(sunrealtype a, N_Vector x, N_Vector y)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y);
  sunindextype n = LEN(x);
  if (a == ONE)
  {
    for (sunindextype i = 0; i < n; i++) { yd[i] += xd[i]; }
    return;
  }
  if (a == -ONE)
  {
    for (sunindextype i = 0; i < n; i++) { yd[i] -= xd[i]; }
    return;
  }
  for (sunindextype i = 0; i < n; i++) { yd[i] += a * xd[i]; }
}
static void vscaleby(sunrealtype a, N_Vector x)
{
  sunrealtype* xd = DATA(x);
  for (sunindextype i = 0; i < LEN(x); i++) { xd[i] *= a; }
}

/* ---- standard ops ------------------------------------------------------- */
static void slinearsum(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z)
{
  sunbooleantype test;
  N_Vector v1, v2;
  sunrealtype c;
  if ((b == ONE) && (z == y))
  {
    vaxpy(a, x, y);
    return;
  }
  if ((a == ONE) && (z == x))
  {
    vaxpy(b, y, x);
    return;
  }
  if ((a == ONE) && (b == ONE))
  {
    vsum(x, y, z);
    return;
  }
  if ((test = ((a == ONE) && (b == -ONE))) || ((a == -ONE) && (b == ONE)))
  {
    v1 = test ? y : x;
    v2 = test ? x : y;
    vdiff(v2, v1, z);
    return;
  }
  if ((test = (a == ONE)) || (b == ONE))
  {
    c  = test ? b : a;
    v1 = test ? y : x;
    v2 = test ? x : y;
    vlin1(c, v1, v2, z);
    return;
  }
  if ((test = (a == -ONE)) || (b == -ONE))
  {
    c  = test ? b : a;
    v1 = test ? y : x;
    v2 = test ? x : y;
    vlin2(c, v1, v2, z);
    return;
  }
  if (a == b)
  {
    vscalesum(a, x, y, z);
    return;
  }
  if (a == -b)
  {
    vscalediff(a, x, y, z);
    return;
  }
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = (a * xd[i]) + (b * yd[i]); }
}

static void sconst(sunrealtype c, N_Vector z)
{
  sunrealtype* zd = DATA(z);
  for (sunindextype i = 0; i < LEN(z); i++) { zd[i] = c; }
}
static void sprod(N_Vector x, N_Vector y, N_Vector z)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = xd[i] * yd[i]; }
}
static void sdiv(N_Vector x, N_Vector y, N_Vector z)
{
  sunrealtype *xd = DATA(x), *yd = DATA(y), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = xd[i] / yd[i]; }
}
static void sscale(sunrealtype c, N_Vector x, N_Vector z)
{
  if (z == x)
  {
    vscaleby(c, x);
    return;
  }
  if (c == ONE) { vcopy(x, z); }
  else if (c == -ONE) { vneg(x, z); }
  else
  {
    sunrealtype *xd = DATA(x), *zd = DATA(z);
    for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = c * xd[i]; }
  }
}
static void sabs(N_Vector x, N_Vector z)
{
  sunrealtype *xd = DATA(x), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = fabs(xd[i]); }
}
static void sinv(N_Vector x, N_Vector z)
{
  sunrealtype *xd = DATA(x), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = ONE / xd[i]; }
}
static void saddconst(N_Vector x, sunrealtype b, N_Vector z)
{
  sunrealtype *xd = DATA(x), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = xd[i] + b; }
}
static sunrealtype sdotprod(N_Vector x, N_Vector y)
{
  sunrealtype sum = ZERO, *xd = DATA(x), *yd = DATA(y);
  for (sunindextype i = 0; i < LEN(x); i++) { sum += xd[i] * yd[i]; }
  return sum;
}
static sunrealtype smaxnorm(N_Vector x)
{
  sunrealtype max = ZERO, *xd = DATA(x);
  for (sunindextype i = 0; i < LEN(x); i++)
  {
    if (fabs(xd[i]) > max) { max = fabs(xd[i]); }
  }
  return max;
}
static sunrealtype swsqrsum(N_Vector x, N_Vector w)
{
  sunrealtype sum = ZERO, prodi, *xd = DATA(x), *wd = DATA(w);
  for (sunindextype i = 0; i < LEN(x); i++)
  {
    prodi = xd[i] * wd[i];
    sum += prodi * prodi;
  }
  return sum;
}
static sunrealtype swsqrsummask(N_Vector x, N_Vector w, N_Vector id)
{
  sunrealtype sum = ZERO, prodi, *xd = DATA(x), *wd = DATA(w), *idd = DATA(id);
  for (sunindextype i = 0; i < LEN(x); i++)
  {
    if (idd[i] > ZERO)
    {
      prodi = xd[i] * wd[i];
      sum += prodi * prodi;
    }
  }
  return sum;
}
static sunrealtype sqrt0(sunrealtype v) { return (v <= ZERO) ? ZERO : sqrt(v); }
static sunrealtype swrmsnorm(N_Vector x, N_Vector w)
{
  return sqrt0(swsqrsum(x, w) / LEN(x));
}
static sunrealtype swrmsnormmask(N_Vector x, N_Vector w, N_Vector id)
{
  return sqrt0(swsqrsummask(x, w, id) / LEN(x));
}
static sunrealtype smin(N_Vector x)
{
  sunrealtype* xd = DATA(x);
  sunrealtype min = xd[0];
  for (sunindextype i = 1; i < LEN(x); i++)
  {
    if (xd[i] < min) { min = xd[i]; }
  }
  return min;
}
static sunrealtype swl2norm(N_Vector x, N_Vector w) { return sqrt0(swsqrsum(x, w)); }
static sunrealtype sl1norm(N_Vector x)
{
  sunrealtype sum = ZERO, *xd = DATA(x);
  for (sunindextype i = 0; i < LEN(x); i++) { sum += fabs(xd[i]); }
  return sum;
}
static void scompare(sunrealtype c, N_Vector x, N_Vector z)
{
  sunrealtype *xd = DATA(x), *zd = DATA(z);
  for (sunindextype i = 0; i < LEN(x); i++) { zd[i] = (fabs(xd[i]) >= c) ? ONE : ZERO; }
}
static sunbooleantype sinvtest(N_Vector x, N_Vector z)
{
  sunrealtype *xd = DATA(x), *zd = DATA(z);
  sunbooleantype no_zero_found = SUNTRUE;
  for (sunindextype i = 0; i < LEN(x); i++)
  {
    if (xd[i] == ZERO) { no_zero_found = SUNFALSE; }
    else { zd[i] = ONE / xd[i]; }
  }
  return no_zero_found;
}
static sunbooleantype sconstrmask(N_Vector c, N_Vector x, N_Vector m)
{
  sunrealtype *cd = DATA(c), *xd = DATA(x), *md = DATA(m);
  sunrealtype temp = ZERO;
  for (sunindextype i = 0; i < LEN(x); i++)
  {
    md[i] = ZERO;
    if (cd[i] == ZERO) { continue; }
    sunbooleantype test = (fabs(cd[i]) > ONEPT5 && xd[i] * cd[i] <= ZERO) ||
                          (fabs(cd[i]) > HALF && xd[i] * cd[i] < ZERO);
    if (test) { temp = md[i] = ONE; }
  }
  return (temp == ONE) ? SUNFALSE : SUNTRUE;
}
static sunrealtype sminquotient(N_Vector num, N_Vector denom)
{
  sunrealtype *nd = DATA(num), *dd = DATA(denom);
  sunbooleantype notEvenOnce = SUNTRUE;
  sunrealtype min            = SUN_BIG_REAL;
  for (sunindextype i = 0; i < LEN(num); i++)
  {
    if (dd[i] == ZERO) { continue; }
    if (!notEvenOnce) { min = fmin(min, nd[i] / dd[i]); }
    else
    {
      min         = nd[i] / dd[i];
      notEvenOnce = SUNFALSE;
    }
  }
  return min;
}

/* ---- fused ops ---------------------------------------------------------- */
static SUNErrCode slincomb(int nvec, sunrealtype* c, N_Vector* X, N_Vector z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  if (nvec == 1)
  {
    sscale(c[0], X[0], z);
    return SUN_SUCCESS;
  }
  if (nvec == 2)
  {
    slinearsum(c[0], X[0], c[1], X[1], z);
    return SUN_SUCCESS;
  }
  sunindextype n = LEN(z);
  sunrealtype* zd = DATA(z);
  if ((X[0] == z) && (c[0] == ONE))
  {
    for (int i = 1; i < nvec; i++)
    {
      sunrealtype* xd = DATA(X[i]);
      for (sunindextype j = 0; j < n; j++) { zd[j] += c[i] * xd[j]; }
    }
    return SUN_SUCCESS;
  }
  if (X[0] == z)
  {
    for (sunindextype j = 0; j < n; j++) { zd[j] *= c[0]; }
    for (int i = 1; i < nvec; i++)
    {
      sunrealtype* xd = DATA(X[i]);
      for (sunindextype j = 0; j < n; j++) { zd[j] += c[i] * xd[j]; }
    }
    return SUN_SUCCESS;
  }
  sunrealtype* xd = DATA(X[0]);
  for (sunindextype j = 0; j < n; j++) { zd[j] = c[0] * xd[j]; }
  for (int i = 1; i < nvec; i++)
  {
    xd = DATA(X[i]);
    for (sunindextype j = 0; j < n; j++) { zd[j] += c[i] * xd[j]; }
  }
  return SUN_SUCCESS;
}

static SUNErrCode sscaleaddmulti(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y,
                                 N_Vector* Z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  if (nvec == 1)
  {
    slinearsum(a[0], x, ONE, Y[0], Z[0]);
    return SUN_SUCCESS;
  }
  sunindextype n  = LEN(x);
  sunrealtype* xd = DATA(x);
  if (Y == Z)
  {
    for (int i = 0; i < nvec; i++)
    {
      sunrealtype* yd = DATA(Y[i]);
      for (sunindextype j = 0; j < n; j++) { yd[j] += a[i] * xd[j]; }
    }
    return SUN_SUCCESS;
  }
  for (int i = 0; i < nvec; i++)
  {
    sunrealtype *yd = DATA(Y[i]), *zd = DATA(Z[i]);
    for (sunindextype j = 0; j < n; j++) { zd[j] = a[i] * xd[j] + yd[j]; }
  }
  return SUN_SUCCESS;
}

static SUNErrCode sdotprodmulti(int nvec, N_Vector x, N_Vector* Y, sunrealtype* d)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  if (nvec == 1)
  {
    d[0] = sdotprod(x, Y[0]);
    return SUN_SUCCESS;
  }
  sunindextype n  = LEN(x);
  sunrealtype* xd = DATA(x);
  for (int i = 0; i < nvec; i++)
  {
    sunrealtype* yd = DATA(Y[i]);
    d[i]            = ZERO;
    for (sunindextype j = 0; j < n; j++) { d[i] += xd[j] * yd[j]; }
  }
  return SUN_SUCCESS;
}

/* vector-array ops: the per-vector loops of the reference give the same
   per-element arithmetic as looping the single-vector op */
static SUNErrCode slinearsumva(int nvec, sunrealtype a, N_Vector* X, sunrealtype b,
                               N_Vector* Y, N_Vector* Z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  for (int i = 0; i < nvec; i++) { slinearsum(a, X[i], b, Y[i], Z[i]); }
  return SUN_SUCCESS;
}
static SUNErrCode sscaleva(int nvec, sunrealtype* c, N_Vector* X, N_Vector* Z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  if (nvec == 1)
  {
    sscale(c[0], X[0], Z[0]);
    return SUN_SUCCESS;
  }
  sunindextype n = LEN(Z[0]);
  for (int i = 0; i < nvec; i++)
  {
    sunrealtype *xd = DATA(X[i]), *zd = DATA(Z[i]);
    if (X == Z)
    {
      for (sunindextype j = 0; j < n; j++) { xd[j] *= c[i]; }
    }
    else
    {
      for (sunindextype j = 0; j < n; j++) { zd[j] = c[i] * xd[j]; }
    }
  }
  return SUN_SUCCESS;
}
static SUNErrCode sconstva(int nvec, sunrealtype c, N_Vector* Z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  for (int i = 0; i < nvec; i++) { sconst(c, Z[i]); }
  return SUN_SUCCESS;
}
static SUNErrCode swrmsva(int nvec, N_Vector* X, N_Vector* W, sunrealtype* nrm)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  if (nvec == 1)
  {
    nrm[0] = swrmsnorm(X[0], W[0]);
    return SUN_SUCCESS;
  }
  sunindextype n = LEN(X[0]);
  for (int i = 0; i < nvec; i++)
  {
    sunrealtype *xd = DATA(X[i]), *wd = DATA(W[i]);
    nrm[i] = ZERO;
    for (sunindextype j = 0; j < n; j++)
    {
      sunrealtype p = xd[j] * wd[j];
      nrm[i] += p * p;
    }
    nrm[i] = sqrt0(nrm[i] / n);
  }
  return SUN_SUCCESS;
}
static SUNErrCode swrmsmaskva(int nvec, N_Vector* X, N_Vector* W, N_Vector id,
                              sunrealtype* nrm)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  if (nvec == 1)
  {
    nrm[0] = swrmsnormmask(X[0], W[0], id);
    return SUN_SUCCESS;
  }
  sunindextype n   = LEN(X[0]);
  sunrealtype* idd = DATA(id);
  for (int i = 0; i < nvec; i++)
  {
    sunrealtype *xd = DATA(X[i]), *wd = DATA(W[i]);
    nrm[i] = ZERO;
    for (sunindextype j = 0; j < n; j++)
    {
      if (idd[j] > ZERO)
      {
        sunrealtype p = xd[j] * wd[j];
        nrm[i] += p * p;
      }
    }
    nrm[i] = sqrt0(nrm[i] / n);
  }
  return SUN_SUCCESS;
}
static SUNErrCode sscaleaddmultiva(int nvec, int nsum, sunrealtype* a, N_Vector* X,
                                   N_Vector** Y, N_Vector** Z)
{
  if (nvec < 1 || nsum < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  sunindextype n = LEN(X[0]);
  for (int i = 0; i < nvec; i++)
  {
    sunrealtype* xd = DATA(X[i]);
    for (int j = 0; j < nsum; j++)
    {
      sunrealtype *yd = DATA(Y[j][i]), *zd = DATA(Z[j][i]);
      if (nvec == 1 && nsum == 1)
      {
        slinearsum(a[0], X[0], ONE, Y[0][0], Z[0][0]);
        return SUN_SUCCESS;
      }
      if (Y == Z)
      {
        for (sunindextype k = 0; k < n; k++) { yd[k] += a[j] * xd[k]; }
      }
      else
      {
        for (sunindextype k = 0; k < n; k++) { zd[k] = a[j] * xd[k] + yd[k]; }
      }
    }
  }
  return SUN_SUCCESS;
}
static SUNErrCode slincombva(int nvec, int nsum, sunrealtype* c, N_Vector** X,
                             N_Vector* Z)
{
  if (nvec < 1 || nsum < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  N_Vector* Yv = (N_Vector*)malloc(nsum * sizeof(N_Vector));
  if (!Yv) { return SUN_ERR_MALLOC_FAIL; }
  for (int i = 0; i < nvec; i++)
  {
    for (int j = 0; j < nsum; j++) { Yv[j] = X[j][i]; }
    slincomb(nsum, c, Yv, Z[i]);
  }
  free(Yv);
  return SUN_SUCCESS;
}

/* ---- buffers, printing --------------------------------------------------- */
static SUNErrCode sbufsize(N_Vector x, sunindextype* size)
{
  *size = LEN(x) * (sunindextype)sizeof(sunrealtype);
  return SUN_SUCCESS;
}
static SUNErrCode sbufpack(N_Vector x, void* buf)
{
  memcpy(buf, DATA(x), (size_t)LEN(x) * sizeof(sunrealtype));
  return SUN_SUCCESS;
}
static SUNErrCode sbufunpack(N_Vector x, void* buf)
{
  memcpy(DATA(x), buf, (size_t)LEN(x) * sizeof(sunrealtype));
  return SUN_SUCCESS;
}
static void sprintfile(N_Vector x, FILE* f)
{
  for (sunindextype i = 0; i < LEN(x); i++) { fprintf(f, "%.16e\n", DATA(x)[i]); }
  fprintf(f, "\n");
}
static void sprint(N_Vector x) { sprintfile(x, stdout); }

/* ---- constructors -------------------------------------------------------- */
N_Vector N_VNewEmpty_Serial(sunindextype length, SUNContext sunctx)
{
  if (length < 0) { return NULL; }
  N_Vector v = N_VNewEmpty(sunctx);
  if (!v) { return NULL; }
  N_Vector_Ops o        = v->ops;
  o->nvgetvectorid      = sid;
  o->nvclone            = sclone;
  o->nvcloneempty       = sclone_empty;
  o->nvdestroy          = sdestroy;
  o->nvspace            = sspace;
  o->nvgetarraypointer  = sptr;
  o->nvsetarraypointer  = ssetptr;
  o->nvgetlength        = slen;
  o->nvgetlocallength   = slen;
  o->nvlinearsum        = slinearsum;
  o->nvconst            = sconst;
  o->nvprod             = sprod;
  o->nvdiv              = sdiv;
  o->nvscale            = sscale;
  o->nvabs              = sabs;
  o->nvinv              = sinv;
  o->nvaddconst         = saddconst;
  o->nvdotprod          = sdotprod;
  o->nvmaxnorm          = smaxnorm;
  o->nvwrmsnormmask     = swrmsnormmask;
  o->nvwrmsnorm         = swrmsnorm;
  o->nvmin              = smin;
  o->nvwl2norm          = swl2norm;
  o->nvl1norm           = sl1norm;
  o->nvcompare          = scompare;
  o->nvinvtest          = sinvtest;
  o->nvconstrmask       = sconstrmask;
  o->nvminquotient      = sminquotient;
  o->nvdotprodlocal     = sdotprod;
  o->nvmaxnormlocal     = smaxnorm;
  o->nvminlocal         = smin;
  o->nvl1normlocal      = sl1norm;
  o->nvinvtestlocal     = sinvtest;
  o->nvconstrmasklocal  = sconstrmask;
  o->nvminquotientlocal = sminquotient;
  o->nvwsqrsumlocal     = swsqrsum;
  o->nvwsqrsummasklocal = swsqrsummask;
  o->nvdotprodmultilocal = sdotprodmulti;
  o->nvbufsize          = sbufsize;
  o->nvbufpack          = sbufpack;
  o->nvbufunpack        = sbufunpack;
  o->nvprint            = sprint;
  o->nvprintfile        = sprintfile;
  SerialContent* c = (SerialContent*)calloc(1, sizeof(SerialContent));
  if (!c)
  {
    N_VFreeEmpty(v);
    return NULL;
  }
  c->length  = length;
  v->content = c;
  return v;
}

N_Vector N_VNew_Serial(sunindextype length, SUNContext sunctx)
{
  N_Vector v = N_VNewEmpty_Serial(length, sunctx);
  if (!v) { return NULL; }
  if (length > 0)
  {
    DATA(v) = (sunrealtype*)malloc((size_t)length * sizeof(sunrealtype));
    if (!DATA(v))
    {
      sdestroy(v);
      return NULL;
    }
    SC(v)->own_data = SUNTRUE;
  }
  return v;
}

N_Vector N_VMake_Serial(sunindextype length, sunrealtype* v_data, SUNContext sunctx)
{
  N_Vector v = N_VNewEmpty_Serial(length, sunctx);
  if (!v) { return NULL; }
  if (length > 0)
  {
    SC(v)->own_data = SUNFALSE;
    DATA(v)         = v_data;
  }
  return v;
}

static N_Vector sclone_empty(N_Vector w)
{
  N_Vector v = N_VNewEmpty(w->sunctx);
  if (!v) { return NULL; }
  N_VCopyOps(w, v);
  SerialContent* c = (SerialContent*)calloc(1, sizeof(SerialContent));
  if (!c)
  {
    N_VFreeEmpty(v);
    return NULL;
  }
  c->length  = LEN(w);
  v->content = c;
  return v;
}

static N_Vector sclone(N_Vector w)
{
  N_Vector v = sclone_empty(w);
  if (!v) { return NULL; }
  if (LEN(w) > 0)
  {
    DATA(v) = (sunrealtype*)malloc((size_t)LEN(w) * sizeof(sunrealtype));
    if (!DATA(v))
    {
      sdestroy(v);
      return NULL;
    }
    SC(v)->own_data = SUNTRUE;
  }
  return v;
}

static void sdestroy(N_Vector v)
{
  if (v == NULL) { return; }
  if (v->content)
  {
    if (SC(v)->own_data && DATA(v)) { free(DATA(v)); }
    free(v->content);
    v->content = NULL;
  }
  N_VFreeEmpty(v);
}

SUNErrCode N_VEnableFusedOps_Serial(N_Vector v, sunbooleantype tf)
{
  if (!v || !v->ops) { return SUN_ERR_ARG_CORRUPT; }
  N_Vector_Ops o                    = v->ops;
  o->nvlinearcombination            = tf ? slincomb : NULL;
  o->nvscaleaddmulti                = tf ? sscaleaddmulti : NULL;
  o->nvdotprodmulti                 = tf ? sdotprodmulti : NULL;
  o->nvlinearsumvectorarray         = tf ? slinearsumva : NULL;
  o->nvscalevectorarray             = tf ? sscaleva : NULL;
  o->nvconstvectorarray             = tf ? sconstva : NULL;
  o->nvwrmsnormvectorarray          = tf ? swrmsva : NULL;
  o->nvwrmsnormmaskvectorarray      = tf ? swrmsmaskva : NULL;
  o->nvscaleaddmultivectorarray     = tf ? sscaleaddmultiva : NULL;
  o->nvlinearcombinationvectorarray = tf ? slincombva : NULL;
  o->nvdotprodmulti                 = tf ? sdotprodmulti : NULL;
  return SUN_SUCCESS;
}
