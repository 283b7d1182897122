/* -----------------------------------------------------------------------------
 * Generic-object layer: context, dispatch wrappers and empty constructors.
 *
 * Semantics follow the reference generic layer:
 *   src/sundials/sundials_nvector.c:46-278   (N_VNewEmpty, N_VFreeEmpty, N_VCopyOps)
 *   src/sundials/sundials_nvector.c:286-545  (standard op dispatch)
 *   src/sundials/sundials_nvector.c:549-831  (fused / array ops and their loop
 *                                             fallbacks when a slot is NULL)
 *   src/sundials/sundials_nvector.c:1042-1090 (vector arrays)
 *   src/sundials/sundials_matrix.c, sundials_linearsolver.c,
 *   sundials_nonlinearsolver.c               (thin wrappers)
 *   src/sundials/sundials_math.c:37-80        (SUNRpowerI, SUNRCompare[Tol])
 * The fallbacks matter for results: an implementation that leaves a fused slot
 * NULL gets the loop below, and a fused implementation must reproduce the
 * loop's rounding to give identical integrator decisions.
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sunamd/sundials_core.h"

struct SUNContext_
{
  SUNErrCode last_err;
  SUNComm comm;
};

SUNErrCode SUNContext_Create(SUNComm comm, SUNContext* sunctx_out)
{
  if (!sunctx_out) { return SUN_ERR_ARG_CORRUPT; }
  SUNContext ctx = (SUNContext)calloc(1, sizeof(struct SUNContext_));
  if (!ctx) { return SUN_ERR_MALLOC_FAIL; }
  ctx->comm     = comm;
  ctx->last_err = SUN_SUCCESS;
  *sunctx_out   = ctx;
  return SUN_SUCCESS;
}

SUNErrCode SUNContext_Free(SUNContext* ctx)
{
  if (!ctx || !*ctx) { return SUN_SUCCESS; }
  free(*ctx);
  *ctx = NULL;
  return SUN_SUCCESS;
}

SUNErrCode SUNContext_GetLastError(SUNContext sunctx)
{
  if (!sunctx) { return SUN_ERR_SUNCTX_CORRUPT; }
  SUNErrCode e     = sunctx->last_err;
  sunctx->last_err = SUN_SUCCESS;
  return e;
}

SUNErrCode SUNContext_PeekLastError(SUNContext sunctx)
{
  if (!sunctx) { return SUN_ERR_SUNCTX_CORRUPT; }
  return sunctx->last_err;
}

void SUNAMD_SetLastError(SUNContext sunctx, SUNErrCode err)
{
  if (sunctx) { sunctx->last_err = err; }
}

/* ------------------------------------------------------------------------- */
/* N_Vector                                                                  */
/* ------------------------------------------------------------------------- */

N_Vector N_VNewEmpty(SUNContext sunctx)
{
  if (!sunctx) { return NULL; }
  N_Vector v = (N_Vector)malloc(sizeof(*v));
  if (!v) { return NULL; }
  N_Vector_Ops ops = (N_Vector_Ops)calloc(1, sizeof(*ops));
  if (!ops)
  {
    free(v);
    return NULL;
  }
  v->ops     = ops;
  v->content = NULL;
  v->sunctx  = sunctx;
  return v;
}

void N_VFreeEmpty(N_Vector v)
{
  if (v == NULL) { return; }
  if (v->ops) { free(v->ops); }
  v->ops = NULL;
  free(v);
}

SUNErrCode N_VCopyOps(N_Vector w, N_Vector v)
{
  if (!w || !v || !w->ops || !v->ops) { return SUN_ERR_ARG_CORRUPT; }
  memcpy(v->ops, w->ops, sizeof(*v->ops));
  return SUN_SUCCESS;
}

N_Vector_ID N_VGetVectorID(N_Vector w) { return w->ops->nvgetvectorid(w); }
N_Vector N_VClone(N_Vector w) { return w->ops->nvclone(w); }
N_Vector N_VCloneEmpty(N_Vector w) { return w->ops->nvcloneempty(w); }

void N_VDestroy(N_Vector v)
{
  if (v == NULL) { return; }
  if (v->ops && v->ops->nvdestroy)
  {
    v->ops->nvdestroy(v);
    return;
  }
  /* no destroy op: free what the generic layer owns */
  free(v->content);
  N_VFreeEmpty(v);
}

void N_VSpace(N_Vector v, sunindextype* lrw, sunindextype* liw)
{
  if (v->ops->nvspace) { v->ops->nvspace(v, lrw, liw); }
  else
  {
    *lrw = 0;
    *liw = 0;
  }
}

sunrealtype* N_VGetArrayPointer(N_Vector v)
{
  return v->ops->nvgetarraypointer ? v->ops->nvgetarraypointer(v) : NULL;
}

sunrealtype* N_VGetDeviceArrayPointer(N_Vector v)
{
  return v->ops->nvgetdevicearraypointer ? v->ops->nvgetdevicearraypointer(v)
                                         : NULL;
}

void N_VSetArrayPointer(sunrealtype* v_data, N_Vector v)
{
  if (v->ops->nvsetarraypointer) { v->ops->nvsetarraypointer(v_data, v); }
}

SUNComm N_VGetCommunicator(N_Vector v)
{
  return v->ops->nvgetcommunicator ? v->ops->nvgetcommunicator(v) : SUN_COMM_NULL;
}

sunindextype N_VGetLength(N_Vector v) { return v->ops->nvgetlength(v); }

sunindextype N_VGetLocalLength(N_Vector v)
{
  return v->ops->nvgetlocallength ? v->ops->nvgetlocallength(v)
                                  : v->ops->nvgetlength(v);
}

void N_VLinearSum(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z)
{
  z->ops->nvlinearsum(a, x, b, y, z);
}
void N_VConst(sunrealtype c, N_Vector z) { z->ops->nvconst(c, z); }
void N_VProd(N_Vector x, N_Vector y, N_Vector z) { z->ops->nvprod(x, y, z); }
void N_VDiv(N_Vector x, N_Vector y, N_Vector z) { z->ops->nvdiv(x, y, z); }
void N_VScale(sunrealtype c, N_Vector x, N_Vector z) { z->ops->nvscale(c, x, z); }
void N_VAbs(N_Vector x, N_Vector z) { z->ops->nvabs(x, z); }
void N_VInv(N_Vector x, N_Vector z) { z->ops->nvinv(x, z); }
void N_VAddConst(N_Vector x, sunrealtype b, N_Vector z)
{
  z->ops->nvaddconst(x, b, z);
}
sunrealtype N_VDotProd(N_Vector x, N_Vector y) { return y->ops->nvdotprod(x, y); }
sunrealtype N_VMaxNorm(N_Vector x) { return x->ops->nvmaxnorm(x); }
sunrealtype N_VWrmsNorm(N_Vector x, N_Vector w) { return x->ops->nvwrmsnorm(x, w); }
sunrealtype N_VWrmsNormMask(N_Vector x, N_Vector w, N_Vector id)
{
  return x->ops->nvwrmsnormmask(x, w, id);
}
sunrealtype N_VMin(N_Vector x) { return x->ops->nvmin(x); }
sunrealtype N_VWL2Norm(N_Vector x, N_Vector w) { return x->ops->nvwl2norm(x, w); }
sunrealtype N_VL1Norm(N_Vector x) { return x->ops->nvl1norm(x); }
void N_VCompare(sunrealtype c, N_Vector x, N_Vector z)
{
  z->ops->nvcompare(c, x, z);
}
sunbooleantype N_VInvTest(N_Vector x, N_Vector z) { return z->ops->nvinvtest(x, z); }
sunbooleantype N_VConstrMask(N_Vector c, N_Vector x, N_Vector m)
{
  return x->ops->nvconstrmask(c, x, m);
}
sunrealtype N_VMinQuotient(N_Vector num, N_Vector denom)
{
  return num->ops->nvminquotient(num, denom);
}

/* fused operations: implementation if present, else the reference loops */
SUNErrCode N_VLinearCombination(int nvec, sunrealtype* c, N_Vector* X, N_Vector z)
{
  if (z->ops->nvlinearcombination) { return z->ops->nvlinearcombination(nvec, c, X, z); }
  z->ops->nvscale(c[0], X[0], z);
  for (int i = 1; i < nvec; i++)
  {
    z->ops->nvlinearsum(c[i], X[i], SUN_RCONST(1.0), z, z);
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VScaleAddMulti(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y,
                            N_Vector* Z)
{
  if (x->ops->nvscaleaddmulti) { return x->ops->nvscaleaddmulti(nvec, a, x, Y, Z); }
  for (int i = 0; i < nvec; i++)
  {
    x->ops->nvlinearsum(a[i], x, SUN_RCONST(1.0), Y[i], Z[i]);
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VDotProdMulti(int nvec, N_Vector x, N_Vector* Y, sunrealtype* dotprods)
{
  if (x->ops->nvdotprodmulti) { return x->ops->nvdotprodmulti(nvec, x, Y, dotprods); }
  for (int i = 0; i < nvec; i++) { dotprods[i] = x->ops->nvdotprod(x, Y[i]); }
  return SUN_SUCCESS;
}

SUNErrCode N_VLinearSumVectorArray(int nvec, sunrealtype a, N_Vector* X,
                                   sunrealtype b, N_Vector* Y, N_Vector* Z)
{
  if (Z[0]->ops->nvlinearsumvectorarray)
  {
    return Z[0]->ops->nvlinearsumvectorarray(nvec, a, X, b, Y, Z);
  }
  for (int i = 0; i < nvec; i++) { Z[0]->ops->nvlinearsum(a, X[i], b, Y[i], Z[i]); }
  return SUN_SUCCESS;
}

SUNErrCode N_VScaleVectorArray(int nvec, sunrealtype* c, N_Vector* X, N_Vector* Z)
{
  if (Z[0]->ops->nvscalevectorarray)
  {
    return Z[0]->ops->nvscalevectorarray(nvec, c, X, Z);
  }
  for (int i = 0; i < nvec; i++) { Z[0]->ops->nvscale(c[i], X[i], Z[i]); }
  return SUN_SUCCESS;
}

SUNErrCode N_VConstVectorArray(int nvec, sunrealtype c, N_Vector* Z)
{
  if (Z[0]->ops->nvconstvectorarray) { return Z[0]->ops->nvconstvectorarray(nvec, c, Z); }
  for (int i = 0; i < nvec; i++) { Z[0]->ops->nvconst(c, Z[i]); }
  return SUN_SUCCESS;
}

SUNErrCode N_VWrmsNormVectorArray(int nvec, N_Vector* X, N_Vector* W, sunrealtype* nrm)
{
  if (X[0]->ops->nvwrmsnormvectorarray)
  {
    return X[0]->ops->nvwrmsnormvectorarray(nvec, X, W, nrm);
  }
  for (int i = 0; i < nvec; i++) { nrm[i] = X[0]->ops->nvwrmsnorm(X[i], W[i]); }
  return SUN_SUCCESS;
}

SUNErrCode N_VWrmsNormMaskVectorArray(int nvec, N_Vector* X, N_Vector* W,
                                      N_Vector id, sunrealtype* nrm)
{
  if (id->ops->nvwrmsnormmaskvectorarray)
  {
    return id->ops->nvwrmsnormmaskvectorarray(nvec, X, W, id, nrm);
  }
  for (int i = 0; i < nvec; i++) { nrm[i] = id->ops->nvwrmsnormmask(X[i], W[i], id); }
  return SUN_SUCCESS;
}

SUNErrCode N_VScaleAddMultiVectorArray(int nvec, int nsum, sunrealtype* a,
                                       N_Vector* X, N_Vector** Y, N_Vector** Z)
{
  int ier = SUN_SUCCESS;
  if (X[0]->ops->nvscaleaddmultivectorarray)
  {
    return X[0]->ops->nvscaleaddmultivectorarray(nvec, nsum, a, X, Y, Z);
  }
  if (X[0]->ops->nvscaleaddmulti)
  {
    N_Vector* YY = (N_Vector*)malloc(nsum * sizeof(N_Vector));
    N_Vector* ZZ = (N_Vector*)malloc(nsum * sizeof(N_Vector));
    if (!YY || !ZZ)
    {
      free(YY);
      free(ZZ);
      return SUN_ERR_MALLOC_FAIL;
    }
    for (int i = 0; i < nvec; i++)
    {
      for (int j = 0; j < nsum; j++)
      {
        YY[j] = Y[j][i];
        ZZ[j] = Z[j][i];
      }
      ier = X[0]->ops->nvscaleaddmulti(nsum, a, X[i], YY, ZZ);
      if (ier != 0) { break; }
    }
    free(YY);
    free(ZZ);
    return ier;
  }
  for (int i = 0; i < nvec; i++)
  {
    for (int j = 0; j < nsum; j++)
    {
      X[0]->ops->nvlinearsum(a[j], X[i], SUN_RCONST(1.0), Y[j][i], Z[j][i]);
    }
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VLinearCombinationVectorArray(int nvec, int nsum, sunrealtype* c,
                                           N_Vector** X, N_Vector* Z)
{
  int ier = SUN_SUCCESS;
  if (Z[0]->ops->nvlinearcombinationvectorarray)
  {
    return Z[0]->ops->nvlinearcombinationvectorarray(nvec, nsum, c, X, Z);
  }
  if (Z[0]->ops->nvlinearcombination)
  {
    N_Vector* Yv = (N_Vector*)malloc(nsum * sizeof(N_Vector));
    if (!Yv) { return SUN_ERR_MALLOC_FAIL; }
    for (int i = 0; i < nvec; i++)
    {
      for (int j = 0; j < nsum; j++) { Yv[j] = X[j][i]; }
      ier = Z[0]->ops->nvlinearcombination(nsum, c, Yv, Z[i]);
      if (ier != 0) { break; }
    }
    free(Yv);
    return ier;
  }
  for (int i = 0; i < nvec; i++)
  {
    Z[0]->ops->nvscale(c[0], X[0][i], Z[i]);
    for (int j = 1; j < nsum; j++)
    {
      Z[0]->ops->nvlinearsum(c[j], X[j][i], SUN_RCONST(1.0), Z[i], Z[i]);
    }
  }
  return SUN_SUCCESS;
}

/* local reductions: only dispatch (no default in the reference) */
sunrealtype N_VDotProdLocal(N_Vector x, N_Vector y) { return y->ops->nvdotprodlocal(x, y); }
sunrealtype N_VMaxNormLocal(N_Vector x) { return x->ops->nvmaxnormlocal(x); }
sunrealtype N_VMinLocal(N_Vector x) { return x->ops->nvminlocal(x); }
sunrealtype N_VL1NormLocal(N_Vector x) { return x->ops->nvl1normlocal(x); }
sunrealtype N_VWSqrSumLocal(N_Vector x, N_Vector w) { return x->ops->nvwsqrsumlocal(x, w); }
sunrealtype N_VWSqrSumMaskLocal(N_Vector x, N_Vector w, N_Vector id)
{
  return x->ops->nvwsqrsummasklocal(x, w, id);
}
sunbooleantype N_VInvTestLocal(N_Vector x, N_Vector z) { return z->ops->nvinvtestlocal(x, z); }
sunbooleantype N_VConstrMaskLocal(N_Vector c, N_Vector x, N_Vector m)
{
  return x->ops->nvconstrmasklocal(c, x, m);
}
sunrealtype N_VMinQuotientLocal(N_Vector num, N_Vector denom)
{
  return num->ops->nvminquotientlocal(num, denom);
}
SUNErrCode N_VDotProdMultiLocal(int nvec, N_Vector x, N_Vector* Y, sunrealtype* d)
{
  if (!x->ops->nvdotprodmultilocal) { return SUN_ERR_NOT_IMPLEMENTED; }
  return x->ops->nvdotprodmultilocal(nvec, x, Y, d);
}
SUNErrCode N_VDotProdMultiAllReduce(int nvec_total, N_Vector x, sunrealtype* sum)
{
  if (!x->ops->nvdotprodmultiallreduce) { return SUN_ERR_NOT_IMPLEMENTED; }
  return x->ops->nvdotprodmultiallreduce(nvec_total, x, sum);
}
SUNErrCode N_VBufSize(N_Vector x, sunindextype* size)
{
  if (!x->ops->nvbufsize) { return SUN_ERR_NOT_IMPLEMENTED; }
  return x->ops->nvbufsize(x, size);
}
SUNErrCode N_VBufPack(N_Vector x, void* buf)
{
  if (!x->ops->nvbufpack) { return SUN_ERR_NOT_IMPLEMENTED; }
  return x->ops->nvbufpack(x, buf);
}
SUNErrCode N_VBufUnpack(N_Vector x, void* buf)
{
  if (!x->ops->nvbufunpack) { return SUN_ERR_NOT_IMPLEMENTED; }
  return x->ops->nvbufunpack(x, buf);
}

N_Vector* N_VNewVectorArray(int count, SUNContext sunctx)
{
  (void)sunctx;
  if (count <= 0) { return NULL; }
  return (N_Vector*)calloc((size_t)count, sizeof(N_Vector));
}

N_Vector* N_VCloneEmptyVectorArray(int count, N_Vector w)
{
  if (count <= 0) { return NULL; }
  N_Vector* vs = (N_Vector*)malloc((size_t)count * sizeof(N_Vector));
  if (!vs) { return NULL; }
  for (int j = 0; j < count; j++)
  {
    vs[j] = N_VCloneEmpty(w);
    if (!vs[j])
    {
      N_VDestroyVectorArray(vs, j);
      return NULL;
    }
  }
  return vs;
}

N_Vector* N_VCloneVectorArray(int count, N_Vector w)
{
  if (count <= 0) { return NULL; }
  N_Vector* vs = (N_Vector*)malloc((size_t)count * sizeof(N_Vector));
  if (!vs) { return NULL; }
  for (int j = 0; j < count; j++)
  {
    vs[j] = N_VClone(w);
    if (!vs[j])
    {
      N_VDestroyVectorArray(vs, j);
      return NULL;
    }
  }
  return vs;
}

void N_VDestroyVectorArray(N_Vector* vs, int count)
{
  if (vs == NULL) { return; }
  for (int j = 0; j < count; j++)
  {
    N_VDestroy(vs[j]);
    vs[j] = NULL;
  }
  free(vs);
}

void N_VPrint(N_Vector v)
{
  if (v->ops->nvprint) { v->ops->nvprint(v); }
}
void N_VPrintFile(N_Vector v, FILE* outfile)
{
  if (v->ops->nvprintfile) { v->ops->nvprintfile(v, outfile); }
}

/* ------------------------------------------------------------------------- */
/* SUNMatrix                                                                 */
/* ------------------------------------------------------------------------- */

SUNMatrix SUNMatNewEmpty(SUNContext sunctx)
{
  if (!sunctx) { return NULL; }
  SUNMatrix A = (SUNMatrix)malloc(sizeof(*A));
  if (!A) { return NULL; }
  SUNMatrix_Ops ops = (SUNMatrix_Ops)calloc(1, sizeof(*ops));
  if (!ops)
  {
    free(A);
    return NULL;
  }
  A->ops     = ops;
  A->content = NULL;
  A->sunctx  = sunctx;
  return A;
}

void SUNMatFreeEmpty(SUNMatrix A)
{
  if (A == NULL) { return; }
  free(A->ops);
  A->ops = NULL;
  free(A);
}

SUNErrCode SUNMatCopyOps(SUNMatrix A, SUNMatrix B)
{
  if (!A || !B || !A->ops || !B->ops) { return SUN_ERR_ARG_CORRUPT; }
  memcpy(B->ops, A->ops, sizeof(*B->ops));
  return SUN_SUCCESS;
}

SUNMatrix_ID SUNMatGetID(SUNMatrix A) { return A->ops->getid(A); }
SUNMatrix SUNMatClone(SUNMatrix A) { return A->ops->clone(A); }
void SUNMatDestroy(SUNMatrix A)
{
  if (A == NULL) { return; }
  if (A->ops && A->ops->destroy)
  {
    A->ops->destroy(A);
    return;
  }
  SUNMatFreeEmpty(A);
}
SUNErrCode SUNMatZero(SUNMatrix A) { return A->ops->zero(A); }
SUNErrCode SUNMatCopy(SUNMatrix A, SUNMatrix B) { return A->ops->copy(A, B); }
SUNErrCode SUNMatScaleAdd(sunrealtype c, SUNMatrix A, SUNMatrix B)
{
  return A->ops->scaleadd(c, A, B);
}
SUNErrCode SUNMatScaleAddI(sunrealtype c, SUNMatrix A) { return A->ops->scaleaddi(c, A); }
SUNErrCode SUNMatMatvecSetup(SUNMatrix A)
{
  return A->ops->matvecsetup ? A->ops->matvecsetup(A) : SUN_SUCCESS;
}
SUNErrCode SUNMatMatvec(SUNMatrix A, N_Vector x, N_Vector y)
{
  return A->ops->matvec(A, x, y);
}
SUNErrCode SUNMatSpace(SUNMatrix A, long int* lenrw, long int* leniw)
{
  if (A->ops->space) { return A->ops->space(A, lenrw, leniw); }
  *lenrw = 0;
  *leniw = 0;
  return SUN_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* SUNLinearSolver                                                           */
/* ------------------------------------------------------------------------- */

SUNLinearSolver SUNLinSolNewEmpty(SUNContext sunctx)
{
  if (!sunctx) { return NULL; }
  SUNLinearSolver S = (SUNLinearSolver)malloc(sizeof(*S));
  if (!S) { return NULL; }
  SUNLinearSolver_Ops ops = (SUNLinearSolver_Ops)calloc(1, sizeof(*ops));
  if (!ops)
  {
    free(S);
    return NULL;
  }
  S->ops     = ops;
  S->content = NULL;
  S->sunctx  = sunctx;
  return S;
}

void SUNLinSolFreeEmpty(SUNLinearSolver S)
{
  if (S == NULL) { return; }
  free(S->ops);
  S->ops = NULL;
  free(S);
}

SUNLinearSolver_Type SUNLinSolGetType(SUNLinearSolver S) { return S->ops->gettype(S); }
SUNLinearSolver_ID SUNLinSolGetID(SUNLinearSolver S)
{
  return S->ops->getid ? S->ops->getid(S) : SUNLINEARSOLVER_CUSTOM;
}
SUNErrCode SUNLinSolSetATimes(SUNLinearSolver S, void* A_data, SUNATimesFn ATimes)
{
  return S->ops->setatimes ? S->ops->setatimes(S, A_data, ATimes) : SUN_SUCCESS;
}
SUNErrCode SUNLinSolSetPreconditioner(SUNLinearSolver S, void* P_data,
                                      SUNPSetupFn Pset, SUNPSolveFn Psol)
{
  return S->ops->setpreconditioner ? S->ops->setpreconditioner(S, P_data, Pset, Psol)
                                   : SUN_SUCCESS;
}
SUNErrCode SUNLinSolSetScalingVectors(SUNLinearSolver S, N_Vector s1, N_Vector s2)
{
  return S->ops->setscalingvectors ? S->ops->setscalingvectors(S, s1, s2)
                                   : SUN_SUCCESS;
}
SUNErrCode SUNLinSolSetZeroGuess(SUNLinearSolver S, sunbooleantype onoff)
{
  return S->ops->setzeroguess ? S->ops->setzeroguess(S, onoff) : SUN_SUCCESS;
}
SUNErrCode SUNLinSolInitialize(SUNLinearSolver S)
{
  return S->ops->initialize ? S->ops->initialize(S) : SUN_SUCCESS;
}
int SUNLinSolSetup(SUNLinearSolver S, SUNMatrix A)
{
  return S->ops->setup ? S->ops->setup(S, A) : SUN_SUCCESS;
}
int SUNLinSolSolve(SUNLinearSolver S, SUNMatrix A, N_Vector x, N_Vector b,
                   sunrealtype tol)
{
  return S->ops->solve(S, A, x, b, tol);
}
int SUNLinSolNumIters(SUNLinearSolver S)
{
  return S->ops->numiters ? S->ops->numiters(S) : 0;
}
sunrealtype SUNLinSolResNorm(SUNLinearSolver S)
{
  return S->ops->resnorm ? S->ops->resnorm(S) : SUN_RCONST(0.0);
}
N_Vector SUNLinSolResid(SUNLinearSolver S)
{
  return S->ops->resid ? S->ops->resid(S) : NULL;
}
sunindextype SUNLinSolLastFlag(SUNLinearSolver S)
{
  return S->ops->lastflag ? S->ops->lastflag(S) : SUN_SUCCESS;
}
SUNErrCode SUNLinSolSpace(SUNLinearSolver S, long int* lenrwLS, long int* leniwLS)
{
  if (S->ops->space) { return S->ops->space(S, lenrwLS, leniwLS); }
  *lenrwLS = 0;
  *leniwLS = 0;
  return SUN_SUCCESS;
}
SUNErrCode SUNLinSolFree(SUNLinearSolver S)
{
  if (S == NULL) { return SUN_SUCCESS; }
  if (S->ops && S->ops->free) { return S->ops->free(S); }
  SUNLinSolFreeEmpty(S);
  return SUN_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* SUNNonlinearSolver                                                        */
/* ------------------------------------------------------------------------- */

SUNNonlinearSolver SUNNonlinSolNewEmpty(SUNContext sunctx)
{
  if (!sunctx) { return NULL; }
  SUNNonlinearSolver NLS = (SUNNonlinearSolver)malloc(sizeof(*NLS));
  if (!NLS) { return NULL; }
  SUNNonlinearSolver_Ops ops = (SUNNonlinearSolver_Ops)calloc(1, sizeof(*ops));
  if (!ops)
  {
    free(NLS);
    return NULL;
  }
  NLS->ops     = ops;
  NLS->content = NULL;
  NLS->sunctx  = sunctx;
  return NLS;
}

void SUNNonlinSolFreeEmpty(SUNNonlinearSolver NLS)
{
  if (NLS == NULL) { return; }
  free(NLS->ops);
  NLS->ops = NULL;
  free(NLS);
}

SUNNonlinearSolver_Type SUNNonlinSolGetType(SUNNonlinearSolver NLS)
{
  return NLS->ops->gettype(NLS);
}
SUNErrCode SUNNonlinSolInitialize(SUNNonlinearSolver NLS)
{
  return NLS->ops->initialize ? NLS->ops->initialize(NLS) : SUN_SUCCESS;
}
int SUNNonlinSolSetup(SUNNonlinearSolver NLS, N_Vector y, void* mem)
{
  return NLS->ops->setup ? NLS->ops->setup(NLS, y, mem) : SUN_SUCCESS;
}
int SUNNonlinSolSolve(SUNNonlinearSolver NLS, N_Vector y0, N_Vector y, N_Vector w,
                      sunrealtype tol, sunbooleantype callLSetup, void* mem)
{
  return NLS->ops->solve(NLS, y0, y, w, tol, callLSetup, mem);
}
SUNErrCode SUNNonlinSolFree(SUNNonlinearSolver NLS)
{
  if (NLS == NULL) { return SUN_SUCCESS; }
  if (NLS->ops && NLS->ops->free) { return NLS->ops->free(NLS); }
  SUNNonlinSolFreeEmpty(NLS);
  return SUN_SUCCESS;
}
SUNErrCode SUNNonlinSolSetSysFn(SUNNonlinearSolver NLS, SUNNonlinSolSysFn SysFn)
{
  return NLS->ops->setsysfn(NLS, SysFn);
}
SUNErrCode SUNNonlinSolSetLSetupFn(SUNNonlinearSolver NLS, SUNNonlinSolLSetupFn f)
{
  return NLS->ops->setlsetupfn ? NLS->ops->setlsetupfn(NLS, f) : SUN_SUCCESS;
}
SUNErrCode SUNNonlinSolSetLSolveFn(SUNNonlinearSolver NLS, SUNNonlinSolLSolveFn f)
{
  return NLS->ops->setlsolvefn ? NLS->ops->setlsolvefn(NLS, f) : SUN_SUCCESS;
}
SUNErrCode SUNNonlinSolSetConvTestFn(SUNNonlinearSolver NLS,
                                     SUNNonlinSolConvTestFn f, void* ctest_data)
{
  return NLS->ops->setctestfn ? NLS->ops->setctestfn(NLS, f, ctest_data)
                              : SUN_SUCCESS;
}
SUNErrCode SUNNonlinSolSetMaxIters(SUNNonlinearSolver NLS, int maxiters)
{
  return NLS->ops->setmaxiters ? NLS->ops->setmaxiters(NLS, maxiters) : SUN_SUCCESS;
}
SUNErrCode SUNNonlinSolGetNumIters(SUNNonlinearSolver NLS, long int* niters)
{
  if (NLS->ops->getnumiters) { return NLS->ops->getnumiters(NLS, niters); }
  *niters = 0;
  return SUN_SUCCESS;
}
SUNErrCode SUNNonlinSolGetCurIter(SUNNonlinearSolver NLS, int* iter)
{
  if (NLS->ops->getcuriter) { return NLS->ops->getcuriter(NLS, iter); }
  *iter = -1;
  return SUN_ERR_NOT_IMPLEMENTED;
}
SUNErrCode SUNNonlinSolGetNumConvFails(SUNNonlinearSolver NLS, long int* nconvfails)
{
  if (NLS->ops->getnumconvfails) { return NLS->ops->getnumconvfails(NLS, nconvfails); }
  *nconvfails = 0;
  return SUN_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* math helpers                                                              */
/* ------------------------------------------------------------------------- */

sunrealtype SUNRpowerI(sunrealtype base, int exponent)
{
  sunrealtype prod = SUN_RCONST(1.0);
  int expt         = abs(exponent);
  for (int i = 1; i <= expt; i++) { prod *= base; }
  if (exponent < 0) { prod = SUN_RCONST(1.0) / prod; }
  return prod;
}

sunbooleantype SUNRCompareTol(sunrealtype a, sunrealtype b, sunrealtype tol)
{
  if (a == b) { return SUNFALSE; }
  sunrealtype diff = fabs(a - b);
  sunrealtype norm = fmin(fabs(a + b), SUN_BIG_REAL);
  sunrealtype lim  = 10 * SUN_UNIT_ROUNDOFF;
  if (tol * norm > lim) { lim = tol * norm; }
  return !isless(diff, lim);
}

sunbooleantype SUNRCompare(sunrealtype a, sunrealtype b)
{
  return SUNRCompareTol(a, b, 10 * SUN_UNIT_ROUNDOFF);
}
