/* -----------------------------------------------------------------------------
 * Scaled, preconditioned GMRES(maxl) with restarts: restatement of
 * src/sunlinsol/spgmr/sunlinsol_spgmr.c
 *   constructor and option setters   :61-350
 *   initialize (workspace)           :371-436
 *   setup (preconditioner setup)     :470-497
 *   solve (Arnoldi loop, QR, update) :499-930
 * Defaults from include/sunlinsol/sunlinsol_spgmr.h:44-46 (maxl 5, no
 * restarts, modified Gram-Schmidt).  The vector work runs through the
 * N_Vector ops of whatever vector the solver is built on; the Gram-Schmidt
 * step is SUNModifiedGS / SUNClassicalGS unless a caller installs another
 * routine with the classical signature (SUNLinSol_SPGMRSetClassicalGSFn),
 * which is how a device vector supplies a fused Gram-Schmidt step.
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>

#include "sunamd/sundials_core.h"
#include "sunamd/sundials_iterative.h"
#include "sunamd/sunlinsol_spgmr.h"

typedef struct
{
  int maxl, pretype, gstype, max_restarts;
  sunbooleantype zeroguess;
  int numiters;
  sunrealtype resnorm;
  int last_flag;
  SUNATimesFn ATimes;
  void* ATData;
  SUNPSetupFn Psetup;
  SUNPSolveFn Psolve;
  void* PData;
  N_Vector s1, s2;
  N_Vector* V;
  sunrealtype** Hes;
  sunrealtype* givens;
  N_Vector xcor;
  sunrealtype* yg;
  N_Vector vtemp;
  sunrealtype* cv;
  N_Vector* Xv;
  SUNAMD_ClassicalGSFn cgs;
} SpgmrContent;

#define SPG(S) ((SpgmrContent*)((S)->content))

static SUNLinearSolver_Type spgmr_gettype(SUNLinearSolver S)
{
  (void)S;
  return SUNLINEARSOLVER_ITERATIVE;
}

static SUNLinearSolver_ID spgmr_getid(SUNLinearSolver S)
{
  (void)S;
  return SUNLINEARSOLVER_SPGMR;
}

static SUNErrCode spgmr_setatimes(SUNLinearSolver S, void* data, SUNATimesFn fn)
{
  SPG(S)->ATimes = fn;
  SPG(S)->ATData = data;
  return SUN_SUCCESS;
}

static SUNErrCode spgmr_setprec(SUNLinearSolver S, void* data, SUNPSetupFn setup,
                                SUNPSolveFn solve)
{
  SPG(S)->Psetup = setup;
  SPG(S)->Psolve = solve;
  SPG(S)->PData  = data;
  return SUN_SUCCESS;
}

static SUNErrCode spgmr_setscaling(SUNLinearSolver S, N_Vector s1, N_Vector s2)
{
  SPG(S)->s1 = s1;
  SPG(S)->s2 = s2;
  return SUN_SUCCESS;
}

static SUNErrCode spgmr_setzeroguess(SUNLinearSolver S, sunbooleantype onoff)
{
  SPG(S)->zeroguess = onoff;
  return SUN_SUCCESS;
}

static SUNErrCode spgmr_initialize(SUNLinearSolver S)
{
  SpgmrContent* c = SPG(S);
  if (c->max_restarts < 0) { c->max_restarts = SUNSPGMR_MAXRS_DEFAULT; }
  if (!c->ATimes) { return SUN_ERR_ARG_CORRUPT; }
  if ((c->pretype != SUN_PREC_LEFT) && (c->pretype != SUN_PREC_RIGHT) &&
      (c->pretype != SUN_PREC_BOTH))
  {
    c->pretype = SUN_PREC_NONE;
  }
  if ((c->pretype != SUN_PREC_NONE) && !c->Psolve) { return SUN_ERR_ARG_CORRUPT; }
  if (!c->V)
  {
    c->V = N_VCloneVectorArray(c->maxl + 1, c->vtemp);
    if (!c->V) { return SUN_ERR_MEM_FAIL; }
  }
  if (!c->Hes)
  {
    c->Hes = (sunrealtype**)calloc((size_t)c->maxl + 1, sizeof(sunrealtype*));
    if (!c->Hes) { return SUN_ERR_MALLOC_FAIL; }
    for (int k = 0; k <= c->maxl; k++)
    {
      c->Hes[k] = (sunrealtype*)malloc(sizeof(sunrealtype) * (size_t)c->maxl);
      if (!c->Hes[k]) { return SUN_ERR_MALLOC_FAIL; }
    }
  }
  if (!c->givens) { c->givens = (sunrealtype*)malloc(sizeof(sunrealtype) * 2 * (size_t)c->maxl); }
  if (!c->yg) { c->yg = (sunrealtype*)malloc(sizeof(sunrealtype) * ((size_t)c->maxl + 1)); }
  if (!c->cv) { c->cv = (sunrealtype*)malloc(sizeof(sunrealtype) * ((size_t)c->maxl + 1)); }
  if (!c->Xv) { c->Xv = (N_Vector*)malloc(sizeof(N_Vector) * ((size_t)c->maxl + 1)); }
  if (!c->givens || !c->yg || !c->cv || !c->Xv) { return SUN_ERR_MALLOC_FAIL; }
  return SUN_SUCCESS;
}

static int spgmr_setup(SUNLinearSolver S, SUNMatrix A)
{
  (void)A;
  SpgmrContent* c = SPG(S);
  if (c->Psetup)
  {
    int status = c->Psetup(c->PData);
    if (status != 0)
    {
      c->last_flag = (status < 0) ? SUNLS_PSET_FAIL_UNREC : SUNLS_PSET_FAIL_REC;
      return c->last_flag;
    }
  }
  c->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}

/* record a failed callback and leave the zero-guess flag cleared */
static int spgmr_fail(SpgmrContent* c, int flag)
{
  c->zeroguess = SUNFALSE;
  c->last_flag = flag;
  return flag;
}

/* x = x0 + P2^{-1} S2^{-1} xcor (x0 = 0 under a zero guess) */
static int spgmr_finish(SpgmrContent* c, N_Vector x, sunrealtype delta, sunbooleantype preR,
                        sunbooleantype scale2, int flag)
{
  if (scale2) { N_VDiv(c->xcor, c->s2, c->xcor); }
  if (preR)
  {
    int status = c->Psolve(c->PData, c->xcor, c->vtemp, delta, SUN_PREC_RIGHT);
    if (status != 0)
    {
      return spgmr_fail(c, (status < 0) ? SUNLS_PSOLVE_FAIL_UNREC : SUNLS_PSOLVE_FAIL_REC);
    }
  }
  else { N_VScale(1.0, c->xcor, c->vtemp); }
  if (c->zeroguess) { N_VScale(1.0, c->vtemp, x); }
  else { N_VLinearSum(1.0, x, 1.0, c->vtemp, x); }
  c->zeroguess = SUNFALSE;
  c->last_flag = flag;
  return flag;
}

static int spgmr_solve(SUNLinearSolver S, SUNMatrix A, N_Vector x, N_Vector b, sunrealtype delta)
{
  (void)A;
  SpgmrContent* c   = SPG(S);
  const int l_max   = c->maxl;
  N_Vector* V       = c->V;
  sunrealtype** Hes = c->Hes;
  sunrealtype* q    = c->givens;
  sunrealtype* yg   = c->yg;
  N_Vector vtemp    = c->vtemp;
  const sunbooleantype preL   = (c->pretype == SUN_PREC_LEFT) || (c->pretype == SUN_PREC_BOTH);
  const sunbooleantype preR   = (c->pretype == SUN_PREC_RIGHT) || (c->pretype == SUN_PREC_BOTH);
  const sunbooleantype scale1 = (c->s1 != NULL);
  const sunbooleantype scale2 = (c->s2 != NULL);
  int status, krydim = 0;
  sunbooleantype converged = SUNFALSE;

  c->numiters = 0;
  if (!c->ATimes) { return spgmr_fail(c, SUN_ERR_ARG_CORRUPT); }
  if ((preL || preR) && !c->Psolve) { return spgmr_fail(c, SUN_ERR_ARG_CORRUPT); }

  /* r0 = b - A x0 */
  if (c->zeroguess) { N_VScale(1.0, b, vtemp); }
  else
  {
    status = c->ATimes(c->ATData, x, vtemp);
    if (status != 0)
    {
      return spgmr_fail(c, (status < 0) ? SUNLS_ATIMES_FAIL_UNREC : SUNLS_ATIMES_FAIL_REC);
    }
    N_VLinearSum(1.0, b, -1.0, vtemp, vtemp);
  }
  N_VScale(1.0, vtemp, V[0]);

  /* V[0] = S1 P1^{-1} r0 */
  if (preL)
  {
    status = c->Psolve(c->PData, V[0], vtemp, delta, SUN_PREC_LEFT);
    if (status != 0)
    {
      return spgmr_fail(c, (status < 0) ? SUNLS_PSOLVE_FAIL_UNREC : SUNLS_PSOLVE_FAIL_REC);
    }
  }
  else { N_VScale(1.0, V[0], vtemp); }
  if (scale1) { N_VProd(c->s1, vtemp, V[0]); }
  else { N_VScale(1.0, vtemp, V[0]); }

  double r_norm = N_VDotProd(V[0], V[0]);
  r_norm        = (r_norm <= 0.0) ? 0.0 : sqrt(r_norm);
  const double beta = r_norm;
  c->resnorm        = r_norm;
  if (r_norm <= delta)
  {
    c->zeroguess = SUNFALSE;
    c->last_flag = SUN_SUCCESS;
    return SUN_SUCCESS;
  }
  double rho = beta;
  N_VConst(0.0, c->xcor);

  for (int ntries = 0; ntries <= c->max_restarts; ntries++)
  {
    for (int i = 0; i <= l_max; i++)
    {
      for (int j = 0; j < l_max; j++) { Hes[i][j] = 0.0; }
    }
    double rotation_product = 1.0;
    N_VScale(1.0 / r_norm, V[0], V[0]);

    for (int l = 0; l < l_max; l++)
    {
      c->numiters++;
      const int lp1 = l + 1;
      krydim        = lp1;

      /* V[l+1] = S1 P1^{-1} A P2^{-1} S2^{-1} V[l] */
      if (scale2) { N_VDiv(V[l], c->s2, vtemp); }
      else { N_VScale(1.0, V[l], vtemp); }
      if (preR)
      {
        N_VScale(1.0, vtemp, V[lp1]);
        status = c->Psolve(c->PData, V[lp1], vtemp, delta, SUN_PREC_RIGHT);
        if (status != 0)
        {
          return spgmr_fail(c, (status < 0) ? SUNLS_PSOLVE_FAIL_UNREC : SUNLS_PSOLVE_FAIL_REC);
        }
      }
      status = c->ATimes(c->ATData, vtemp, V[lp1]);
      if (status != 0)
      {
        return spgmr_fail(c, (status < 0) ? SUNLS_ATIMES_FAIL_UNREC : SUNLS_ATIMES_FAIL_REC);
      }
      if (preL)
      {
        status = c->Psolve(c->PData, V[lp1], vtemp, delta, SUN_PREC_LEFT);
        if (status != 0)
        {
          return spgmr_fail(c, (status < 0) ? SUNLS_PSOLVE_FAIL_UNREC : SUNLS_PSOLVE_FAIL_REC);
        }
      }
      else { N_VScale(1.0, V[lp1], vtemp); }
      if (scale1) { N_VProd(c->s1, vtemp, V[lp1]); }
      else { N_VScale(1.0, vtemp, V[lp1]); }

      /* orthogonalise against V[0..l]; the new norm lands in Hes[l+1][l] */
      SUNErrCode err;
      if (c->gstype == SUN_CLASSICAL_GS)
      {
        err = c->cgs(V, Hes, lp1, l_max, &(Hes[lp1][l]), c->cv, c->Xv);
      }
      else { err = SUNModifiedGS(V, Hes, lp1, l_max, &(Hes[lp1][l])); }
      if (err != SUN_SUCCESS) { return spgmr_fail(c, err); }

      if (SUNQRfact(krydim, Hes, q, l) != 0) { return spgmr_fail(c, SUNLS_QRFACT_FAIL); }

      rotation_product *= q[2 * l + 1];
      rho        = fabs(rotation_product * r_norm);
      c->resnorm = rho;
      if (rho <= delta)
      {
        converged = SUNTRUE;
        break;
      }
      N_VScale(1.0 / Hes[lp1][l], V[lp1], V[lp1]);
    }

    /* xcor += V y with R y = Q^T (r_norm e1) */
    yg[0] = r_norm;
    for (int i = 1; i <= krydim; i++) { yg[i] = 0.0; }
    if (SUNQRsol(krydim, Hes, q, yg) != 0) { return spgmr_fail(c, SUNLS_QRSOL_FAIL); }
    c->cv[0] = 1.0;
    c->Xv[0] = c->xcor;
    for (int k = 0; k < krydim; k++)
    {
      c->cv[k + 1] = yg[k];
      c->Xv[k + 1] = V[k];
    }
    SUNErrCode err = N_VLinearCombination(krydim + 1, c->cv, c->Xv, c->xcor);
    if (err != SUN_SUCCESS) { return spgmr_fail(c, err); }

    if (converged) { return spgmr_finish(c, x, delta, preR, scale2, SUN_SUCCESS); }
    if (ntries == c->max_restarts) { break; }

    /* restart from the last residual: V[0] = V_(krydim+1) * (r_norm * Q e_last) */
    double s_product = 1.0;
    for (int i = krydim; i > 0; i--)
    {
      yg[i] = s_product * q[2 * i - 2];
      s_product *= q[2 * i - 1];
    }
    yg[0] = s_product;
    r_norm *= s_product;
    for (int i = 0; i <= krydim; i++) { yg[i] *= r_norm; }
    r_norm = fabs(r_norm);
    for (int k = 0; k <= krydim; k++)
    {
      c->cv[k] = yg[k];
      c->Xv[k] = V[k];
    }
    err = N_VLinearCombination(krydim + 1, c->cv, c->Xv, V[0]);
    if (err != SUN_SUCCESS) { return spgmr_fail(c, err); }
  }

  /* no convergence: keep the iterate if it reduced the residual */
  if (rho < beta) { return spgmr_finish(c, x, delta, preR, scale2, SUNLS_RES_REDUCED); }
  return spgmr_fail(c, SUNLS_CONV_FAIL);
}

static int spgmr_numiters(SUNLinearSolver S) { return SPG(S)->numiters; }

static sunrealtype spgmr_resnorm(SUNLinearSolver S) { return SPG(S)->resnorm; }

static N_Vector spgmr_resid(SUNLinearSolver S) { return SPG(S)->vtemp; }

static sunindextype spgmr_lastflag(SUNLinearSolver S) { return SPG(S)->last_flag; }

static SUNErrCode spgmr_space(SUNLinearSolver S, long int* lenrw, long int* leniw)
{
  SpgmrContent* c = SPG(S);
  sunindextype lrw1 = 0, liw1 = 0;
  if (c->vtemp->ops->nvspace) { N_VSpace(c->vtemp, &lrw1, &liw1); }
  long int lrw = (long int)lrw1 * (c->maxl + 5) + c->maxl * (c->maxl + 5) + 2;
  long int liw = (long int)liw1 * (c->maxl + 5);
  if (lenrw) { *lenrw = lrw; }
  if (leniw) { *leniw = liw; }
  return SUN_SUCCESS;
}

static SUNErrCode spgmr_free(SUNLinearSolver S)
{
  if (!S) { return SUN_SUCCESS; }
  SpgmrContent* c = SPG(S);
  if (c)
  {
    if (c->xcor) { N_VDestroy(c->xcor); }
    if (c->vtemp) { N_VDestroy(c->vtemp); }
    if (c->V) { N_VDestroyVectorArray(c->V, c->maxl + 1); }
    if (c->Hes)
    {
      for (int k = 0; k <= c->maxl; k++) { free(c->Hes[k]); }
      free(c->Hes);
    }
    free(c->givens);
    free(c->yg);
    free(c->cv);
    free(c->Xv);
    free(c);
    S->content = NULL;
  }
  SUNLinSolFreeEmpty(S);
  return SUN_SUCCESS;
}

SUNLinearSolver SUNLinSol_SPGMR(N_Vector y, int pretype, int maxl, SUNContext sunctx)
{
  if ((pretype != SUN_PREC_NONE) && (pretype != SUN_PREC_LEFT) && (pretype != SUN_PREC_RIGHT) &&
      (pretype != SUN_PREC_BOTH))
  {
    pretype = SUN_PREC_NONE;
  }
  if (maxl <= 0) { maxl = SUNSPGMR_MAXL_DEFAULT; }
  if (!y || !y->ops->nvclone || !y->ops->nvdestroy || !y->ops->nvlinearsum ||
      !y->ops->nvconst || !y->ops->nvprod || !y->ops->nvdiv || !y->ops->nvscale ||
      !y->ops->nvdotprod)
  {
    return NULL;
  }
  SUNLinearSolver S = SUNLinSolNewEmpty(sunctx);
  if (!S) { return NULL; }
  S->ops->gettype           = spgmr_gettype;
  S->ops->getid             = spgmr_getid;
  S->ops->setatimes         = spgmr_setatimes;
  S->ops->setpreconditioner = spgmr_setprec;
  S->ops->setscalingvectors = spgmr_setscaling;
  S->ops->setzeroguess      = spgmr_setzeroguess;
  S->ops->initialize        = spgmr_initialize;
  S->ops->setup             = spgmr_setup;
  S->ops->solve             = spgmr_solve;
  S->ops->numiters          = spgmr_numiters;
  S->ops->resnorm           = spgmr_resnorm;
  S->ops->resid             = spgmr_resid;
  S->ops->lastflag          = spgmr_lastflag;
  S->ops->space             = spgmr_space;
  S->ops->free              = spgmr_free;
  SpgmrContent* c = (SpgmrContent*)calloc(1, sizeof(SpgmrContent));
  if (!c)
  {
    SUNLinSolFreeEmpty(S);
    return NULL;
  }
  S->content      = c;
  c->maxl         = maxl;
  c->pretype      = pretype;
  c->gstype       = SUNSPGMR_GSTYPE_DEFAULT;
  c->max_restarts = SUNSPGMR_MAXRS_DEFAULT;
  c->cgs          = SUNClassicalGS;
  c->xcor         = N_VClone(y);
  c->vtemp        = N_VClone(y);
  if (!c->xcor || !c->vtemp)
  {
    spgmr_free(S);
    return NULL;
  }
  return S;
}

SUNErrCode SUNLinSol_SPGMRSetPrecType(SUNLinearSolver S, int pretype)
{
  if ((pretype != SUN_PREC_NONE) && (pretype != SUN_PREC_LEFT) && (pretype != SUN_PREC_RIGHT) &&
      (pretype != SUN_PREC_BOTH))
  {
    return SUN_ERR_ARG_OUTOFRANGE;
  }
  SPG(S)->pretype = pretype;
  return SUN_SUCCESS;
}

SUNErrCode SUNLinSol_SPGMRSetGSType(SUNLinearSolver S, int gstype)
{
  if ((gstype != SUN_MODIFIED_GS) && (gstype != SUN_CLASSICAL_GS)) { return SUN_ERR_ARG_OUTOFRANGE; }
  SPG(S)->gstype = gstype;
  return SUN_SUCCESS;
}

SUNErrCode SUNLinSol_SPGMRSetMaxRestarts(SUNLinearSolver S, int maxrs)
{
  if (maxrs < 0) { maxrs = SUNSPGMR_MAXRS_DEFAULT; }
  SPG(S)->max_restarts = maxrs;
  return SUN_SUCCESS;
}

SUNErrCode SUNLinSol_SPGMRSetClassicalGSFn(SUNLinearSolver S, SUNAMD_ClassicalGSFn fn)
{
  SPG(S)->cgs = fn ? fn : SUNClassicalGS;
  return SUN_SUCCESS;
}
