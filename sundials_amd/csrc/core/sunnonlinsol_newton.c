/* -----------------------------------------------------------------------------
 * Modified Newton nonlinear solver (SUNNONLINEARSOLVER_ROOTFIND).
 *
 * Restates src/sunnonlinsol/newton/sunnonlinsol_newton.c:
 *   constructor / defaults (maxiters = 3)        :44-91
 *   initialize (resets counters and jcur)        :133-146
 *   solve loop                                    :179-324
 *     - Sys -> [LSetup] -> { delta = -delta; LSolve; ycor += delta; CTest;
 *       Sys } until converged / diverged / maxiters
 *     - on a recoverable failure with a stale Jacobian: zero ycor, call
 *       LSetup with jbad = TRUE and retry (:296-309)
 * The per-iteration vector work is N_VScale(-1) + N_VLinearSum(1,1), both of
 * which run on whatever N_Vector the integrator hands in.
 * ---------------------------------------------------------------------------*/
#include <stdlib.h>
#include <string.h>

#include "sunamd/sundials_core.h"

typedef struct
{
  SUNNonlinSolSysFn Sys;
  SUNNonlinSolLSetupFn LSetup;
  SUNNonlinSolLSolveFn LSolve;
  SUNNonlinSolConvTestFn CTest;
  N_Vector delta;
  sunbooleantype jcur;
  int curiter;
  int maxiters;
  long int niters;
  long int nconvfails;
  void* ctest_data;
  sunbooleantype negated_lsolve; /* LSolve returns the solution for -b */
} NewtonContent;

#define NC(NLS) ((NewtonContent*)((NLS)->content))

static SUNNonlinearSolver_Type newton_gettype(SUNNonlinearSolver NLS)
{
  (void)NLS;
  return SUNNONLINEARSOLVER_ROOTFIND;
}

static SUNErrCode newton_initialize(SUNNonlinearSolver NLS)
{
  NewtonContent* c = NC(NLS);
  if (!c->Sys || !c->CTest || !c->LSolve) { return SUN_ERR_ARG_CORRUPT; }
  c->niters     = 0;
  c->nconvfails = 0;
  c->jcur       = SUNFALSE;
  return SUN_SUCCESS;
}

static int newton_solve(SUNNonlinearSolver NLS, N_Vector y0, N_Vector ycor,
                        N_Vector w, sunrealtype tol, sunbooleantype callLSetup,
                        void* mem)
{
  (void)y0;
  NewtonContent* c = NC(NLS);
  int retval;
  sunbooleantype jbad = SUNFALSE;
  N_Vector delta      = c->delta;

  if (!c->Sys || !c->CTest || !c->LSolve) { return SUN_ERR_ARG_CORRUPT; }
  if (callLSetup && !c->LSetup) { return SUN_ERR_ARG_CORRUPT; }

  c->niters     = 0;
  c->nconvfails = 0;

  for (;;)
  {
    c->curiter = 0;

    /* failures of the first residual or of the setup end the solve */
    retval = c->Sys(ycor, delta, mem);
    if (retval != SUN_SUCCESS) { break; }

    if (callLSetup)
    {
      retval = c->LSetup(jbad, &c->jcur, mem);
      if (retval != SUN_SUCCESS) { break; }
    }

    for (;;)
    {
      c->niters++;
      if (!c->negated_lsolve) { N_VScale(SUN_RCONST(-1.0), delta, delta); }

      retval = c->LSolve(delta, mem);
      if (retval != SUN_SUCCESS) { break; }

      N_VLinearSum(SUN_RCONST(1.0), ycor, SUN_RCONST(1.0), delta, ycor);

      retval = c->CTest(NLS, ycor, delta, tol, w, c->ctest_data);
      c->curiter++;

      if (retval == SUN_SUCCESS)
      {
        c->jcur = SUNFALSE;
        return SUN_SUCCESS;
      }
      if (retval != SUN_NLS_CONTINUE) { break; }
      if (c->curiter >= c->maxiters)
      {
        retval = SUN_NLS_CONV_RECVR;
        break;
      }
      retval = c->Sys(ycor, delta, mem);
      if (retval != SUN_SUCCESS) { break; }
    }

    /* recoverable failure with stale Jacobian data: re-setup and retry */
    if ((retval > 0) && !(c->jcur) && (c->LSetup))
    {
      c->nconvfails++;
      callLSetup = SUNTRUE;
      jbad       = SUNTRUE;
      N_VConst(SUN_RCONST(0.0), ycor);
      continue;
    }
    break;
  }

  c->nconvfails++;
  return retval;
}

/* No counterpart in the reference.  With on, the solve loop hands LSolve the
   residual itself instead of its negation (the N_VScale(-1) of
   sunnonlinsol_newton.c:239 is skipped): for an integrator whose LSolve solves
   J x = -b, the negation folded into the solve.  The solution of -b is the
   negated solution of b, bit for bit, so the iterates are unchanged. */
SUNErrCode SUNNonlinSol_NewtonSetNegatedLSolve(SUNNonlinearSolver NLS, sunbooleantype on)
{
  if (!NLS || !NLS->content || NLS->ops->solve != newton_solve) { return SUN_ERR_ARG_INCOMPATIBLE; }
  NC(NLS)->negated_lsolve = on ? SUNTRUE : SUNFALSE;
  return SUN_SUCCESS;
}

static SUNErrCode newton_free(SUNNonlinearSolver NLS)
{
  if (NLS == NULL) { return SUN_SUCCESS; }
  if (NLS->content)
  {
    NewtonContent* c = NC(NLS);
    if (c->delta) { N_VDestroy(c->delta); }
    free(NLS->content);
    NLS->content = NULL;
  }
  SUNNonlinSolFreeEmpty(NLS);
  return SUN_SUCCESS;
}

static SUNErrCode newton_setsys(SUNNonlinearSolver NLS, SUNNonlinSolSysFn f)
{
  if (!f) { return SUN_ERR_ARG_CORRUPT; }
  NC(NLS)->Sys = f;
  return SUN_SUCCESS;
}
static SUNErrCode newton_setlsetup(SUNNonlinearSolver NLS, SUNNonlinSolLSetupFn f)
{
  NC(NLS)->LSetup = f;
  return SUN_SUCCESS;
}
static SUNErrCode newton_setlsolve(SUNNonlinearSolver NLS, SUNNonlinSolLSolveFn f)
{
  if (!f) { return SUN_ERR_ARG_CORRUPT; }
  NC(NLS)->LSolve = f;
  return SUN_SUCCESS;
}
static SUNErrCode newton_setctest(SUNNonlinearSolver NLS, SUNNonlinSolConvTestFn f,
                                  void* ctest_data)
{
  if (!f) { return SUN_ERR_ARG_CORRUPT; }
  NC(NLS)->CTest      = f;
  NC(NLS)->ctest_data = ctest_data;
  return SUN_SUCCESS;
}
static SUNErrCode newton_setmaxiters(SUNNonlinearSolver NLS, int maxiters)
{
  if (maxiters < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  NC(NLS)->maxiters = maxiters;
  return SUN_SUCCESS;
}
static SUNErrCode newton_getnumiters(SUNNonlinearSolver NLS, long int* n)
{
  *n = NC(NLS)->niters;
  return SUN_SUCCESS;
}
static SUNErrCode newton_getcuriter(SUNNonlinearSolver NLS, int* iter)
{
  *iter = NC(NLS)->curiter;
  return SUN_SUCCESS;
}
static SUNErrCode newton_getnumconvfails(SUNNonlinearSolver NLS, long int* n)
{
  *n = NC(NLS)->nconvfails;
  return SUN_SUCCESS;
}

SUNNonlinearSolver SUNNonlinSol_Newton(N_Vector y, SUNContext sunctx)
{
  if (!y || !y->ops->nvclone || !y->ops->nvdestroy || !y->ops->nvscale ||
      !y->ops->nvlinearsum)
  {
    SUNAMD_SetLastError(sunctx, SUN_ERR_ARG_INCOMPATIBLE);
    return NULL;
  }
  SUNNonlinearSolver NLS = SUNNonlinSolNewEmpty(sunctx);
  if (!NLS) { return NULL; }
  NLS->ops->gettype         = newton_gettype;
  NLS->ops->initialize      = newton_initialize;
  NLS->ops->solve           = newton_solve;
  NLS->ops->free            = newton_free;
  NLS->ops->setsysfn        = newton_setsys;
  NLS->ops->setlsetupfn     = newton_setlsetup;
  NLS->ops->setlsolvefn     = newton_setlsolve;
  NLS->ops->setctestfn      = newton_setctest;
  NLS->ops->setmaxiters     = newton_setmaxiters;
  NLS->ops->getnumiters     = newton_getnumiters;
  NLS->ops->getcuriter      = newton_getcuriter;
  NLS->ops->getnumconvfails = newton_getnumconvfails;

  NewtonContent* c = (NewtonContent*)calloc(1, sizeof(NewtonContent));
  if (!c)
  {
    SUNNonlinSolFreeEmpty(NLS);
    return NULL;
  }
  c->maxiters  = 3;
  c->delta     = N_VClone(y);
  NLS->content = c;
  if (!c->delta)
  {
    newton_free(NLS);
    return NULL;
  }
  return NLS;
}
