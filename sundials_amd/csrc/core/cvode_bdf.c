/* -----------------------------------------------------------------------------
 * Variable-order, variable-step BDF integrator in Nordsieck form with modified
 * Newton iteration and a matrix-based linear-solver interface.
 *
 * This is the host-side caller of the vector / block-matrix hot path.  Every
 * decision that sets the step, order and iteration counters follows the
 * reference CVODE so that a run on any conforming N_Vector / SUNMatrix /
 * SUNLinearSolver takes the same steps:
 *   defaults / constants          src/cvode/cvode_impl.h:53-127, cvode.c:42-125,
 *                                 cvode_ls_impl.h:43-45, cvode_nls.c:32-34
 *   CVodeCreate / CVodeInit       src/cvode/cvode.c:255-540
 *   CVode main loop               src/cvode/cvode.c:1037-1670
 *   CVodeGetDky                   src/cvode/cvode.c:1671-1753
 *   cvInitialSetup                src/cvode/cvode.c:2018-2110
 *   cvHin/cvUpperBoundH0/cvYddNorm src/cvode/cvode.c:2161-2340
 *   cvStep                        src/cvode/cvode.c:2367-2487
 *   cvAdjustParams/Order, cvIncreaseBDF, cvDecreaseBDF, cvRescale, cvPredict
 *                                 src/cvode/cvode.c:2505-2771
 *   cvSet, cvSetBDF, cvSetTqBDF   src/cvode/cvode.c:2791-3053
 *   cvNls                         src/cvode/cvode.c:3068-3148
 *   cvHandleNFlag, cvRestore, cvDoErrorTest
 *                                 src/cvode/cvode.c:3257-3447
 *   cvCompleteStep, cvPrepareNextStep, cvSetEta, cvComputeEtaqm1/qp1,
 *   cvChooseEta                   src/cvode/cvode.c:3467-3711
 *   cvEwtSetSS/SV                 src/cvode/cvode.c:4791-4866
 *   cvNlsResidual/ConvTest/LSetup/LSolve   src/cvode/cvode_nls.c:256-399
 *   cvLsLinSys/cvLsSetup/cvLsSolve src/cvode/cvode_ls.c:1311-1850
 * Not restated: Adams, rootfinding, projection, constraints, stability limit
 * detection, difference-quotient Jacobians, logging and monitoring.
 * ---------------------------------------------------------------------------*/
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sunamd/cvode_sunamd.h"
#include "sunamd_pow.h"

/* ---- constants (values from cvode_impl.h, cvode.c, cvode_ls_impl.h) ------ */
#define BDF_Q_MAX 5
#define L_MAX_BDF (BDF_Q_MAX + 1)
#define NUM_TESTS 5

#define HMIN_DEFAULT     0.0
#define HMAX_INV_DEFAULT 0.0
#define MXHNIL_DEFAULT   10
#define MXSTEP_DEFAULT   500
#define MSBP_DEFAULT     20
#define DGMAX_LSETUP_DEFAULT 0.3

#define ETA_MIN_FX_DEFAULT 0.0
#define ETA_MAX_FX_DEFAULT 1.5
#define ETA_MAX_FS_DEFAULT 10000.0
#define ETA_MAX_ES_DEFAULT 10.0
#define ETA_MAX_GS_DEFAULT 10.0
#define ETA_MIN_DEFAULT    0.1
#define ETA_MAX_EF_DEFAULT 0.2
#define ETA_MIN_EF_DEFAULT 0.1
#define ETA_CF_DEFAULT     0.25
#define SMALL_NST_DEFAULT  10
#define SMALL_NEF_DEFAULT  2
#define ONEPSM             1.000001

#define ADDON 0.000001
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define LONG_WAIT 10

#define MXNCF  10
#define MXNEF  7
#define MXNEF1 3

#define DO_ERROR_TEST  +2
#define PREDICT_AGAIN  +3
#define TRY_AGAIN      +5
#define FIRST_CALL     +6
#define PREV_CONV_FAIL +7
#define PREV_ERR_FAIL  +9
#define RHSFUNC_RECVR  +10

#define CV_NO_FAILURES 0
#define CV_FAIL_BAD_J  1
#define CV_FAIL_OTHER  2

#define CV_NN 0
#define CV_SS 1
#define CV_SV 2
#define CV_WF 3

#define FUZZ_FACTOR 100.0
#define HLB_FACTOR  100.0
#define HUB_FACTOR  0.1
#define H_BIAS      0.5
#define MAX_ITERS   4
#define CORTES      0.1

#define NLS_MAXCOR 3
#define CRDOWN     0.3
#define RDIV       2.0

#define CVLS_MSBJ  51
#define CVLS_DGMAX 0.2

#define ZERO   0.0
#define PT1    0.1
#define POINT2 0.2
#define HALF   0.5
#define ONE    1.0
#define TWO    2.0
#define FOUR   4.0

#define SUNMAX(a, b) ((a) > (b) ? (a) : (b))
#define SUNMIN(a, b) ((a) < (b) ? (a) : (b))
#define SUNRabs(x)   fabs(x)
#define SUNRsqrt(x)  ((x) <= 0.0 ? 0.0 : sqrt((x)))
/* SUNRpowerR is pow() in the reference (sundials_math.h:235); the correctly
   rounded sunamd_pow gives the same result wherever libm's pow is correctly
   rounded, and the same result on host and device everywhere.  A build with
   SUNAMD_SUNRPOWER_LIBM defined uses the C library's pow instead, as the
   reference does (used only to measure what the substitution changes). */
#ifdef SUNAMD_SUNRPOWER_LIBM
#define SUNRpowerR(b, e) pow((b), (e))
#else
#define SUNRpowerR(b, e) sunamd_pow((b), (e))
#endif

/* ---- integrator state ---------------------------------------------------- */
typedef struct
{
  /* linear-solver interface state (cvode_ls_impl.h CVLsMemRec subset) */
  SUNLinearSolver LS;
  SUNMatrix A;
  SUNMatrix savedJ;
  CVLsJacFn jac;
  sunbooleantype iterative, matrixbased, scalesol;
  sunbooleantype jbad;
  long int msbj, nstlj, nje, nli, ncfl;
  sunrealtype dgmax_jbad, tnlj;
  N_Vector x; /* solution vector for the LS */
  int last_flag;
} LsState;

typedef struct CVodeMemRec
{
  SUNContext sunctx;
  sunrealtype uround;

  CVRhsFn f;
  void* user_data;
  int lmm;
  int itol;
  sunrealtype reltol, Sabstol;
  N_Vector Vabstol;
  sunbooleantype atolmin0;
  sunbooleantype user_efun;
  CVEwtFn efun_user;

  /* Nordsieck history and work vectors */
  N_Vector zn[L_MAX_BDF];
  N_Vector ewt, y, acor, tempv, ftemp, vtemp1, vtemp2, vtemp3;
  int qmax;
  long int mxstep;
  int mxhnil;
  sunrealtype hin, hmin, hmax_inv;
  sunrealtype eta_min_fx, eta_max_fx, eta_max_fs, eta_max_es, eta_max_gs;
  sunrealtype eta_min, eta_min_ef, eta_max_ef, eta_cf;
  long int small_nst;
  int small_nef;
  int maxnef, maxncf;
  sunrealtype nlscoef;
  long int msbp;
  sunrealtype dgmax_lsetup;
  sunbooleantype tstopset;
  sunrealtype tstop;

  /* step data */
  int q, qprime, next_q, qwait, L;
  sunrealtype hscale, h, hprime, next_h, eta, tn, tretlast;
  sunrealtype tau[L_MAX_BDF + 1];
  sunrealtype tq[NUM_TESTS + 1];
  sunrealtype l[L_MAX_BDF];
  sunrealtype rl1, gamma, gammap, gamrat;
  sunrealtype crate, delp, acnrm;
  sunbooleantype acnrmcur;
  sunrealtype etaqm1, etaq, etaqp1, etamax;
  sunrealtype saved_tq5;
  int indx_acor;
  sunbooleantype jcur;
  sunrealtype tolsf;
  int convfail;
  sunrealtype cvals[L_MAX_BDF + 2];
  N_Vector Xvecs[L_MAX_BDF + 2];

  /* counters */
  long int nst, nfe, ncfn, netf, nni, nnf, nsetups, nhnil, nstlp, nscon;
  sunrealtype h0u, hu;
  int qu;
  long int ninit_used;
  sunbooleantype nonzero_start; /* this step's Newton iteration started from a proposal */

  /* nonlinear solver */
  SUNNonlinearSolver NLS;
  sunbooleantype ownNLS;

  /* linear solver */
  sunbooleantype have_lsetup, have_lsolve;
  LsState ls;

  CVInitialCorrectionFn initcor;

  /* fused error-weight / residual kernels of the vector (cvode_io.c:1006-1031) */
  sunbooleantype usefused;
  int (*fused_ewt_ss)(sunbooleantype, sunrealtype, sunrealtype, N_Vector, N_Vector, N_Vector);
  int (*fused_ewt_sv)(sunbooleantype, sunrealtype, N_Vector, N_Vector, N_Vector, N_Vector);
  int (*fused_resid)(sunrealtype, sunrealtype, N_Vector, N_Vector, N_Vector, N_Vector);
  int (*fused_predict)(int, N_Vector*); /* Nordsieck predictor / its undo in one pass */
  int (*fused_restore)(int, N_Vector*);
  /* copy + ScaleAddI + factorisation of the Newton matrix in one pass (NULL:
     the library has none); declines with SUN_ERR_ARG_INCOMPATIBLE */
  int (*fused_lsetup)(SUNLinearSolver, SUNMatrix, SUNMatrix, sunrealtype, int);
  int (*fused_lsolve)(SUNLinearSolver, SUNMatrix, N_Vector, sunrealtype);
  sunbooleantype nls_negated; /* the Newton solver hands LSolve r, not -r */

  sunbooleantype MallocDone;
} CVodeMemRec;

typedef CVodeMemRec* CVodeMem;

static int cvEwtSet(N_Vector ycur, N_Vector weight, CVodeMem cv);
static int cvNlsResidual(N_Vector ycor, N_Vector res, void* mem);
static int cvNlsConvTest(SUNNonlinearSolver NLS, N_Vector ycor, N_Vector delta,
                         sunrealtype tol, N_Vector ewt, void* mem);
static int cvNlsLSetup(sunbooleantype jbad, sunbooleantype* jcur, void* mem);
static int cvNlsLSolve(N_Vector delta, void* mem);

/* ------------------------------------------------------------------------- */
/* creation / initialisation                                                 */
/* ------------------------------------------------------------------------- */

void* CVodeCreate(int lmm, SUNContext sunctx)
{
  if (lmm != CV_BDF || sunctx == NULL) { return NULL; }
  CVodeMem cv = (CVodeMem)calloc(1, sizeof(CVodeMemRec));
  if (!cv) { return NULL; }
  cv->sunctx       = sunctx;
  cv->lmm          = lmm;
  cv->uround       = SUN_UNIT_ROUNDOFF;
  cv->itol         = CV_NN;
  cv->atolmin0     = SUNTRUE;
  cv->qmax         = BDF_Q_MAX;
  cv->mxstep       = MXSTEP_DEFAULT;
  cv->mxhnil       = MXHNIL_DEFAULT;
  cv->hin          = ZERO;
  cv->hmin         = HMIN_DEFAULT;
  cv->hmax_inv     = HMAX_INV_DEFAULT;
  cv->eta_min_fx   = ETA_MIN_FX_DEFAULT;
  cv->eta_max_fx   = ETA_MAX_FX_DEFAULT;
  cv->eta_max_fs   = ETA_MAX_FS_DEFAULT;
  cv->eta_max_es   = ETA_MAX_ES_DEFAULT;
  cv->eta_max_gs   = ETA_MAX_GS_DEFAULT;
  cv->eta_min      = ETA_MIN_DEFAULT;
  cv->eta_min_ef   = ETA_MIN_EF_DEFAULT;
  cv->eta_max_ef   = ETA_MAX_EF_DEFAULT;
  cv->eta_cf       = ETA_CF_DEFAULT;
  cv->small_nst    = SMALL_NST_DEFAULT;
  cv->small_nef    = SMALL_NEF_DEFAULT;
  cv->maxnef       = MXNEF;
  cv->maxncf       = MXNCF;
  cv->nlscoef      = CORTES;
  cv->msbp         = MSBP_DEFAULT;
  cv->dgmax_lsetup = DGMAX_LSETUP_DEFAULT;
  cv->convfail     = CV_NO_FAILURES;
  return cv;
}

static void cvFreeVectors(CVodeMem cv)
{
  N_Vector* vs[] = {&cv->ewt, &cv->acor, &cv->tempv, &cv->ftemp,
                    &cv->vtemp1, &cv->vtemp2, &cv->vtemp3, &cv->Vabstol};
  for (size_t i = 0; i < sizeof(vs) / sizeof(vs[0]); i++)
  {
    if (*vs[i]) { N_VDestroy(*vs[i]); }
    *vs[i] = NULL;
  }
  for (int j = 0; j <= cv->qmax; j++)
  {
    if (cv->zn[j]) { N_VDestroy(cv->zn[j]); }
    cv->zn[j] = NULL;
  }
}

static int cvAllocVectors(CVodeMem cv, N_Vector tmpl)
{
  N_Vector* vs[] = {&cv->ewt, &cv->acor, &cv->tempv, &cv->ftemp,
                    &cv->vtemp1, &cv->vtemp2, &cv->vtemp3};
  for (size_t i = 0; i < sizeof(vs) / sizeof(vs[0]); i++)
  {
    *vs[i] = N_VClone(tmpl);
    if (!*vs[i])
    {
      cvFreeVectors(cv);
      return 0;
    }
  }
  for (int j = 0; j <= cv->qmax; j++)
  {
    cv->zn[j] = N_VClone(tmpl);
    if (!cv->zn[j])
    {
      cvFreeVectors(cv);
      return 0;
    }
  }
  return 1;
}

static void cvResetCounters(CVodeMem cv, sunrealtype t0)
{
  cv->tn      = t0;
  cv->q       = 1;
  cv->L       = 2;
  cv->qwait   = cv->L;
  cv->etamax  = cv->eta_max_fs;
  cv->qu      = 0;
  cv->hu      = ZERO;
  cv->tolsf   = ONE;
  cv->nst     = 0;
  cv->nfe     = 0;
  cv->ncfn    = 0;
  cv->netf    = 0;
  cv->nni     = 0;
  cv->nnf     = 0;
  cv->nsetups = 0;
  cv->nhnil   = 0;
  cv->nstlp   = 0;
  cv->nscon   = 0;
  cv->h0u     = ZERO;
  cv->next_h  = ZERO;
  cv->next_q  = 0;
  cv->ninit_used = 0;
  cv->saved_tq5  = ZERO;
  cv->indx_acor  = 0;
}

int CVodeSetNonlinearSolver(void* cvode_mem, SUNNonlinearSolver NLS)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (!NLS || !NLS->ops || !NLS->ops->gettype || !NLS->ops->solve ||
      !NLS->ops->setsysfn)
  {
    return CV_ILL_INPUT;
  }
  if (cv->NLS && cv->ownNLS) { SUNNonlinSolFree(cv->NLS); }
  cv->NLS    = NLS;
  cv->ownNLS = SUNFALSE;
  cv->nls_negated = SUNFALSE; /* the new solver has not been told */
  if (SUNNonlinSolGetType(NLS) != SUNNONLINEARSOLVER_ROOTFIND) { return CV_ILL_INPUT; }
  if (SUNNonlinSolSetSysFn(NLS, cvNlsResidual) != SUN_SUCCESS) { return CV_ILL_INPUT; }
  if (SUNNonlinSolSetConvTestFn(NLS, cvNlsConvTest, cv) != SUN_SUCCESS)
  {
    return CV_ILL_INPUT;
  }
  if (SUNNonlinSolSetMaxIters(NLS, NLS_MAXCOR) != SUN_SUCCESS) { return CV_ILL_INPUT; }
  cv->acnrmcur = SUNFALSE;
  return CV_SUCCESS;
}

int CVodeInit(void* cvode_mem, CVRhsFn f, sunrealtype t0, N_Vector y0)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (!y0 || !f) { return CV_ILL_INPUT; }
  if (!y0->ops->nvclone || !y0->ops->nvdestroy || !y0->ops->nvlinearsum ||
      !y0->ops->nvconst || !y0->ops->nvprod || !y0->ops->nvdiv ||
      !y0->ops->nvscale || !y0->ops->nvabs || !y0->ops->nvinv ||
      !y0->ops->nvaddconst || !y0->ops->nvmaxnorm || !y0->ops->nvwrmsnorm)
  {
    return CV_ILL_INPUT; /* required ops, cvode.c:1838-1850 */
  }
  if (!cvAllocVectors(cv, y0)) { return CV_MEM_FAIL; }
  cv->f  = f;
  cv->tn = t0;
  N_VScale(ONE, y0, cv->zn[0]);

  SUNNonlinearSolver NLS = SUNNonlinSol_Newton(y0, cv->sunctx);
  if (!NLS)
  {
    cvFreeVectors(cv);
    return CV_MEM_FAIL;
  }
  if (CVodeSetNonlinearSolver(cv, NLS) != CV_SUCCESS)
  {
    cvFreeVectors(cv);
    SUNNonlinSolFree(NLS);
    return CV_MEM_FAIL;
  }
  cv->ownNLS = SUNTRUE;
  cvResetCounters(cv, t0);
  cv->MallocDone = SUNTRUE;
  return CV_SUCCESS;
}

int CVodeReInit(void* cvode_mem, sunrealtype t0, N_Vector y0)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (!cv->MallocDone) { return CV_NO_MALLOC; }
  if (!y0) { return CV_ILL_INPUT; }
  N_VScale(ONE, y0, cv->zn[0]);
  cvResetCounters(cv, t0);
  return CV_SUCCESS;
}

int CVodeSStolerances(void* cvode_mem, sunrealtype reltol, sunrealtype abstol)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (!cv->MallocDone) { return CV_NO_MALLOC; }
  if (reltol < ZERO || abstol < ZERO) { return CV_ILL_INPUT; }
  cv->reltol    = reltol;
  cv->Sabstol   = abstol;
  cv->atolmin0  = (abstol == ZERO);
  cv->itol      = CV_SS;
  cv->user_efun = SUNFALSE;
  return CV_SUCCESS;
}

int CVodeSVtolerances(void* cvode_mem, sunrealtype reltol, N_Vector abstol)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (!cv->MallocDone) { return CV_NO_MALLOC; }
  if (reltol < ZERO || !abstol) { return CV_ILL_INPUT; }
  if (!abstol->ops->nvmin) { return CV_ILL_INPUT; }
  sunrealtype atolmin = N_VMin(abstol);
  if (atolmin < ZERO) { return CV_ILL_INPUT; }
  if (!cv->Vabstol)
  {
    cv->Vabstol = N_VClone(abstol);
    if (!cv->Vabstol) { return CV_MEM_FAIL; }
  }
  cv->reltol = reltol;
  N_VScale(ONE, abstol, cv->Vabstol);
  cv->atolmin0  = (atolmin == ZERO);
  cv->itol      = CV_SV;
  cv->user_efun = SUNFALSE;
  return CV_SUCCESS;
}

int CVodeWFtolerances(void* cvode_mem, CVEwtFn efun)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (!cv->MallocDone) { return CV_NO_MALLOC; }
  cv->itol      = CV_WF;
  cv->user_efun = SUNTRUE;
  cv->efun_user = efun;
  return CV_SUCCESS;
}

int CVodeSetUserData(void* cvode_mem, void* user_data)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  cv->user_data = user_data;
  return CV_SUCCESS;
}

int CVodeSetMaxNumSteps(void* cvode_mem, long int mxsteps)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  cv->mxstep = (mxsteps == 0) ? MXSTEP_DEFAULT : mxsteps;
  return CV_SUCCESS;
}

int CVodeSetMaxOrd(void* cvode_mem, int maxord)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (maxord <= 0 || maxord > BDF_Q_MAX) { return CV_ILL_INPUT; }
  if (cv->MallocDone && maxord > cv->qmax) { return CV_ILL_INPUT; }
  cv->qmax = maxord;
  return CV_SUCCESS;
}

int CVodeSetInitStep(void* cvode_mem, sunrealtype hin)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  cv->hin = hin;
  return CV_SUCCESS;
}

int CVodeSetMinStep(void* cvode_mem, sunrealtype hmin)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (hmin < ZERO) { return CV_ILL_INPUT; }
  if (hmin == ZERO)
  {
    cv->hmin = HMIN_DEFAULT;
    return CV_SUCCESS;
  }
  if (hmin * cv->hmax_inv > ONE) { return CV_ILL_INPUT; }
  cv->hmin = hmin;
  return CV_SUCCESS;
}

/* cvode_io.c:847: 0 restores the default (20 steps), negative is an error */
int CVodeSetLSetupFrequency(void* cvode_mem, long int msbp)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (msbp < 0) { return CV_ILL_INPUT; }
  cv->msbp = (msbp == 0) ? MSBP_DEFAULT : msbp;
  return CV_SUCCESS;
}

int CVodeSetMaxStep(void* cvode_mem, sunrealtype hmax)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (hmax < ZERO) { return CV_ILL_INPUT; }
  if (hmax == ZERO)
  {
    cv->hmax_inv = HMAX_INV_DEFAULT;
    return CV_SUCCESS;
  }
  sunrealtype hmax_inv = ONE / hmax;
  if (hmax_inv * cv->hmin > ONE) { return CV_ILL_INPUT; }
  cv->hmax_inv = hmax_inv;
  return CV_SUCCESS;
}

int CVodeSetStopTime(void* cvode_mem, sunrealtype tstop)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (cv->nst > 0 && (tstop - cv->tn) * cv->h < ZERO) { return CV_ILL_INPUT; }
  cv->tstop    = tstop;
  cv->tstopset = SUNTRUE;
  return CV_SUCCESS;
}

/* The reference enables its HIP/CUDA fused kernels only for its own GPU
   vectors (cvode_io.c:1006-1031).  Here the kernels are looked up in the
   loaded MI355X library and enabled only when the state vector is one of its
   vectors; otherwise the call fails and the unfused op sequences stay on. */
int CVodeSetUseIntegratorFusedKernels(void* cvode_mem, sunbooleantype onoff)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (!onoff)
  {
    cv->usefused = SUNFALSE;
    return CV_SUCCESS;
  }
  if (!cv->MallocDone) { return CV_ILL_INPUT; }
  int (*is_mi355)(N_Vector) = (int (*)(N_Vector))dlsym(RTLD_DEFAULT, "SUNAMD_NVectorIsMi355");
  void* ss = dlsym(RTLD_DEFAULT, "cvEwtSetSS_fused_Mi355");
  void* sv = dlsym(RTLD_DEFAULT, "cvEwtSetSV_fused_Mi355");
  void* rs = dlsym(RTLD_DEFAULT, "cvNlsResid_fused_Mi355");
  void* pr = dlsym(RTLD_DEFAULT, "cvPredict_fused_Mi355");
  void* rr = dlsym(RTLD_DEFAULT, "cvRestore_fused_Mi355");
  if (!is_mi355 || !ss || !sv || !rs || !pr || !rr || !is_mi355(cv->ewt)) { return CV_ILL_INPUT; }
  cv->fused_predict = (int (*)(int, N_Vector*))pr;
  cv->fused_restore = (int (*)(int, N_Vector*))rr;
  cv->fused_ewt_ss = (int (*)(sunbooleantype, sunrealtype, sunrealtype, N_Vector, N_Vector,
                              N_Vector))ss;
  cv->fused_ewt_sv = (int (*)(sunbooleantype, sunrealtype, N_Vector, N_Vector, N_Vector,
                              N_Vector))sv;
  cv->fused_resid  = (int (*)(sunrealtype, sunrealtype, N_Vector, N_Vector, N_Vector,
                              N_Vector))rs;
  cv->fused_lsetup = (int (*)(SUNLinearSolver, SUNMatrix, SUNMatrix, sunrealtype,
                              int))dlsym(RTLD_DEFAULT, "SUNLinSol_Mi355BlockLUSetupScaled");
  cv->fused_lsolve = (int (*)(SUNLinearSolver, SUNMatrix, N_Vector,
                              sunrealtype))dlsym(RTLD_DEFAULT, "SUNLinSol_Mi355BlockLUSolveScaled");
  cv->usefused     = SUNTRUE;
  return CV_SUCCESS;
}

int CVodeSetInitialCorrectionFn(void* cvode_mem, CVInitialCorrectionFn fn)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  cv->initcor = fn;
  return CV_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* linear solver interface (cvode_ls.c)                                      */
/* ------------------------------------------------------------------------- */

static int cvLsInitialize(CVodeMem cv);
static int cvLsSetup(CVodeMem cv, int convfail, N_Vector ypred, N_Vector fpred,
                     sunbooleantype* jcurPtr, N_Vector vtemp1, N_Vector vtemp2,
                     N_Vector vtemp3);
static int cvLsSolve(CVodeMem cv, N_Vector b, N_Vector weight, N_Vector ynow,
                     N_Vector fnow);

int CVodeSetLinearSolver(void* cvode_mem, SUNLinearSolver LS, SUNMatrix A)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CVLS_MEM_NULL; }
  if (!LS || !LS->ops->gettype || !LS->ops->solve) { return CVLS_ILL_INPUT; }
  SUNLinearSolver_Type t = SUNLinSolGetType(LS);
  LsState* ls            = &cv->ls;
  if (ls->x) { N_VDestroy(ls->x); }
  if (ls->savedJ) { SUNMatDestroy(ls->savedJ); }
  memset(ls, 0, sizeof(*ls));
  ls->iterative   = (t != SUNLINEARSOLVER_DIRECT);
  ls->matrixbased = (t != SUNLINEARSOLVER_ITERATIVE) &&
                    (t != SUNLINEARSOLVER_MATRIX_EMBEDDED);
  if (ls->iterative) { return CVLS_ILL_INPUT; /* direct solvers only */ }
  if (ls->matrixbased && A == NULL) { return CVLS_ILL_INPUT; }
  ls->LS         = LS;
  ls->A          = A;
  ls->msbj       = CVLS_MSBJ;
  ls->dgmax_jbad = CVLS_DGMAX;
  ls->scalesol   = (ls->matrixbased && cv->lmm == CV_BDF);
  ls->x          = N_VClone(cv->tempv);
  if (!ls->x) { return CVLS_MEM_FAIL; }
  cv->have_lsetup = SUNTRUE;
  cv->have_lsolve = SUNTRUE;
  return CVLS_SUCCESS;
}

int CVodeSetJacFn(void* cvode_mem, CVLsJacFn jac)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CVLS_MEM_NULL; }
  if (!cv->ls.LS) { return CVLS_LMEM_NULL; }
  cv->ls.jac = jac;
  return CVLS_SUCCESS;
}

static int cvLsInitialize(CVodeMem cv)
{
  LsState* ls = &cv->ls;
  if (!ls->LS) { return CVLS_LMEM_NULL; }
  if (ls->A != NULL)
  {
    /* no difference-quotient Jacobian: a Jacobian function is required */
    if (!ls->jac) { return CVLS_ILL_INPUT; }
    if (ls->savedJ == NULL)
    {
      ls->savedJ = SUNMatClone(ls->A);
      if (!ls->savedJ) { return CVLS_MEM_FAIL; }
    }
  }
  ls->nje = ls->nli = ls->ncfl = 0;
  ls->nstlj                    = 0;
  if (SUNLinSolGetType(ls->LS) == SUNLINEARSOLVER_MATRIX_EMBEDDED)
  {
    cv->have_lsetup = SUNFALSE;
    ls->scalesol    = SUNFALSE;
  }
  ls->last_flag = SUNLinSolInitialize(ls->LS);
  return ls->last_flag;
}

/* A = I - gamma*J, reusing the saved J when jok (cvode_ls.c:1311-1395) */
static int cvLsLinSys(CVodeMem cv, sunrealtype t, N_Vector y, N_Vector fy,
                      SUNMatrix A, sunbooleantype jok, sunbooleantype* jcur,
                      sunrealtype gamma, N_Vector vtemp1, N_Vector vtemp2,
                      N_Vector vtemp3, sunbooleantype* fused)
{
  LsState* ls = &cv->ls;
  int retval;
  /* with the fused kernels on, a solver that offers it forms and factors
     I - gamma*J in one pass from the saved (jok) or the fresh Jacobian;
     *fused receives SUNTRUE and the setup's flag lands in ls->last_flag */
  const sunbooleantype try_fused = cv->usefused && cv->fused_lsetup && fused;
  if (fused) { *fused = SUNFALSE; }
  if (jok)
  {
    *jcur = SUNFALSE;
    if (try_fused)
    {
      retval = cv->fused_lsetup(ls->LS, A, ls->savedJ, -gamma, 0);
      if (retval != SUN_ERR_ARG_INCOMPATIBLE)
      {
        *fused        = SUNTRUE;
        ls->last_flag = retval;
        return CVLS_SUCCESS;
      }
    }
    retval = SUNMatCopy(ls->savedJ, A);
    if (retval)
    {
      ls->last_flag = CVLS_SUNMAT_FAIL;
      return ls->last_flag;
    }
  }
  else
  {
    *jcur = SUNTRUE;
    if (SUNLinSolGetType(ls->LS) == SUNLINEARSOLVER_DIRECT)
    {
      retval = SUNMatZero(A);
      if (retval)
      {
        ls->last_flag = CVLS_SUNMAT_FAIL;
        return ls->last_flag;
      }
    }
    retval = ls->jac(t, y, fy, A, cv->user_data, vtemp1, vtemp2, vtemp3);
    if (retval < 0)
    {
      ls->last_flag = CVLS_JACFUNC_UNRECVR;
      return -1;
    }
    if (retval > 0)
    {
      ls->last_flag = CVLS_JACFUNC_RECVR;
      return 1;
    }
    if (try_fused)
    {
      retval = cv->fused_lsetup(ls->LS, A, ls->savedJ, -gamma, 1);
      if (retval != SUN_ERR_ARG_INCOMPATIBLE)
      {
        *fused        = SUNTRUE;
        ls->last_flag = retval;
        return CVLS_SUCCESS;
      }
    }
    retval = SUNMatCopy(A, ls->savedJ);
    if (retval)
    {
      ls->last_flag = CVLS_SUNMAT_FAIL;
      return ls->last_flag;
    }
  }
  retval = SUNMatScaleAddI(-gamma, A);
  if (retval)
  {
    ls->last_flag = CVLS_SUNMAT_FAIL;
    return ls->last_flag;
  }
  return CVLS_SUCCESS;
}

static int cvLsSetup(CVodeMem cv, int convfail, N_Vector ypred, N_Vector fpred,
                     sunbooleantype* jcurPtr, N_Vector vtemp1, N_Vector vtemp2,
                     N_Vector vtemp3)
{
  LsState* ls = &cv->ls;
  int retval;
  if (SUNLinSolGetType(ls->LS) == SUNLINEARSOLVER_MATRIX_EMBEDDED)
  {
    ls->last_flag = CVLS_SUCCESS;
    return ls->last_flag;
  }
  sunrealtype dgamma = SUNRabs((cv->gamma / cv->gammap) - ONE);
  ls->jbad = (cv->nst == 0) || (cv->nst >= ls->nstlj + ls->msbj) ||
             ((convfail == CV_FAIL_BAD_J) && (dgamma < ls->dgmax_jbad)) ||
             (convfail == CV_FAIL_OTHER);
  sunbooleantype fused = SUNFALSE;
  if (ls->A != NULL)
  {
    retval = cvLsLinSys(cv, cv->tn, ypred, fpred, ls->A, !(ls->jbad), jcurPtr,
                        cv->gamma, vtemp1, vtemp2, vtemp3, &fused);
    if (*jcurPtr)
    {
      ls->nje++;
      ls->nstlj = cv->nst;
      ls->tnlj  = cv->tn;
    }
    if (retval != CVLS_SUCCESS) { return retval; }
  }
  else { *jcurPtr = ls->jbad; }
  if (!fused) { ls->last_flag = SUNLinSolSetup(ls->LS, ls->A); }
  return ls->last_flag;
}

static int cvLsSolve(CVodeMem cv, N_Vector b, N_Vector weight, N_Vector ynow,
                     N_Vector fnow)
{
  (void)weight;
  (void)ynow;
  (void)fnow;
  LsState* ls = &cv->ls;
  int retval;
  int curiter = 0;
  (void)SUNNonlinSolGetCurIter(cv->NLS, &curiter);

  /* the block LU's one-pass solve (fused kernels on): b solved in place and
     scaled as it is written, the same bits as the calls below; declined
     (SUN_ERR_ARG_INCOMPATIBLE) where they must run */
  retval = SUN_ERR_ARG_INCOMPATIBLE;
  if (cv->usefused && cv->fused_lsolve)
  {
    const sunrealtype scale = (ls->scalesol && cv->gamrat != ONE) ? TWO / (ONE + cv->gamrat)
                                                                   : ONE;
    retval = cv->fused_lsolve(ls->LS, ls->A, b, cv->nls_negated ? -scale : scale);
  }
  if (retval == SUN_ERR_ARG_INCOMPATIBLE)
  {
    /* the Newton solver left the residual un-negated for the one-pass solve */
    if (cv->nls_negated) { N_VScale(-ONE, b, b); }
    /* direct solver: zero tolerance, zero initial guess */
    N_VConst(ZERO, ls->x);
    retval = SUNLinSolSetZeroGuess(ls->LS, SUNTRUE);
    if (retval != SUN_SUCCESS) { return -1; }
    retval = SUNLinSolSolve(ls->LS, ls->A, ls->x, b, ZERO);
    N_VScale(ONE, ls->x, b);
    if (ls->scalesol && cv->gamrat != ONE)
    {
      N_VScale(TWO / (ONE + cv->gamrat), b, b);
    }
  }
  if (retval != SUN_SUCCESS) { ls->ncfl++; }
  ls->last_flag = retval;
  switch (retval)
  {
  case SUN_SUCCESS: return 0;
  case SUNLS_RES_REDUCED: return (curiter == 0) ? 0 : 1;
  case SUNLS_CONV_FAIL:
  case SUNLS_ATIMES_FAIL_REC:
  case SUNLS_PSOLVE_FAIL_REC:
  case SUNLS_PACKAGE_FAIL_REC:
  case SUNLS_QRFACT_FAIL:
  case SUNLS_LUFACT_FAIL: return 1;
  case SUN_ERR_ARG_CORRUPT:
  case SUN_ERR_ARG_INCOMPATIBLE:
  case SUN_ERR_MEM_FAIL:
  case SUNLS_GS_FAIL:
  case SUNLS_QRSOL_FAIL:
  case SUN_ERR_EXT_FAIL:
  case SUNLS_ATIMES_FAIL_UNREC:
  case SUNLS_PSOLVE_FAIL_UNREC: return -1;
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* nonlinear solver interface (cvode_nls.c)                                  */
/* ------------------------------------------------------------------------- */

static int cvNlsInit(CVodeMem cv)
{
  if (SUNNonlinSolSetLSetupFn(cv->NLS, cv->have_lsetup ? cvNlsLSetup : NULL) !=
      SUN_SUCCESS)
  {
    return CV_NLS_INIT_FAIL;
  }
  if (SUNNonlinSolSetLSolveFn(cv->NLS, cv->have_lsolve ? cvNlsLSolve : NULL) !=
      SUN_SUCCESS)
  {
    return CV_NLS_INIT_FAIL;
  }
  if (SUNNonlinSolInitialize(cv->NLS) != SUN_SUCCESS) { return CV_NLS_INIT_FAIL; }
  return CV_SUCCESS;
}

static int cvNlsLSetup(sunbooleantype jbad, sunbooleantype* jcur, void* mem)
{
  CVodeMem cv = (CVodeMem)mem;
  if (jbad) { cv->convfail = CV_FAIL_BAD_J; }
  int retval = cvLsSetup(cv, cv->convfail, cv->y, cv->ftemp, &cv->jcur, cv->vtemp1,
                         cv->vtemp2, cv->vtemp3);
  cv->nsetups++;
  *jcur      = cv->jcur;
  cv->gamrat = ONE;
  cv->gammap = cv->gamma;
  cv->crate  = ONE;
  cv->nstlp  = cv->nst;
  if (retval < 0) { return CV_LSETUP_FAIL; }
  if (retval > 0) { return SUN_NLS_CONV_RECVR; }
  return CV_SUCCESS;
}

static int cvNlsLSolve(N_Vector delta, void* mem)
{
  CVodeMem cv = (CVodeMem)mem;
  int retval  = cvLsSolve(cv, delta, cv->ewt, cv->y, cv->ftemp);
  if (retval < 0) { return CV_LSOLVE_FAIL; }
  if (retval > 0) { return SUN_NLS_CONV_RECVR; }
  return CV_SUCCESS;
}

static int cvNlsConvTest(SUNNonlinearSolver NLS, N_Vector ycor, N_Vector delta,
                         sunrealtype tol, N_Vector ewt, void* mem)
{
  CVodeMem cv = (CVodeMem)mem;
  int m;
  sunrealtype del = N_VWrmsNorm(delta, ewt);
  if (SUNNonlinSolGetCurIter(NLS, &m) != SUN_SUCCESS) { return CV_MEM_NULL; }
  if (m > 0) { cv->crate = SUNMAX(CRDOWN * cv->crate, del / cv->delp); }
  sunrealtype dcon = del * SUNMIN(ONE, cv->crate) / tol;
  if (dcon <= ONE)
  {
    /* after one iteration from a zero start the correction is the update
       itself (the reference's shortcut); from a proposed start its norm has
       to be taken (after a restart from zero it equals del bit for bit) */
    cv->acnrm    = (m == 0 && !cv->nonzero_start) ? del : N_VWrmsNorm(ycor, ewt);
    cv->acnrmcur = SUNTRUE;
    return CV_SUCCESS;
  }
  if ((m >= 1) && (del > RDIV * cv->delp)) { return SUN_NLS_CONV_RECVR; }
  cv->delp = del;
  return SUN_NLS_CONTINUE;
}

static int cvNlsResidual(N_Vector ycor, N_Vector res, void* mem)
{
  CVodeMem cv = (CVodeMem)mem;
  N_VLinearSum(ONE, cv->zn[0], ONE, ycor, cv->y);
  int retval = cv->f(cv->tn, cv->y, cv->ftemp, cv->user_data);
  cv->nfe++;
  if (retval < 0) { return CV_RHSFUNC_FAIL; }
  if (retval > 0) { return RHSFUNC_RECVR; }
  /* cvode_nls.c:385-391; the fused kernel declines (non-zero) for a vector
     it does not handle or on a launch error: the op sequence then runs */
  if (cv->usefused && cv->fused_resid(cv->rl1, -cv->gamma, cv->zn[1], ycor, cv->ftemp, res) == 0)
  {
    return CV_SUCCESS;
  }
  N_VLinearSum(cv->rl1, cv->zn[1], ONE, ycor, res);
  N_VLinearSum(-cv->gamma, cv->ftemp, ONE, res, res);
  return CV_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* error weights (cvode.c:4791-4866)                                         */
/* ------------------------------------------------------------------------- */

static int cvEwtSet(N_Vector ycur, N_Vector weight, CVodeMem cv)
{
  if (cv->usefused && (cv->itol == CV_SS || cv->itol == CV_SV))
  {
    /* cvode.c:4794-4806, 4834-4846: weights are formed regardless of the
       component test, which then looks at tempv; a declined fused call
       (non-zero) falls back to the op sequence below */
    const int rc = (cv->itol == CV_SS)
                     ? cv->fused_ewt_ss(cv->atolmin0, cv->reltol, cv->Sabstol, ycur, cv->tempv,
                                        weight)
                     : cv->fused_ewt_sv(cv->atolmin0, cv->reltol, cv->Vabstol, ycur, cv->tempv,
                                        weight);
    if (rc == 0)
    {
      if (cv->atolmin0 && N_VMin(cv->tempv) <= ZERO) { return -1; }
      return 0;
    }
  }
  switch (cv->itol)
  {
  case CV_SS:
    N_VAbs(ycur, cv->tempv);
    N_VScale(cv->reltol, cv->tempv, cv->tempv);
    N_VAddConst(cv->tempv, cv->Sabstol, cv->tempv);
    if (cv->atolmin0 && N_VMin(cv->tempv) <= ZERO) { return -1; }
    N_VInv(cv->tempv, weight);
    return 0;
  case CV_SV:
    N_VAbs(ycur, cv->tempv);
    N_VLinearSum(cv->reltol, cv->tempv, ONE, cv->Vabstol, cv->tempv);
    if (cv->atolmin0 && N_VMin(cv->tempv) <= ZERO) { return -1; }
    N_VInv(cv->tempv, weight);
    return 0;
  case CV_WF: return cv->efun_user(ycur, weight, cv->user_data);
  }
  return -1;
}

/* ------------------------------------------------------------------------- */
/* step machinery                                                            */
/* ------------------------------------------------------------------------- */

static int cvInitialSetup(CVodeMem cv)
{
  if (cv->itol == CV_NN) { return CV_ILL_INPUT; }
  if (!cv->user_efun && cv->atolmin0 && !cv->tempv->ops->nvmin) { return CV_ILL_INPUT; }
  if (cvEwtSet(cv->zn[0], cv->ewt, cv) != 0) { return CV_ILL_INPUT; }
  if (cv->ls.LS)
  {
    if (cvLsInitialize(cv) != 0) { return CV_LINIT_FAIL; }
  }
  if (cvNlsInit(cv) != 0) { return CV_NLS_INIT_FAIL; }
  return CV_SUCCESS;
}

static sunrealtype cvUpperBoundH0(CVodeMem cv, sunrealtype tdist)
{
  N_Vector temp1 = cv->tempv;
  N_Vector temp2 = cv->acor;
  N_VAbs(cv->zn[0], temp2);
  cvEwtSet(cv->zn[0], temp1, cv);
  N_VInv(temp1, temp1);
  N_VLinearSum(HUB_FACTOR, temp2, ONE, temp1, temp1);
  N_VAbs(cv->zn[1], temp2);
  N_VDiv(temp2, temp1, temp1);
  sunrealtype hub_inv = N_VMaxNorm(temp1);
  sunrealtype hub     = HUB_FACTOR * tdist;
  if (hub * hub_inv > ONE) { hub = ONE / hub_inv; }
  return hub;
}

static int cvYddNorm(CVodeMem cv, sunrealtype hg, sunrealtype* yddnrm)
{
  N_VLinearSum(hg, cv->zn[1], ONE, cv->zn[0], cv->y);
  int retval = cv->f(cv->tn + hg, cv->y, cv->tempv, cv->user_data);
  cv->nfe++;
  if (retval < 0) { return CV_RHSFUNC_FAIL; }
  if (retval > 0) { return RHSFUNC_RECVR; }
  N_VLinearSum(ONE / hg, cv->tempv, -ONE / hg, cv->zn[1], cv->tempv);
  *yddnrm = N_VWrmsNorm(cv->tempv, cv->ewt);
  return CV_SUCCESS;
}

static int cvHin(CVodeMem cv, sunrealtype tout)
{
  int sign, count1, count2, retval;
  sunrealtype tdiff, tdist, tround, hlb, hub, hg, hgs, hs, hnew = 0, hrat, h0, yddnrm;
  sunbooleantype hgOK;

  if ((tdiff = tout - cv->tn) == ZERO) { return CV_TOO_CLOSE; }
  sign   = (tdiff > ZERO) ? 1 : -1;
  tdist  = SUNRabs(tdiff);
  tround = cv->uround * SUNMAX(SUNRabs(cv->tn), SUNRabs(tout));
  if (tdist < TWO * tround) { return CV_TOO_CLOSE; }

  hlb = HLB_FACTOR * tround;
  hub = cvUpperBoundH0(cv, tdist);
  hg  = SUNRsqrt(hlb * hub);
  if (hub < hlb)
  {
    cv->h = (sign == -1) ? -hg : hg;
    return CV_SUCCESS;
  }

  hs = hg;
  for (count1 = 1; count1 <= MAX_ITERS; count1++)
  {
    hgOK = SUNFALSE;
    for (count2 = 1; count2 <= MAX_ITERS; count2++)
    {
      hgs    = hg * sign;
      retval = cvYddNorm(cv, hgs, &yddnrm);
      if (retval < 0) { return CV_RHSFUNC_FAIL; }
      if (retval == CV_SUCCESS)
      {
        hgOK = SUNTRUE;
        break;
      }
      hg *= POINT2;
    }
    if (!hgOK)
    {
      if (count1 <= 2) { return CV_REPTD_RHSFUNC_ERR; }
      hnew = hs;
      break;
    }
    hs   = hg;
    hnew = (yddnrm * hub * hub > TWO) ? SUNRsqrt(TWO / yddnrm) : SUNRsqrt(hg * hub);
    if (count1 == MAX_ITERS) { break; }
    hrat = hnew / hg;
    if ((hrat > HALF) && (hrat < TWO)) { break; }
    if ((count1 > 1) && (hrat > TWO))
    {
      hnew = hg;
      break;
    }
    hg = hnew;
  }

  h0 = H_BIAS * hnew;
  if (h0 < hlb) { h0 = hlb; }
  if (h0 > hub) { h0 = hub; }
  if (sign == -1) { h0 = -h0; }
  cv->h = h0;
  return CV_SUCCESS;
}

static void cvRescale(CVodeMem cv)
{
  cv->cvals[0] = cv->eta;
  for (int j = 1; j <= cv->q; j++) { cv->cvals[j] = cv->eta * cv->cvals[j - 1]; }
  (void)N_VScaleVectorArray(cv->q, cv->cvals, cv->zn + 1, cv->zn + 1);
  cv->h      = cv->hscale * cv->eta;
  cv->next_h = cv->h;
  cv->hscale = cv->h;
  cv->nscon  = 0;
}

static void cvIncreaseBDF(CVodeMem cv)
{
  sunrealtype alpha0, alpha1, prod, xi, xiold, hsum, A1;
  for (int i = 0; i <= cv->qmax; i++) { cv->l[i] = ZERO; }
  cv->l[2] = alpha1 = prod = xiold = ONE;
  alpha0                           = -ONE;
  hsum                             = cv->hscale;
  if (cv->q > 1)
  {
    for (int j = 1; j < cv->q; j++)
    {
      hsum += cv->tau[j + 1];
      xi = hsum / cv->hscale;
      prod *= xi;
      alpha0 -= ONE / (j + 1);
      alpha1 += ONE / xi;
      for (int i = j + 2; i >= 2; i--) { cv->l[i] = cv->l[i] * xiold + cv->l[i - 1]; }
      xiold = xi;
    }
  }
  A1 = (-alpha0 - alpha1) / prod;
  N_VScale(A1, cv->zn[cv->indx_acor], cv->zn[cv->L]);
  if (cv->q > 1)
  {
    (void)N_VScaleAddMulti(cv->q - 1, cv->l + 2, cv->zn[cv->L], cv->zn + 2, cv->zn + 2);
  }
}

static void cvDecreaseBDF(CVodeMem cv)
{
  sunrealtype hsum, xi;
  for (int i = 0; i <= cv->qmax; i++) { cv->l[i] = ZERO; }
  cv->l[2] = ONE;
  hsum     = ZERO;
  for (int j = 1; j <= cv->q - 2; j++)
  {
    hsum += cv->tau[j];
    xi = hsum / cv->hscale;
    for (int i = j + 2; i >= 2; i--) { cv->l[i] = cv->l[i] * xi + cv->l[i - 1]; }
  }
  for (int j = 2; j < cv->q; j++) { cv->cvals[j - 2] = -cv->l[j]; }
  if (cv->q > 2)
  {
    (void)N_VScaleAddMulti(cv->q - 2, cv->cvals, cv->zn[cv->q], cv->zn + 2, cv->zn + 2);
  }
}

static void cvAdjustOrder(CVodeMem cv, int deltaq)
{
  if ((cv->q == 2) && (deltaq != 1)) { return; }
  if (deltaq == 1) { cvIncreaseBDF(cv); }
  else if (deltaq == -1) { cvDecreaseBDF(cv); }
}

static void cvAdjustParams(CVodeMem cv)
{
  if (cv->qprime != cv->q)
  {
    cvAdjustOrder(cv, cv->qprime - cv->q);
    cv->q     = cv->qprime;
    cv->L     = cv->q + 1;
    cv->qwait = cv->L;
  }
  cvRescale(cv);
}

static void cvPredict(CVodeMem cv)
{
  cv->tn += cv->h;
  if (cv->tstopset)
  {
    if ((cv->tn - cv->tstop) * cv->h > ZERO) { cv->tn = cv->tstop; }
  }
  if (cv->usefused && cv->fused_predict(cv->q, cv->zn) == 0) { return; }
  for (int k = 1; k <= cv->q; k++)
  {
    for (int j = cv->q; j >= k; j--)
    {
      N_VLinearSum(ONE, cv->zn[j - 1], ONE, cv->zn[j], cv->zn[j - 1]);
    }
  }
}

static void cvSetTqBDF(CVodeMem cv, sunrealtype hsum, sunrealtype alpha0,
                       sunrealtype alpha0_hat, sunrealtype xi_inv,
                       sunrealtype xistar_inv)
{
  sunrealtype A1, A2, A3, A4, A5, A6, C, Cpinv, Cppinv;
  A1        = ONE - alpha0_hat + alpha0;
  A2        = ONE + cv->q * A1;
  cv->tq[2] = SUNRabs(A1 / (alpha0 * A2));
  cv->tq[5] = SUNRabs(A2 * xistar_inv / (cv->l[cv->q] * xi_inv));
  if (cv->qwait == 1)
  {
    if (cv->q > 1)
    {
      C         = xistar_inv / cv->l[cv->q];
      A3        = alpha0 + ONE / cv->q;
      A4        = alpha0_hat + xi_inv;
      Cpinv     = (ONE - A4 + A3) / A3;
      cv->tq[1] = SUNRabs(C * Cpinv);
    }
    else { cv->tq[1] = ONE; }
    hsum += cv->tau[cv->q];
    xi_inv    = cv->h / hsum;
    A5        = alpha0 - (ONE / (cv->q + 1));
    A6        = alpha0_hat - xi_inv;
    Cppinv    = (ONE - A6 + A5) / A2;
    cv->tq[3] = SUNRabs(Cppinv / (xi_inv * (cv->q + 2) * A5));
  }
  cv->tq[4] = cv->nlscoef / cv->tq[2];
}

static void cvSetBDF(CVodeMem cv)
{
  sunrealtype alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
  cv->l[0] = cv->l[1] = xi_inv = xistar_inv = ONE;
  for (int i = 2; i <= cv->q; i++) { cv->l[i] = ZERO; }
  alpha0 = alpha0_hat = -ONE;
  hsum                = cv->h;
  if (cv->q > 1)
  {
    for (int j = 2; j < cv->q; j++)
    {
      hsum += cv->tau[j - 1];
      xi_inv = cv->h / hsum;
      alpha0 -= ONE / j;
      for (int i = j; i >= 1; i--) { cv->l[i] += cv->l[i - 1] * xi_inv; }
    }
    alpha0 -= ONE / cv->q;
    xistar_inv = -cv->l[1] - alpha0;
    hsum += cv->tau[cv->q - 1];
    xi_inv     = cv->h / hsum;
    alpha0_hat = -cv->l[1] - xi_inv;
    for (int i = cv->q; i >= 1; i--) { cv->l[i] += cv->l[i - 1] * xistar_inv; }
  }
  cvSetTqBDF(cv, hsum, alpha0, alpha0_hat, xi_inv, xistar_inv);
}

static void cvSet(CVodeMem cv)
{
  cvSetBDF(cv);
  cv->rl1   = ONE / cv->l[1];
  cv->gamma = cv->h * cv->rl1;
  if (cv->nst == 0) { cv->gammap = cv->gamma; }
  cv->gamrat = (cv->nst > 0) ? cv->gamma / cv->gammap : ONE;
}

static int cvNls(CVodeMem cv, int nflag)
{
  int flag = CV_SUCCESS;
  sunbooleantype callSetup;
  long int nni_inc = 0, nnf_inc = 0;

  if (cv->have_lsetup)
  {
    cv->convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES
                                                                       : CV_FAIL_OTHER;
    callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) ||
                (cv->nst == 0) || (cv->nst >= cv->nstlp + cv->msbp) ||
                (SUNRabs(cv->gamrat - ONE) > cv->dgmax_lsetup);
  }
  else
  {
    cv->crate = ONE;
    callSetup = SUNFALSE;
  }

  /* initial guess for the correction: zero, or the caller's proposal */
  N_VConst(ZERO, cv->acor);
  cv->nonzero_start = SUNFALSE;
  if (cv->initcor)
  {
    if (cv->initcor(cv->tn, cv->h, cv->gamma, cv->zn[0], cv->acor, cv->user_data) == 0)
    {
      cv->ninit_used++;
      cv->nonzero_start = SUNTRUE;
    }
    else { N_VConst(ZERO, cv->acor); }
  }

  if (cv->NLS->ops->setup)
  {
    flag = SUNNonlinSolSetup(cv->NLS, cv->acor, cv);
    if (flag < 0) { return CV_NLS_SETUP_FAIL; }
    if (flag > 0) { return SUN_NLS_CONV_RECVR; }
  }

  /* fused mode with the one-pass solve: the residual's negation moves into
     the solve (the solution of -r is the negated solution of r, bit for bit) */
  {
    const sunbooleantype neg = cv->usefused && cv->fused_lsolve && cv->ownNLS;
    if (neg != cv->nls_negated)
    {
      /* a solver of another type negates the residual itself */
      cv->nls_negated = (SUNNonlinSol_NewtonSetNegatedLSolve(cv->NLS, neg) == SUN_SUCCESS)
                          ? neg
                          : SUNFALSE;
    }
  }
  flag = SUNNonlinSolSolve(cv->NLS, cv->zn[0], cv->acor, cv->ewt, cv->tq[4],
                           callSetup, cv);

  (void)SUNNonlinSolGetNumIters(cv->NLS, &nni_inc);
  cv->nni += nni_inc;
  (void)SUNNonlinSolGetNumConvFails(cv->NLS, &nnf_inc);
  cv->nnf += nnf_inc;

  if (flag != SUN_SUCCESS) { return flag; }

  N_VLinearSum(ONE, cv->zn[0], ONE, cv->acor, cv->y);
  if (!cv->acnrmcur) { cv->acnrm = N_VWrmsNorm(cv->acor, cv->ewt); }
  cv->jcur = SUNFALSE;
  return flag;
}

static void cvRestore(CVodeMem cv, sunrealtype saved_t)
{
  cv->tn = saved_t;
  if (cv->usefused && cv->fused_restore(cv->q, cv->zn) == 0) { return; }
  for (int k = 1; k <= cv->q; k++)
  {
    for (int j = cv->q; j >= k; j--)
    {
      N_VLinearSum(ONE, cv->zn[j - 1], -ONE, cv->zn[j], cv->zn[j - 1]);
    }
  }
}

static int cvHandleNFlag(CVodeMem cv, int* nflagPtr, sunrealtype saved_t, int* ncfPtr)
{
  int nflag = *nflagPtr;
  if (nflag == CV_SUCCESS) { return DO_ERROR_TEST; }
  cv->ncfn++;
  cvRestore(cv, saved_t);
  if (nflag < 0)
  {
    if (nflag == CV_LSETUP_FAIL) { return CV_LSETUP_FAIL; }
    if (nflag == CV_LSOLVE_FAIL) { return CV_LSOLVE_FAIL; }
    if (nflag == CV_RHSFUNC_FAIL) { return CV_RHSFUNC_FAIL; }
    return CV_NLS_FAIL;
  }
  (*ncfPtr)++;
  cv->etamax = ONE;
  if ((SUNRabs(cv->h) <= cv->hmin * ONEPSM) || (*ncfPtr == cv->maxncf))
  {
    if (nflag == SUN_NLS_CONV_RECVR) { return CV_CONV_FAILURE; }
    if (nflag == RHSFUNC_RECVR) { return CV_REPTD_RHSFUNC_ERR; }
  }
  cv->eta   = SUNMAX(cv->eta_cf, cv->hmin / SUNRabs(cv->h));
  *nflagPtr = PREV_CONV_FAIL;
  cvRescale(cv);
  return PREDICT_AGAIN;
}

static int cvDoErrorTest(CVodeMem cv, int* nflagPtr, sunrealtype saved_t,
                         int* nefPtr, sunrealtype* dsmPtr)
{
  sunrealtype dsm = cv->acnrm * cv->tq[2];
  *dsmPtr         = dsm;
  if (dsm <= ONE) { return CV_SUCCESS; }

  (*nefPtr)++;
  cv->netf++;
  *nflagPtr = PREV_ERR_FAIL;
  cvRestore(cv, saved_t);

  if ((SUNRabs(cv->h) <= cv->hmin * ONEPSM) || (*nefPtr == cv->maxnef))
  {
    return CV_ERR_FAILURE;
  }
  cv->etamax = ONE;

  if (*nefPtr <= MXNEF1)
  {
    cv->eta = ONE / (SUNRpowerR(BIAS2 * dsm, ONE / cv->L) + ADDON);
    cv->eta = SUNMAX(cv->eta_min_ef, SUNMAX(cv->eta, cv->hmin / SUNRabs(cv->h)));
    if (*nefPtr >= cv->small_nef) { cv->eta = SUNMIN(cv->eta, cv->eta_max_ef); }
    cvRescale(cv);
    return TRY_AGAIN;
  }

  if (cv->q > 1)
  {
    cv->eta = SUNMAX(cv->eta_min_ef, cv->hmin / SUNRabs(cv->h));
    cvAdjustOrder(cv, -1);
    cv->L = cv->q;
    cv->q--;
    cv->qwait = cv->L;
    cvRescale(cv);
    return TRY_AGAIN;
  }

  cv->eta = SUNMAX(cv->eta_min_ef, cv->hmin / SUNRabs(cv->h));
  cv->h *= cv->eta;
  cv->next_h = cv->h;
  cv->hscale = cv->h;
  cv->qwait  = LONG_WAIT;
  cv->nscon  = 0;
  int retval = cv->f(cv->tn, cv->zn[0], cv->tempv, cv->user_data);
  cv->nfe++;
  if (retval < 0) { return CV_RHSFUNC_FAIL; }
  if (retval > 0) { return CV_UNREC_RHSFUNC_ERR; }
  N_VScale(cv->h, cv->tempv, cv->zn[1]);
  return TRY_AGAIN;
}

static void cvCompleteStep(CVodeMem cv)
{
  cv->nst++;
  cv->nscon++;
  cv->hu = cv->h;
  cv->qu = cv->q;
  for (int i = cv->q; i >= 2; i--) { cv->tau[i] = cv->tau[i - 1]; }
  if ((cv->q == 1) && (cv->nst > 1)) { cv->tau[2] = cv->tau[1]; }
  cv->tau[1] = cv->h;
  (void)N_VScaleAddMulti(cv->q + 1, cv->l, cv->acor, cv->zn, cv->zn);
  cv->qwait--;
  if ((cv->qwait == 1) && (cv->q != cv->qmax))
  {
    N_VScale(ONE, cv->acor, cv->zn[cv->qmax]);
    cv->saved_tq5 = cv->tq[5];
    cv->indx_acor = cv->qmax;
  }
}

static void cvSetEta(CVodeMem cv)
{
  if ((cv->eta > cv->eta_min_fx) && (cv->eta < cv->eta_max_fx))
  {
    cv->eta    = ONE;
    cv->hprime = cv->h;
  }
  else
  {
    if (cv->eta >= cv->eta_max_fx)
    {
      cv->eta = SUNMIN(cv->eta, cv->etamax);
      cv->eta /= SUNMAX(ONE, SUNRabs(cv->h) * cv->hmax_inv * cv->eta);
    }
    else
    {
      cv->eta = SUNMAX(cv->eta, cv->eta_min);
      cv->eta = SUNMAX(cv->eta, cv->hmin / SUNRabs(cv->h));
    }
    cv->hprime = cv->h * cv->eta;
    if (cv->qprime < cv->q) { cv->nscon = 0; }
  }
}

static sunrealtype cvComputeEtaqm1(CVodeMem cv)
{
  cv->etaqm1 = ZERO;
  if (cv->q > 1)
  {
    sunrealtype ddn = N_VWrmsNorm(cv->zn[cv->q], cv->ewt) * cv->tq[1];
    cv->etaqm1      = ONE / (SUNRpowerR(BIAS1 * ddn, ONE / cv->q) + ADDON);
  }
  return cv->etaqm1;
}

static sunrealtype cvComputeEtaqp1(CVodeMem cv)
{
  cv->etaqp1 = ZERO;
  if (cv->q != cv->qmax)
  {
    if (cv->saved_tq5 == ZERO) { return cv->etaqp1; }
    sunrealtype cquot = (cv->tq[5] / cv->saved_tq5) *
                        SUNRpowerI(cv->h / cv->tau[2], cv->L);
    N_VLinearSum(-cquot, cv->zn[cv->qmax], ONE, cv->acor, cv->tempv);
    sunrealtype dup = N_VWrmsNorm(cv->tempv, cv->ewt) * cv->tq[3];
    cv->etaqp1      = ONE / (SUNRpowerR(BIAS3 * dup, ONE / (cv->L + 1)) + ADDON);
  }
  return cv->etaqp1;
}

static void cvChooseEta(CVodeMem cv)
{
  sunrealtype etam = SUNMAX(cv->etaqm1, SUNMAX(cv->etaq, cv->etaqp1));
  if ((etam > cv->eta_min_fx) && (etam < cv->eta_max_fx))
  {
    cv->eta    = ONE;
    cv->qprime = cv->q;
    return;
  }
  if (etam == cv->etaq)
  {
    cv->eta    = cv->etaq;
    cv->qprime = cv->q;
  }
  else if (etam == cv->etaqm1)
  {
    cv->eta    = cv->etaqm1;
    cv->qprime = cv->q - 1;
  }
  else
  {
    cv->eta    = cv->etaqp1;
    cv->qprime = cv->q + 1;
    /* Delta_n saved in zn[qmax] for the order increase */
    N_VScale(ONE, cv->acor, cv->zn[cv->qmax]);
  }
}

static void cvPrepareNextStep(CVodeMem cv, sunrealtype dsm)
{
  if (cv->etamax == ONE)
  {
    cv->qwait  = SUNMAX(cv->qwait, 2);
    cv->qprime = cv->q;
    cv->hprime = cv->h;
    cv->eta    = ONE;
    return;
  }
  cv->etaq = ONE / (SUNRpowerR(BIAS2 * dsm, ONE / cv->L) + ADDON);
  if (cv->qwait != 0)
  {
    cv->eta    = cv->etaq;
    cv->qprime = cv->q;
    cvSetEta(cv);
  }
  else
  {
    cv->qwait  = 2;
    cv->etaqm1 = cvComputeEtaqm1(cv);
    cv->etaqp1 = cvComputeEtaqp1(cv);
    cvChooseEta(cv);
    cvSetEta(cv);
  }
}

static int cvStep(CVodeMem cv)
{
  sunrealtype saved_t, dsm = ZERO;
  int ncf = 0, nef = 0, nflag, kflag, eflag;

  if ((cv->nst > 0) && (cv->hprime != cv->h)) { cvAdjustParams(cv); }

  saved_t = cv->tn;
  nflag   = FIRST_CALL;
  for (;;)
  {
    cvPredict(cv);
    cvSet(cv);
    nflag = cvNls(cv, nflag);
    kflag = cvHandleNFlag(cv, &nflag, saved_t, &ncf);
    if (kflag == PREDICT_AGAIN) { continue; }
    if (kflag != DO_ERROR_TEST) { return kflag; }
    eflag = cvDoErrorTest(cv, &nflag, saved_t, &nef, &dsm);
    if (eflag == TRY_AGAIN) { continue; }
    if (eflag != CV_SUCCESS) { return eflag; }
    break;
  }

  cvCompleteStep(cv);
  cvPrepareNextStep(cv, dsm);
  cv->etamax = (cv->nst <= cv->small_nst) ? cv->eta_max_es : cv->eta_max_gs;
  N_VScale(cv->tq[2], cv->acor, cv->acor);
  return CV_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* driver                                                                    */
/* ------------------------------------------------------------------------- */

int CVodeGetDky(void* cvode_mem, sunrealtype t, int k, N_Vector dky)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if (!dky) { return CV_BAD_DKY; }
  if ((k < 0) || (k > cv->q)) { return CV_BAD_K; }
  sunrealtype tfuzz = FUZZ_FACTOR * cv->uround * (SUNRabs(cv->tn) + SUNRabs(cv->hu));
  if (cv->hu < ZERO) { tfuzz = -tfuzz; }
  sunrealtype tp  = cv->tn - cv->hu - tfuzz;
  sunrealtype tn1 = cv->tn + tfuzz;
  if ((t - tp) * (t - tn1) > ZERO) { return CV_BAD_T; }

  int nvec      = 0;
  sunrealtype s = (t - cv->tn) / cv->h;
  for (int j = cv->q; j >= k; j--)
  {
    cv->cvals[nvec] = ONE;
    for (int i = j; i >= j - k + 1; i--) { cv->cvals[nvec] *= i; }
    for (int i = 0; i < j - k; i++) { cv->cvals[nvec] *= s; }
    cv->Xvecs[nvec] = cv->zn[j];
    nvec += 1;
  }
  if (N_VLinearCombination(nvec, cv->cvals, cv->Xvecs, dky) != SUN_SUCCESS)
  {
    return CV_VECTOROP_ERR;
  }
  if (k == 0) { return CV_SUCCESS; }
  sunrealtype r = SUNRpowerI(cv->h, -k);
  N_VScale(r, dky, dky);
  return CV_SUCCESS;
}

int CVode(void* cvode_mem, sunrealtype tout, N_Vector yout, sunrealtype* tret, int itask)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  long int nstloc;
  int kflag, istate = CV_SUCCESS, ier;
  sunrealtype troundoff, tout_hin, rh, nrm;

  if (!cv) { return CV_MEM_NULL; }
  if (!cv->MallocDone) { return CV_NO_MALLOC; }
  if ((cv->y = yout) == NULL) { return CV_ILL_INPUT; }
  if (tret == NULL) { return CV_ILL_INPUT; }
  if ((itask != CV_NORMAL) && (itask != CV_ONE_STEP)) { return CV_ILL_INPUT; }

  if (cv->nst == 0)
  {
    cv->tretlast = *tret = cv->tn;
    ier                  = cvInitialSetup(cv);
    if (ier != CV_SUCCESS) { return ier; }

    int retval = cv->f(cv->tn, cv->zn[0], cv->zn[1], cv->user_data);
    cv->nfe++;
    if (retval < 0) { return CV_RHSFUNC_FAIL; }
    if (retval > 0) { return CV_FIRST_RHSFUNC_ERR; }

    if (cv->tstopset)
    {
      if ((cv->tstop - cv->tn) * (tout - cv->tn) <= ZERO) { return CV_ILL_INPUT; }
    }

    cv->h = cv->hin;
    if ((cv->h != ZERO) && ((tout - cv->tn) * cv->h < ZERO)) { return CV_ILL_INPUT; }
    if (cv->h == ZERO)
    {
      tout_hin = tout;
      if (cv->tstopset && (tout - cv->tn) * (tout - cv->tstop) > ZERO)
      {
        tout_hin = cv->tstop;
      }
      int hflag = cvHin(cv, tout_hin);
      if (hflag != CV_SUCCESS) { return hflag; }
    }

    rh = SUNRabs(cv->h) * cv->hmax_inv;
    if (rh > ONE) { cv->h /= rh; }
    if (SUNRabs(cv->h) < cv->hmin) { cv->h *= cv->hmin / SUNRabs(cv->h); }

    if (cv->tstopset)
    {
      if ((cv->tn + cv->h - cv->tstop) * cv->h > ZERO)
      {
        cv->h = (cv->tstop - cv->tn) * (ONE - FOUR * cv->uround);
      }
    }

    cv->hscale = cv->h;
    cv->h0u    = cv->h;
    cv->hprime = cv->h;
    N_VScale(cv->h, cv->zn[1], cv->zn[1]);
  }

  if (cv->nst > 0)
  {
    troundoff = FUZZ_FACTOR * cv->uround * (SUNRabs(cv->tn) + SUNRabs(cv->h));
    if (cv->tstopset)
    {
      if (SUNRabs(cv->tn - cv->tstop) <= troundoff)
      {
        if ((tout - cv->tstop) * cv->h >= ZERO ||
            SUNRabs(tout - cv->tstop) <= troundoff)
        {
          N_VScale(ONE, cv->zn[0], yout);
          cv->tretlast = *tret = cv->tstop;
          cv->tstopset         = SUNFALSE;
          return CV_TSTOP_RETURN;
        }
      }
      else if ((cv->tn + cv->hprime - cv->tstop) * cv->h > ZERO)
      {
        cv->hprime = (cv->tstop - cv->tn) * (ONE - FOUR * cv->uround);
        cv->eta    = cv->hprime / cv->h;
      }
    }
    if ((itask == CV_NORMAL) && ((cv->tn - tout) * cv->h >= ZERO))
    {
      cv->tretlast = *tret = tout;
      ier                  = CVodeGetDky(cv, tout, 0, yout);
      if (ier != CV_SUCCESS) { return CV_ILL_INPUT; }
      return CV_SUCCESS;
    }
    if (itask == CV_ONE_STEP && SUNRabs(cv->tn - cv->tretlast) > troundoff)
    {
      cv->tretlast = *tret = cv->tn;
      N_VScale(ONE, cv->zn[0], yout);
      return CV_SUCCESS;
    }
  }

  nstloc = 0;
  for (;;)
  {
    cv->next_h = cv->h;
    cv->next_q = cv->q;

    if (cv->nst > 0)
    {
      if (cvEwtSet(cv->zn[0], cv->ewt, cv) != 0)
      {
        istate       = CV_ILL_INPUT;
        cv->tretlast = *tret = cv->tn;
        N_VScale(ONE, cv->zn[0], yout);
        break;
      }
    }

    if ((cv->mxstep > 0) && (nstloc >= cv->mxstep))
    {
      istate       = CV_TOO_MUCH_WORK;
      cv->tretlast = *tret = cv->tn;
      N_VScale(ONE, cv->zn[0], yout);
      break;
    }

    nrm       = N_VWrmsNorm(cv->zn[0], cv->ewt);
    cv->tolsf = cv->uround * nrm;
    if (cv->tolsf > ONE)
    {
      istate       = CV_TOO_MUCH_ACC;
      cv->tretlast = *tret = cv->tn;
      N_VScale(ONE, cv->zn[0], yout);
      cv->tolsf *= TWO;
      break;
    }
    else { cv->tolsf = ONE; }

    if (cv->tn + cv->h == cv->tn) { cv->nhnil++; }

    kflag = cvStep(cv);
    if (kflag != CV_SUCCESS)
    {
      istate       = kflag;
      cv->tretlast = *tret = cv->tn;
      N_VScale(ONE, cv->zn[0], yout);
      break;
    }
    nstloc++;

    if (cv->tstopset)
    {
      troundoff = FUZZ_FACTOR * cv->uround * (SUNRabs(cv->tn) + SUNRabs(cv->h));
      if (SUNRabs(cv->tn - cv->tstop) <= troundoff) { cv->tn = cv->tstop; }
    }

    if (cv->tstopset)
    {
      troundoff = FUZZ_FACTOR * cv->uround * (SUNRabs(cv->tn) + SUNRabs(cv->h));
      if (SUNRabs(cv->tn - cv->tstop) <= troundoff)
      {
        if ((tout - cv->tstop) * cv->h >= ZERO ||
            SUNRabs(tout - cv->tstop) <= troundoff)
        {
          N_VScale(ONE, cv->zn[0], yout);
          cv->tretlast = *tret = cv->tstop;
          cv->tstopset         = SUNFALSE;
          istate               = CV_TSTOP_RETURN;
          break;
        }
      }
      else if ((cv->tn + cv->hprime - cv->tstop) * cv->h > ZERO)
      {
        cv->hprime = (cv->tstop - cv->tn) * (ONE - FOUR * cv->uround);
        cv->eta    = cv->hprime / cv->h;
      }
    }

    if ((itask == CV_NORMAL) && (cv->tn - tout) * cv->h >= ZERO)
    {
      istate       = CV_SUCCESS;
      cv->tretlast = *tret = tout;
      (void)CVodeGetDky(cv, tout, 0, yout);
      cv->next_q = cv->qprime;
      cv->next_h = cv->hprime;
      break;
    }

    if (itask == CV_ONE_STEP)
    {
      istate       = CV_SUCCESS;
      cv->tretlast = *tret = cv->tn;
      N_VScale(ONE, cv->zn[0], yout);
      cv->next_q = cv->qprime;
      cv->next_h = cv->hprime;
      break;
    }
  }
  return istate;
}

/* ------------------------------------------------------------------------- */
/* statistics                                                                */
/* ------------------------------------------------------------------------- */

#define CV_GET(name, type, expr)                  \
  int name(void* cvode_mem, type* out)            \
  {                                               \
    CVodeMem cv = (CVodeMem)cvode_mem;            \
    if (!cv) { return CV_MEM_NULL; }              \
    *out = (expr);                                \
    return CV_SUCCESS;                            \
  }

CV_GET(CVodeGetNumSteps, long int, cv->nst)
CV_GET(CVodeGetNumRhsEvals, long int, cv->nfe)
CV_GET(CVodeGetNumLinSolvSetups, long int, cv->nsetups)
CV_GET(CVodeGetNumErrTestFails, long int, cv->netf)
CV_GET(CVodeGetNumNonlinSolvIters, long int, cv->nni)
CV_GET(CVodeGetNumNonlinSolvConvFails, long int, cv->nnf)
CV_GET(CVodeGetNumStepSolveFails, long int, cv->ncfn)
CV_GET(CVodeGetNumJacEvals, long int, cv->ls.nje)
CV_GET(CVodeGetLastOrder, int, cv->qu)
CV_GET(CVodeGetCurrentOrder, int, cv->next_q)
CV_GET(CVodeGetActualInitStep, sunrealtype, cv->h0u)
CV_GET(CVodeGetLastStep, sunrealtype, cv->hu)
CV_GET(CVodeGetCurrentStep, sunrealtype, cv->next_h)
CV_GET(CVodeGetCurrentTime, sunrealtype, cv->tn)
CV_GET(CVodeGetNumInitialCorrectionsUsed, long int, cv->ninit_used)

int CVodeGetAllStats(void* cvode_mem, long int* counts, int* orders, sunrealtype* steps)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  counts[0] = cv->nst;
  counts[1] = cv->nfe;
  counts[2] = cv->nsetups;
  counts[3] = cv->netf;
  counts[4] = cv->nni;
  counts[5] = cv->nnf;
  counts[6] = cv->ncfn;
  counts[7] = cv->ls.nje;
  orders[0] = cv->qu;
  orders[1] = cv->next_q;
  steps[0]  = cv->h0u;
  steps[1]  = cv->hu;
  steps[2]  = cv->next_h;
  steps[3]  = cv->tn;
  return CV_SUCCESS;
}

/* CVodePrintAllStats (cvode_io.c:1626-1714) with the reference's number
   formats (SUN_FORMAT_G "%.15g", SUN_FORMAT_E "% .15e", sundials_types.h:113-114)
   and table width (SUN_TABLE_WIDTH 29, sundials_utils.h:31) */
#define STATS_TABLE_WIDTH 29
static void stats_real(FILE* fp, SUNOutputFormat fmt, int start, const char* name,
                       sunrealtype value)
{
  if (fmt == SUN_OUTPUTFORMAT_TABLE) { fprintf(fp, "%-*s = %.15g\n", STATS_TABLE_WIDTH, name, value); }
  else { fprintf(fp, "%s%s,% .15e", start ? "" : ",", name, value); }
}

static void stats_long(FILE* fp, SUNOutputFormat fmt, int start, const char* name, long value)
{
  if (fmt == SUN_OUTPUTFORMAT_TABLE) { fprintf(fp, "%-*s = %ld\n", STATS_TABLE_WIDTH, name, value); }
  else { fprintf(fp, "%s%s,%ld", start ? "" : ",", name, value); }
}

int CVodePrintAllStats(void* cvode_mem, FILE* outfile, SUNOutputFormat fmt)
{
  CVodeMem cv = (CVodeMem)cvode_mem;
  if (!cv) { return CV_MEM_NULL; }
  if ((fmt != SUN_OUTPUTFORMAT_TABLE && fmt != SUN_OUTPUTFORMAT_CSV) || !outfile)
  {
    return CV_ILL_INPUT;
  }
  stats_real(outfile, fmt, 1, "Current time", cv->tn);
  stats_long(outfile, fmt, 0, "Steps", cv->nst);
  stats_long(outfile, fmt, 0, "Error test fails", cv->netf);
  stats_long(outfile, fmt, 0, "NLS step fails", cv->ncfn);
  stats_real(outfile, fmt, 0, "Initial step size", cv->h0u);
  stats_real(outfile, fmt, 0, "Last step size", cv->hu);
  stats_real(outfile, fmt, 0, "Current step size", cv->next_h);
  stats_long(outfile, fmt, 0, "Last method order", cv->qu);
  stats_long(outfile, fmt, 0, "Current method order", cv->next_q);
  stats_long(outfile, fmt, 0, "Stab. lim. order reductions", 0L);
  stats_long(outfile, fmt, 0, "RHS fn evals", cv->nfe);
  stats_long(outfile, fmt, 0, "NLS iters", cv->nni);
  stats_long(outfile, fmt, 0, "NLS fails", cv->nnf);
  if (cv->nst > 0)
  {
    stats_real(outfile, fmt, 0, "NLS iters per step", (sunrealtype)cv->nni / (sunrealtype)cv->nst);
  }
  stats_long(outfile, fmt, 0, "LS setups", cv->nsetups);
  if (cv->ls.LS)
  {
    /* direct solvers: no difference-quotient RHS, preconditioner or
       Jacobian-times evaluations */
    stats_long(outfile, fmt, 0, "Jac fn evals", cv->ls.nje);
    stats_long(outfile, fmt, 0, "LS RHS fn evals", 0L);
    stats_long(outfile, fmt, 0, "Prec setup evals", 0L);
    stats_long(outfile, fmt, 0, "Prec solves", 0L);
    stats_long(outfile, fmt, 0, "LS iters", cv->ls.nli);
    stats_long(outfile, fmt, 0, "LS fails", cv->ls.ncfl);
    stats_long(outfile, fmt, 0, "Jac-times setups", 0L);
    stats_long(outfile, fmt, 0, "Jac-times evals", 0L);
    if (cv->nni > 0)
    {
      stats_real(outfile, fmt, 0, "LS iters per NLS iter",
                 (sunrealtype)cv->ls.nli / (sunrealtype)cv->nni);
      stats_real(outfile, fmt, 0, "Jac evals per NLS iter",
                 (sunrealtype)cv->ls.nje / (sunrealtype)cv->nni);
      stats_real(outfile, fmt, 0, "Prec evals per NLS iter", ZERO);
    }
  }
  stats_long(outfile, fmt, 0, "Root fn evals", 0L);
  return CV_SUCCESS;
}

void CVodeFree(void** cvode_mem)
{
  if (!cvode_mem || !*cvode_mem) { return; }
  CVodeMem cv = (CVodeMem)(*cvode_mem);
  cvFreeVectors(cv);
  if (cv->ownNLS && cv->NLS) { SUNNonlinSolFree(cv->NLS); }
  if (cv->ls.x) { N_VDestroy(cv->ls.x); }
  if (cv->ls.savedJ) { SUNMatDestroy(cv->ls.savedJ); }
  free(cv);
  *cvode_mem = NULL;
}

/* the x^y used by the step-size controller, exported so its rounding can be
   checked (no reference counterpart) */
double SUNAMD_Pow(double x, double y) { return sunamd_pow(x, y); }

int SUNAMD_PowFastParts(double x, double y, double* res, double* hi, double* lo, double* bound,
                        int* scale)
{
  *res = 0.0;
  return sunamd_pow_fast(x, y, res, hi, lo, bound, scale);
}

double SUNAMD_PowExact(double x, double y) { return sunamd_pow_exact(x, y); }
