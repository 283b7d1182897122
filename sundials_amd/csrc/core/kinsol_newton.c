/* -----------------------------------------------------------------------------
 * KINSOL's Newton iteration (strategies KIN_NONE and KIN_LINESEARCH) and its
 * interface to a matrix-based direct SUNLinearSolver.  Restatement of the
 * reference's control flow:
 *
 *   KINSol Newton loop               kinsol.c:560-760
 *   KINSolInit                       kinsol.c:1017-1224
 *   KINLinSolDrv                     kinsol.c:1231-1290
 *   KINFullNewton                    kinsol.c:1293-1390
 *   KINLineSearch (alpha/beta)       kinsol.c:1440-1765
 *   KINConstraint                    kinsol.c:1769-1810
 *   KINStop (with residual monitor)  kinsol.c:1813-1972
 *   KINScSNorm                       kinsol.c:2051-2062
 *   KINSetLinearSolver, KINSetJacFn  kinsol_ls.c:44-270
 *   kinLsDQJac, kinLsDenseDQJac      kinsol_ls.c:713-850
 *   kinLsInitialize/Setup/Solve      kinsol_ls.c:1006-1350
 *   option setters                   kinsol_io.c (KINSetConstraints,
 *     KINSetScaledStepTol, KINSetMaxSetupCalls, KINSetMaxSubSetupCalls,
 *     KINSetMaxNewtonStep, KINSetNoInitSetup, KINSetNoResMon,
 *     KINSetMaxBetaFails, KINSetRelErrFunc, KINSetResMonParams,
 *     KINSetResMonConstVal)
 *
 * Every vector operation goes through the N_Vector ops in the reference's
 * order, and the linear algebra through the SUNLinearSolver ops, so with the
 * MI355X vector and the block-diagonal matrix / batched LU the whole Newton
 * iteration runs on the device.  Only iterative (matrix-free) linear solvers
 * and the Picard strategy are outside this library.
 * ---------------------------------------------------------------------------*/
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kinsol_impl.h"
#include "sunamd_pow.h"

#define POINT1    0.1
#define POINT01   0.01
#define POINT9    0.9
#define POINT99   0.99
#define POINT0001 0.0001
#define THREE     3.0
#define TWELVE    12.0
#define THOUSAND  1000.0
#define TWOPT5    2.5
#define MAX_RECVR 5

#define KMEM()                          \
  KinMem* k = (KinMem*)kinmem;          \
  if (!k) { return KIN_MEM_NULL; }

/* ------------------------------------------------------------------------- */
/* option setters (kinsol_io.c)                                              */
/* ------------------------------------------------------------------------- */
int KINSetConstraints(void* kinmem, N_Vector constraints)
{
  KMEM();
  if (!constraints)
  {
    if (k->constraints) { N_VDestroy(k->constraints); }
    k->constraints = NULL;
    return KIN_SUCCESS;
  }
  if (N_VMaxNorm(constraints) > TWOPT5) { return KIN_ILL_INPUT; }
  if (!k->constraints)
  {
    k->constraints = N_VClone(constraints);
    if (!k->constraints) { return KIN_MEM_FAIL; }
  }
  N_VScale(ONE, constraints, k->constraints);
  return KIN_SUCCESS;
}

int KINSetScaledStepTol(void* kinmem, sunrealtype scsteptol)
{
  KMEM();
  if (scsteptol < ZERO) { return KIN_ILL_INPUT; }
  k->scsteptol = (scsteptol == ZERO) ? sunamd_pow(UROUND, TWOTHIRDS) : scsteptol;
  return KIN_SUCCESS;
}

int KINSetMaxSetupCalls(void* kinmem, long int msbset)
{
  KMEM();
  if (msbset < 0) { return KIN_ILL_INPUT; }
  k->msbset = (msbset == 0) ? MSBSET_DEFAULT : msbset;
  return KIN_SUCCESS;
}

int KINSetMaxSubSetupCalls(void* kinmem, long int msbsetsub)
{
  KMEM();
  if (msbsetsub < 0) { return KIN_ILL_INPUT; }
  k->msbset_sub = (msbsetsub == 0) ? MSBSET_SUB_DEFAULT : msbsetsub;
  return KIN_SUCCESS;
}

/* 0 selects the default 1000 * ||u0||_{uscale}, computed at each KINSol */
int KINSetMaxNewtonStep(void* kinmem, sunrealtype mxnewtstep)
{
  KMEM();
  if (mxnewtstep < ZERO) { return KIN_ILL_INPUT; }
  k->mxnstepin = mxnewtstep;
  return KIN_SUCCESS;
}

int KINSetNoInitSetup(void* kinmem, sunbooleantype noInitSetup)
{
  KMEM();
  k->no_init_setup = noInitSetup ? 1 : 0;
  return KIN_SUCCESS;
}

int KINSetNoResMon(void* kinmem, sunbooleantype noResMon)
{
  KMEM();
  k->no_res_mon = noResMon ? 1 : 0;
  return KIN_SUCCESS;
}

int KINSetMaxBetaFails(void* kinmem, long int mxnbcf)
{
  KMEM();
  if (mxnbcf < 0) { return KIN_ILL_INPUT; }
  k->mxnbcf = (mxnbcf == 0) ? MXNBCF_DEFAULT : mxnbcf;
  return KIN_SUCCESS;
}

int KINSetRelErrFunc(void* kinmem, sunrealtype relfunc)
{
  KMEM();
  if (relfunc < ZERO) { return KIN_ILL_INPUT; }
  k->sqrt_relfunc = sqrt(relfunc == ZERO ? UROUND : relfunc);
  return KIN_SUCCESS;
}

int KINSetResMonParams(void* kinmem, sunrealtype omegamin, sunrealtype omegamax)
{
  KMEM();
  if (omegamin < ZERO || omegamax < ZERO) { return KIN_ILL_INPUT; }
  const sunrealtype lo = (omegamin == ZERO) ? OMEGA_MIN_DEFAULT : omegamin;
  const sunrealtype hi = (omegamax == ZERO) ? OMEGA_MAX_DEFAULT : omegamax;
  if (lo > hi) { return KIN_ILL_INPUT; }
  k->omega_min = lo;
  k->omega_max = hi;
  return KIN_SUCCESS;
}

/* a fixed residual-monitoring constant; 0 selects the fnorm-based formula */
int KINSetResMonConstVal(void* kinmem, sunrealtype omegaconst)
{
  KMEM();
  if (omegaconst < ZERO) { return KIN_ILL_INPUT; }
  k->omega = omegaconst;
  return KIN_SUCCESS;
}

/* The scaled norms as one device pass each when the vectors are the MI355X
   library's (looked up in the loaded library, like CVODE's fused kernels);
   bit-identical to the op chains.  Fails, leaving the op chains on, when the
   library or its vector is absent. */
int KINSetUseFusedKernels(void* kinmem, sunbooleantype onoff)
{
  KMEM();
  if (!onoff)
  {
    k->usefused = 0;
    return KIN_SUCCESS;
  }
  if (!k->malloc_done) { return KIN_ILL_INPUT; }
  int (*is_mi355)(N_Vector) = (int (*)(N_Vector))dlsym(RTLD_DEFAULT, "SUNAMD_NVectorIsMi355");
  void* ss = dlsym(RTLD_DEFAULT, "kinScSNorm_fused_Mi355");
  void* sf = dlsym(RTLD_DEFAULT, "kinScFNorm_fused_Mi355");
  if (!is_mi355 || !ss || !sf || !is_mi355(k->vtemp1)) { return KIN_ILL_INPUT; }
  k->fused_ssnorm = (sunrealtype(*)(N_Vector, N_Vector, N_Vector))ss;
  k->fused_sfnorm = (sunrealtype(*)(N_Vector, N_Vector))sf;
  k->usefused     = 1;
  return KIN_SUCCESS;
}

int KINGetNumBetaCondFails(void* kinmem, long int* nbcfails)
{
  KMEM();
  *nbcfails = k->nbcf;
  return KIN_SUCCESS;
}

int KINGetNumBacktrackOps(void* kinmem, long int* nbacktr)
{
  KMEM();
  *nbacktr = k->nbktrk;
  return KIN_SUCCESS;
}

int KINGetStepLength(void* kinmem, sunrealtype* steplength)
{
  KMEM();
  *steplength = k->stepl;
  return KIN_SUCCESS;
}

/* ------------------------------------------------------------------------- */
/* linear solver interface (kinsol_ls.c)                                      */
/* ------------------------------------------------------------------------- */

/* Public content layout of a SUNMATRIX_DENSE matrix
   (include/sunmatrix/sunmatrix_dense.h:48-55): the difference-quotient
   Jacobian writes each column in place through cols[j]. */
struct dense_matrix_layout
{
  sunindextype M;
  sunindextype N;
  sunrealtype* data;
  sunindextype ldata;
  sunrealtype** cols;
};

/* kinsol_ls.c:777-850: column j of J is (F(u + sigma_j e_j) - F(u)) / sigma_j
   with sigma_j = sqrt(relfunc) * max(|u_j|, 1/uscale_j), signed like u_j */
static int dense_dq_jac(N_Vector u, N_Vector fu, SUNMatrix Jac, KinMem* k, N_Vector tmp1,
                        N_Vector tmp2)
{
  const struct dense_matrix_layout* c = (const struct dense_matrix_layout*)Jac->content;
  sunrealtype* tmp2_data = N_VGetArrayPointer(tmp2);
  sunrealtype* ud        = N_VGetArrayPointer(u);
  sunrealtype* usd       = N_VGetArrayPointer(k->uscale);
  int r                  = 0;
  for (sunindextype j = 0; j < c->N; j++)
  {
    N_VSetArrayPointer(c->cols[j], tmp2);
    const sunrealtype saved = ud[j];
    const sunrealtype scl   = ONE / usd[j];
    const sunrealtype sgn   = (saved >= ZERO) ? ONE : -ONE;
    const sunrealtype mag   = fabs(saved);
    const sunrealtype inc   = k->sqrt_relfunc * (mag > scl ? mag : scl) * sgn;
    ud[j] += inc;
    r = k->func(u, tmp1, k->user_data);
    k->ls->nfeDQ++;
    if (r != 0) { break; }
    ud[j]                   = saved;
    const sunrealtype inv   = ONE / inc;
    N_VLinearSum(inv, tmp1, -inv, fu, tmp2);
  }
  N_VSetArrayPointer(tmp2_data, tmp2);
  return r;
}

static int dq_jac_entry(N_Vector u, N_Vector fu, SUNMatrix J, void* kmem, N_Vector tmp1,
                        N_Vector tmp2)
{
  KinMem* k = (KinMem*)kmem;
  if (SUNMatGetID(J) != SUNMATRIX_DENSE) { return KIN_ILL_INPUT; }
  return dense_dq_jac(u, fu, J, k, tmp1, tmp2);
}

void kin_ls_free(KinMem* k)
{
  free(k->ls);
  k->ls = NULL;
}

int KINSetLinearSolver(void* kinmem, SUNLinearSolver LS, SUNMatrix A)
{
  KinMem* k = (KinMem*)kinmem;
  if (!k) { return KINLS_MEM_NULL; }
  if (!LS || !LS->ops->gettype || !LS->ops->solve) { return KINLS_ILL_INPUT; }
  /* the vector templates come from KINInit */
  if (!k->malloc_done) { return KINLS_ILL_INPUT; }
  if (!k->vtemp1->ops->nvconst || !k->vtemp1->ops->nvdotprod) { return KINLS_ILL_INPUT; }
  /* direct, matrix-based solvers only: iterative and matrix-embedded
     solvers are not part of this library */
  if (SUNLinSolGetType(LS) != SUNLINEARSOLVER_DIRECT || !A) { return KINLS_ILL_INPUT; }
  kin_ls_free(k);
  KinLsMem* m = (KinLsMem*)calloc(1, sizeof(KinLsMem));
  if (!m) { return KINLS_MEM_FAIL; }
  m->LS     = LS;
  m->J      = A;
  m->jacDQ  = 1;
  m->jac    = dq_jac_entry;
  m->J_data = k;
  k->ls     = m;
  return KINLS_SUCCESS;
}

int KINSetJacFn(void* kinmem, KINLsJacFn jac)
{
  KinMem* k = (KinMem*)kinmem;
  if (!k) { return KINLS_MEM_NULL; }
  if (!k->ls) { return KINLS_LMEM_NULL; }
  if (jac)
  {
    k->ls->jacDQ  = 0;
    k->ls->jac    = jac;
    k->ls->J_data = k->user_data;
  }
  else
  {
    k->ls->jacDQ  = 1;
    k->ls->jac    = dq_jac_entry;
    k->ls->J_data = k;
  }
  return KINLS_SUCCESS;
}

#define LSMEM()                                  \
  KinMem* k = (KinMem*)kinmem;                   \
  if (!k) { return KINLS_MEM_NULL; }             \
  if (!k->ls) { return KINLS_LMEM_NULL; }

int KINGetNumJacEvals(void* kinmem, long int* njevals)
{
  LSMEM();
  *njevals = k->ls->nje;
  return KINLS_SUCCESS;
}

int KINGetNumLinFuncEvals(void* kinmem, long int* nfevals)
{
  LSMEM();
  *nfevals = k->ls->nfeDQ;
  return KINLS_SUCCESS;
}

int KINGetLastLinFlag(void* kinmem, long int* flag)
{
  LSMEM();
  *flag = k->ls->last_flag;
  return KINLS_SUCCESS;
}

/* kinsol_ls.c:1006-1145 (direct branch) */
static int ls_initialize(KinMem* k)
{
  KinLsMem* m = k->ls;
  if (m->jacDQ)
  {
    if (!m->J->ops->getid || SUNMatGetID(m->J) != SUNMATRIX_DENSE ||
        !k->vtemp1->ops->nvgetarraypointer || !k->vtemp1->ops->nvsetarraypointer)
    {
      m->last_flag = KINLS_ILL_INPUT;
      return KINLS_ILL_INPUT;
    }
    m->jac    = dq_jac_entry;
    m->J_data = k;
  }
  else { m->J_data = k->user_data; }
  m->nje = m->nfeDQ = 0;
  m->last_flag      = KINLS_SUCCESS;
  if (m->LS->ops->setscalingvectors &&
      SUNLinSolSetScalingVectors(m->LS, k->fscale, k->fscale) != SUN_SUCCESS)
  {
    return KINLS_SUNLS_FAIL;
  }
  m->last_flag = SUNLinSolInitialize(m->LS);
  return (int)m->last_flag;
}

/* kinsol_ls.c:1148-1200 */
static int ls_setup(KinMem* k)
{
  KinLsMem* m = k->ls;
  m->nje++;
  if (SUNMatZero(m->J) != SUN_SUCCESS)
  {
    m->last_flag = KINLS_SUNMAT_FAIL;
    return KINLS_SUNMAT_FAIL;
  }
  if (m->jac(k->uu, k->fval, m->J, m->J_data, k->vtemp1, k->vtemp2) != 0)
  {
    m->last_flag = KINLS_JACFUNC_ERR;
    return KINLS_JACFUNC_ERR;
  }
  m->last_flag = SUNLinSolSetup(m->LS, m->J);
  k->nnilset   = k->nni;
  return (int)m->last_flag;
}

/* kinsol_ls.c:1207-1345 (direct branch: x = J^{-1} b, and for the line
   search sFdotJp = F . (fscale^2 * J p) with J p = b) */
static int ls_solve(KinMem* k, N_Vector x, N_Vector b)
{
  KinLsMem* m = k->ls;
  N_VConst(ZERO, x);
  if (SUNLinSolSetZeroGuess(m->LS, SUNTRUE) != SUN_SUCCESS) { return -1; }
  const int r  = SUNLinSolSolve(m->LS, m->J, x, b, ZERO);
  m->last_flag = r;
  if (r != SUN_SUCCESS && r != SUNLS_RES_REDUCED)
  {
    if (r == SUNLS_ATIMES_FAIL_REC || r == SUNLS_PSOLVE_FAIL_REC) { return 1; }
    return r;
  }
  if (k->strategy == KIN_LINESEARCH)
  {
    N_VProd(b, k->fscale, b);
    N_VProd(b, k->fscale, b);
    k->sFdotJp = N_VDotProd(k->fval, b);
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Newton iteration (kinsol.c)                                                */
/* ------------------------------------------------------------------------- */

/* kinsol.c:2051-2062: max_i |v_i| / (1/uscale_i + |u_i|) */
static sunrealtype sc_snorm(KinMem* k, N_Vector v, N_Vector u)
{
  if (k->usefused)
  {
    const sunrealtype r = k->fused_ssnorm(v, u, k->uscale);
    if (r >= ZERO) { return r; }
  }
  N_VInv(k->uscale, k->vtemp1);
  N_VAbs(u, k->vtemp2);
  N_VLinearSum(ONE, k->vtemp1, ONE, k->vtemp2, k->vtemp1);
  N_VDiv(v, k->vtemp1, k->vtemp1);
  return N_VMaxNorm(k->vtemp1);
}

/* kinsol.c:1017-1224 */
static int sol_init(KinMem* k)
{
  if (!k->uscale || N_VMin(k->uscale) <= ZERO) { return KIN_ILL_INPUT; }
  if (!k->fscale || N_VMin(k->fscale) <= ZERO) { return KIN_ILL_INPUT; }
  if (!k->ls) { return KIN_ILL_INPUT; }
  if (k->constraints)
  {
    if (!k->constraints->ops->nvconstrmask || !k->constraints->ops->nvminquotient)
    {
      return KIN_ILL_INPUT;
    }
    if (!N_VConstrMask(k->constraints, k->uu, k->vtemp1)) { return KIN_ILL_INPUT; }
  }
  k->mxnewtstep = (k->mxnstepin == ZERO) ? THOUSAND * N_VWL2Norm(k->uu, k->uscale)
                                         : k->mxnstepin;
  if (k->mxnewtstep < ONE) { k->mxnewtstep = ONE; }

  k->nfe = k->nnilset = k->nnilset_sub = k->nni = k->nbcf = k->nbktrk = 0;
  const int r = k->func(k->uu, k->fval, k->user_data);
  k->nfe++;
  if (r < 0) { return KIN_SYSFUNC_FAIL; }
  if (r > 0) { return KIN_FIRST_SYSFUNC_ERR; }
  if (kin_sc_fnorm(k, k->fval, k->fscale) <= POINT01 * k->fnormtol)
  {
    k->fnorm = N_VWL2Norm(k->fval, k->fscale);
    return KIN_INITIAL_GUESS_OK;
  }
  if (ls_initialize(k) != 0) { return KIN_LINIT_FAIL; }
  k->fnorm     = N_VWL2Norm(k->fval, k->fscale);
  k->f1norm    = HALF * k->fnorm * k->fnorm;
  k->fnorm_sub = k->fnorm;
  return KIN_SUCCESS;
}

/* kinsol.c:1231-1290: (re)factor when due, then solve J p = -F */
static int lin_sol_drv(KinMem* k)
{
  if (k->nni - k->nnilset >= k->msbset)
  {
    k->sthrsh           = TWO;
    k->update_fnorm_sub = 1;
  }
  for (;;)
  {
    k->jac_current = 0;
    if (k->sthrsh > 1.5)
    {
      const int r      = ls_setup(k);
      k->jac_current   = 1;
      k->nnilset       = k->nni;
      k->nnilset_sub   = k->nni;
      if (r != 0) { return KIN_LSETUP_FAIL; }
    }
    N_VScale(-ONE, k->fval, k->unew);
    const int r = ls_solve(k, k->pp, k->unew);
    if (r == 0) { return KIN_SUCCESS; }
    if (r < 0) { return KIN_LSOLVE_FAIL; }
    if (k->jac_current) { return KIN_LINSOLV_NO_RECOVERY; }
    k->sthrsh = TWO;
  }
}

/* kinsol.c:1769-1810: if u + p leaves the feasible set, stepmul is 0.9 times
   the largest fraction of p that keeps every constrained component feasible */
static int constraint_step(KinMem* k)
{
  N_VLinearSum(ONE, k->uu, ONE, k->pp, k->vtemp1);
  if (N_VConstrMask(k->constraints, k->vtemp1, k->vtemp2)) { return KIN_SUCCESS; }
  N_VAbs(k->pp, k->vtemp1);
  N_VProd(k->vtemp2, k->vtemp1, k->vtemp1);
  N_VAbs(k->uu, k->vtemp2);
  k->stepmul = POINT9 * N_VMinQuotient(k->vtemp2, k->vtemp1);
  return CONSTR_VIOLATED;
}

/* shared start of both globalisations: given pnorm = ||p||_{uscale}, cap the
   step at mxnewtstep and cut it back into the feasible set.  Returns
   STEP_TOO_SMALL or 0; ratio and pnorm are updated in place and *capped
   reports that a cut happened (the line search then uses rlmax = 1). */
static int limit_step(KinMem* k, sunrealtype* ratio, sunrealtype* pnorm, int* capped)
{
  *capped  = 0;
  *ratio   = ONE;
  k->stepl = *pnorm;
  if (*pnorm > k->mxnewtstep)
  {
    *ratio = k->mxnewtstep / *pnorm;
    N_VScale(*ratio, k->pp, k->pp);
    *pnorm   = k->mxnewtstep;
    *capped  = 1;
    k->stepl = *pnorm;
  }
  k->stepmul = ONE;
  if (k->constraints && constraint_step(k) == CONSTR_VIOLATED)
  {
    N_VScale(k->stepmul, k->pp, k->pp);
    *ratio *= k->stepmul;
    *pnorm *= k->stepmul;
    *capped  = 1;
    k->stepl = *pnorm;
    if (*pnorm <= k->scsteptol)
    {
      N_VLinearSum(ONE, k->uu, ONE, k->pp, k->unew);
      return STEP_TOO_SMALL;
    }
  }
  return 0;
}

/* evaluate F at u + p, halving p on recoverable failures (MAX_RECVR tries) */
static int eval_full_step(KinMem* k, sunrealtype* ratio, sunrealtype* pnorm, int* capped)
{
  for (int i = 1; i <= MAX_RECVR; i++)
  {
    N_VLinearSum(ONE, k->uu, ONE, k->pp, k->unew);
    const int r = k->func(k->unew, k->fval, k->user_data);
    k->nfe++;
    if (r == 0) { return KIN_SUCCESS; }
    if (r < 0) { return KIN_SYSFUNC_FAIL; }
    N_VScale(HALF, k->pp, k->pp);
    *ratio *= HALF;
    *pnorm *= HALF;
    *capped  = 1;
    k->stepl = *pnorm;
  }
  return KIN_REPTD_SYSFUNC_ERR;
}

/* kinsol.c:1293-1390 */
static int full_newton(KinMem* k, sunrealtype* fnormp, sunrealtype* f1normp, int* max_step)
{
  sunrealtype ratio, pnorm = N_VWL2Norm(k->pp, k->uscale);
  int capped;
  int r = limit_step(k, &ratio, &pnorm, &capped);
  if (r) { return r; }
  r = eval_full_step(k, &ratio, &pnorm, &capped);
  if (r != KIN_SUCCESS) { return r; }
  *fnormp  = N_VWL2Norm(k->fval, k->fscale);
  *f1normp = HALF * (*fnormp) * (*fnormp);
  k->sFdotJp *= ratio;
  k->sJpnorm *= ratio;
  if (pnorm > POINT99 * k->mxnewtstep) { *max_step = 1; }
  return KIN_SUCCESS;
}

/* evaluate F at u + rl p; fnormp / f1normp follow */
static int eval_at(KinMem* k, sunrealtype rl, sunrealtype* fnormp, sunrealtype* f1normp)
{
  N_VLinearSum(ONE, k->uu, rl, k->pp, k->unew);
  const int r = k->func(k->unew, k->fval, k->user_data);
  k->nfe++;
  if (r != 0) { return KIN_SYSFUNC_FAIL; }
  *fnormp  = N_VWL2Norm(k->fval, k->fscale);
  *f1normp = HALF * (*fnormp) * (*fnormp);
  return KIN_SUCCESS;
}

/* kinsol.c:1440-1765: backtracking with quadratic then cubic models until the
   alpha (sufficient decrease) condition holds, then, if the beta condition
   fails, expansion from a full step and bisection between the last two trial
   lengths */
static int line_search(KinMem* k, sunrealtype* fnormp, sunrealtype* f1normp, int* max_step)
{
  const sunrealtype alpha = POINT0001, beta = POINT9;
  sunrealtype ratio, pnorm = N_VWL2Norm(k->pp, k->uscale);
  int capped;
  long int nbk   = 0;
  int first_back = 1;
  sunrealtype rlprev = ZERO, f1nprv = ZERO;
  *max_step = 0;

  /* rlmax is mxnewtstep / pnorm unless the step was cut (then 1) */
  sunrealtype rlmax = k->mxnewtstep / pnorm;
  int r             = limit_step(k, &ratio, &pnorm, &capped);
  if (r) { return r; }
  r = eval_full_step(k, &ratio, &pnorm, &capped);
  if (r != KIN_SUCCESS) { return r; }
  if (capped) { rlmax = ONE; }
  *fnormp  = N_VWL2Norm(k->fval, k->fscale);
  *f1normp = HALF * (*fnormp) * (*fnormp);

  const sunrealtype slpi    = k->sFdotJp * ratio;
  const sunrealtype rlength = sc_snorm(k, k->pp, k->uu);
  const sunrealtype rlmin   = k->scsteptol / rlength;
  sunrealtype rl            = ONE;

  for (;;)
  {
    const sunrealtype alpha_cond = k->f1norm + alpha * slpi * rl;
    if (*f1normp <= alpha_cond) { break; }
    sunrealtype rltmp;
    if (first_back)
    {
      rltmp      = -slpi / (TWO * ((*f1normp) - k->f1norm - slpi));
      first_back = 0;
    }
    else
    {
      sunrealtype t1   = (*f1normp) - k->f1norm - (rl * slpi);
      sunrealtype t2   = f1nprv - k->f1norm - (rlprev * slpi);
      sunrealtype ca   = ((ONE / (rl * rl)) * t1) - ((ONE / (rlprev * rlprev)) * t2);
      sunrealtype cb   = ((-rlprev / (rl * rl)) * t1) + ((rl / (rlprev * rlprev)) * t2);
      const sunrealtype inv = ONE / (rl - rlprev);
      ca *= inv;
      cb *= inv;
      const sunrealtype disc = (cb * cb) - (THREE * ca * slpi);
      if (fabs(ca) < UROUND) { rltmp = -slpi / (TWO * cb); }
      else { rltmp = (-cb + sqrt(disc)) / (THREE * ca); }
    }
    if (rltmp > HALF * rl) { rltmp = HALF * rl; }
    rlprev = rl;
    f1nprv = *f1normp;
    const sunrealtype lo = POINT1 * rl;
    rl                   = (lo > rltmp) ? lo : rltmp;
    nbk++;
    r = eval_at(k, rl, fnormp, f1normp);
    if (r) { return r; }
    if (rl < rlmin)
    {
      N_VScale(ONE, k->uu, k->unew);
      return STEP_TOO_SMALL;
    }
  }

  sunrealtype beta_cond = k->f1norm + beta * slpi * rl;
  if (*f1normp < beta_cond)
  {
    sunrealtype alpha_cond = k->f1norm + alpha * slpi * rl;
    if (rl == ONE && pnorm < k->mxnewtstep)
    {
      do {
        rlprev = rl;
        f1nprv = *f1normp;
        rl     = (TWO * rl < rlmax) ? TWO * rl : rlmax;
        nbk++;
        r = eval_at(k, rl, fnormp, f1normp);
        if (r) { return r; }
        alpha_cond = k->f1norm + alpha * slpi * rl;
        beta_cond  = k->f1norm + beta * slpi * rl;
      }
      while (*f1normp <= alpha_cond && *f1normp < beta_cond && rl < rlmax);
    }
    if (rl < ONE || (rl > ONE && *f1normp > alpha_cond))
    {
      sunrealtype rllo   = (rl < rlprev) ? rl : rlprev;
      sunrealtype rldiff = fabs(rlprev - rl);
      do {
        const sunrealtype rlinc = HALF * rldiff;
        rl                      = rllo + rlinc;
        nbk++;
        r = eval_at(k, rl, fnormp, f1normp);
        if (r) { return r; }
        alpha_cond = k->f1norm + alpha * slpi * rl;
        beta_cond  = k->f1norm + beta * slpi * rl;
        if (*f1normp > alpha_cond) { rldiff = rlinc; }
        else if (*f1normp < beta_cond)
        {
          rllo   = rl;
          rldiff = rldiff - rlinc;
        }
      }
      while (*f1normp > alpha_cond || (*f1normp < beta_cond && rldiff >= rlmin));
      if (*f1normp < beta_cond || (rldiff < rlmin && *f1normp > alpha_cond))
      {
        /* keep the last length that met the alpha condition */
        r = eval_at(k, rllo, fnormp, f1normp);
        if (r) { return r; }
        k->nbcf++;
      }
    }
  }
  k->nbktrk += nbk;
  k->sFdotJp = k->sFdotJp * rl * ratio;
  k->sJpnorm = k->sJpnorm * rl * ratio;
  if (rl * pnorm > POINT99 * k->mxnewtstep) { *max_step = 1; }
  return KIN_SUCCESS;
}

/* kinsol.c:1813-1972 */
static int stop_test(KinMem* k, int max_step, int sflag)
{
  if (sflag == STEP_TOO_SMALL)
  {
    if (!k->jac_current)
    {
      k->sthrsh = TWO;
      return RETRY_ITERATION;
    }
    return (k->strategy == KIN_NONE) ? KIN_STEP_LT_STPTOL : KIN_LINESEARCH_NONCONV;
  }
  if (kin_sc_fnorm(k, k->fval, k->fscale) <= k->fnormtol) { return KIN_SUCCESS; }

  N_Vector delta = k->pp;
  N_VLinearSum(ONE, k->unew, -ONE, k->uu, delta);
  const sunrealtype rlength = sc_snorm(k, delta, k->unew);
  if (rlength <= k->scsteptol)
  {
    if (!k->jac_current)
    {
      k->sthrsh = TWO;
      return CONTINUE_ITERATIONS;
    }
    return KIN_STEP_LT_STPTOL;
  }
  if (k->nni >= k->mxiter) { return KIN_MAXITER_REACHED; }
  k->ncscmx = max_step ? k->ncscmx + 1 : 0;
  if (k->ncscmx == 5) { return KIN_MXNEWT_5X_EXCEEDED; }

  if (!k->no_res_mon)
  {
    if (k->nni - k->nnilset_sub >= k->msbset_sub)
    {
      /* residual monitoring: refactor when ||F|| has not dropped by omega
         since the last check */
      k->nnilset_sub = k->nni;
      if (k->eval_omega)
      {
        sunrealtype omexp = k->fnorm / k->fnormtol - ONE;
        if (omexp < ZERO) { omexp = ZERO; }
        if (omexp > TWELVE) { k->omega = k->omega_max; }
        else
        {
          const sunrealtype w = k->omega_min * exp(omexp);
          k->omega            = (w < k->omega_max) ? w : k->omega_max;
        }
      }
      if (k->fnorm > k->omega * k->fnorm_sub)
      {
        if (!k->jac_current)
        {
          k->sthrsh = TWO;
          return CONTINUE_ITERATIONS;
        }
      }
      else
      {
        k->fnorm_sub = k->fnorm;
        k->sthrsh    = ONE;
      }
    }
    else
    {
      if (k->retry_nni || k->update_fnorm_sub) { k->fnorm_sub = k->fnorm; }
      k->update_fnorm_sub = 0;
      k->sthrsh           = ONE;
    }
  }
  return CONTINUE_ITERATIONS;
}

/* kinsol.c:560-760 (KIN_NONE and KIN_LINESEARCH) */
int kin_newton(KinMem* k)
{
  int ret = sol_init(k);
  if (ret != KIN_SUCCESS) { return ret; }
  k->ncscmx     = 0;
  k->sthrsh     = k->no_init_setup ? ONE : TWO;
  k->eval_omega = (k->omega == ZERO);

  sunrealtype fnormp = -ONE, f1normp = -ONE;
  int max_step       = 0;
  for (;;)
  {
    k->retry_nni = 0;
    k->nni++;
    for (;;) /* repeated when the stop test asks for a retry */
    {
      int sflag = 0;
      ret       = lin_sol_drv(k);
      if (ret != KIN_SUCCESS) { return ret; }
      if (k->strategy == KIN_NONE)
      {
        sflag = full_newton(k, &fnormp, &f1normp, &max_step);
        if (sflag == KIN_SYSFUNC_FAIL || sflag == KIN_REPTD_SYSFUNC_ERR) { return sflag; }
      }
      else
      {
        sflag = line_search(k, &fnormp, &f1normp, &max_step);
        if (sflag == KIN_SYSFUNC_FAIL || sflag == KIN_REPTD_SYSFUNC_ERR) { return sflag; }
        if (k->nbcf > k->mxnbcf) { return KIN_LINESEARCH_BCFAIL; }
      }
      k->fnorm = fnormp;
      ret      = stop_test(k, max_step, sflag);
      if (ret != RETRY_ITERATION) { break; }
      k->retry_nni = 1;
    }
    N_VScale(ONE, k->unew, k->uu);
    k->f1norm = f1normp;
    if (ret != CONTINUE_ITERATIONS) { break; }
  }
  return ret;
}
