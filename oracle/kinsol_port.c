/* -----------------------------------------------------------------------------
 * CPU ORACLE (test infrastructure): the fixed-point problem of
 * include/sunamd/kin_fp.h on the serial vector -- the map with the cell
 * arithmetic of sundials_amd/csrc/problems/kin_cell.h, the damping and depth
 * functions of examples/kinsol/serial/kinAnalytic_fp.c:352-396 -- solved by
 * the KINSOL restatement (core/kinsol_fp.c), the CPU twin of
 * SUNAMD_KinSolve_Mi355.  With ncells = 1, variant 0 and no perturbation it
 * is the reference example itself.
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <time.h>

#include "kin_cell.h"
#include "oracle.h"
#include "sunamd/kin_fp.h"
#include "sunamd/kinsol_sunamd.h"

typedef struct
{
  const SUNAMD_KinOptions* o;
  N_Vector tmp;
} KinHost;

static int h_g(N_Vector uv, N_Vector gv, void* ud)
{
  const KinHost* H = (const KinHost*)ud;
  const double* u  = N_VGetArrayPointer(uv);
  double* g        = N_VGetArrayPointer(gv);
  for (long c = 0; c < H->o->ncells; c++) { kin_cell_g(H->o->variant, u + 3 * c, g + 3 * c); }
  return 0;
}

/* gain-based damping: ||f||^2 = (g - u).(g - u) */
static int h_damping(long int iter, N_Vector u, N_Vector g, sunrealtype* qt_fn, long int depth,
                     void* ud, sunrealtype* beta)
{
  (void)iter;
  const KinHost* H = (const KinHost*)ud;
  if (depth == 0)
  {
    *beta = 0.5;
    return 0;
  }
  double q2 = 0.0;
  for (long int i = 0; i < depth; i++) { q2 += qt_fn[i] * qt_fn[i]; }
  N_VLinearSum(1.0, g, -1.0, u, H->tmp);
  const double f2   = N_VDotProd(H->tmp, H->tmp);
  const double arg  = 1.0 - q2 / f2;
  const double gain = (arg <= 0.0) ? 0.0 : sqrt(arg);
  *beta             = 0.9 - 0.5 * gain;
  return 0;
}

static int h_depth(long int iter, N_Vector u, N_Vector g, N_Vector f, N_Vector* df,
                   sunrealtype* R, long int depth, void* ud, long int* new_depth,
                   sunbooleantype* remove)
{
  (void)u; (void)g; (void)f; (void)df; (void)R; (void)ud; (void)remove;
  *new_depth = (iter < 2) ? 1 : depth;
  return 0;
}

int oracle_kin_solve(const SUNAMD_KinOptions* o, double* stats, double* uout)
{
  SUNContext ctx = NULL;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) { return -100; }
  const long n = 3 * o->ncells;
  N_Vector u = N_VNew_Serial(n, ctx), scale = N_VClone(u);
  KinHost H  = {o, N_VClone(u)};
  double* ud = N_VGetArrayPointer(u);
  for (long c = 0; c < o->ncells; c++) { kin_cell_u0(c, o->perturb, ud + 3 * c); }
  N_VConst(1.0, scale);
  void* kmem = KINCreate(ctx);
  int rc     = KINSetUserData(kmem, &H);
  rc |= KINSetMAA(kmem, o->m_aa);
  rc |= KINSetOrthAA(kmem, o->orth_aa);
  rc |= KINInit(kmem, h_g, u);
  rc |= KINSetFuncNormTol(kmem, o->tol);
  rc |= KINSetNumMaxIters(kmem, o->maxiter);
  if (o->m_aa == 0) { rc |= KINSetDamping(kmem, o->damping_fp); }
  else
  {
    rc |= KINSetDampingAA(kmem, o->damping_aa);
    rc |= KINSetDelayAA(kmem, o->delay_aa);
  }
  if (o->damping_fn) { rc |= KINSetDampingFn(kmem, h_damping); }
  if (o->depth_fn) { rc |= KINSetDepthFn(kmem, h_depth); }
  if (rc) { return -101; }
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  const int flag = KINSol(kmem, u, KIN_FP, scale, scale);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  long nni = 0, nfe = 0;
  double fnorm = 0.0;
  KINGetNumNonlinSolvIters(kmem, &nni);
  KINGetNumFuncEvals(kmem, &nfe);
  KINGetFuncNorm(kmem, &fnorm);
  if (stats)
  {
    stats[0] = flag;
    stats[1] = (double)nni;
    stats[2] = (double)nfe;
    stats[3] = fnorm;
    stats[4] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  }
  if (uout)
  {
    for (long i = 0; i < n; i++) { uout[i] = ud[i]; }
  }
  KINFree(&kmem);
  N_VDestroy(H.tmp);
  N_VDestroy(u);
  N_VDestroy(scale);
  SUNContext_Free(&ctx);
  return flag;
}

/* ------------------------------------------------------------------------- */
/* Newton path: the constrained cell system of problems/kin_fertron.h solved  */
/* by the KINSOL restatement's KIN_NONE / KIN_LINESEARCH (core/               */
/* kinsol_newton.c) on the serial vector, with either the dense DQ Jacobian  */
/* (SUNDenseMatrix + SUNLinSol_Dense, as examples/kinsol/serial/             */
/* kinFerTron_dns.c:165-200) or the analytic block Jacobian (block-dense     */
/* matrix + per-block LU, the CPU twin of SUNAMD_KinNewtonSolve_Mi355).      */
/* ------------------------------------------------------------------------- */
#include "kin_fertron.h"
#include "sunamd/kin_newton.h"

typedef struct
{
  const SUNAMD_KinNewtonOptions* o;
} KinNHost;

static int hn_f(N_Vector uv, N_Vector fv, void* ud)
{
  const KinNHost* H = (const KinNHost*)ud;
  const double* u   = N_VGetArrayPointer(uv);
  double* f         = N_VGetArrayPointer(fv);
  for (long c = 0; c < H->o->ncells; c++) { kft_f(H->o->variant, u + 6 * c, f + 6 * c); }
  return 0;
}

static int hn_jac(N_Vector uv, N_Vector fu, SUNMatrix J, void* ud, N_Vector t1, N_Vector t2)
{
  (void)fu; (void)t1; (void)t2;
  const KinNHost* H = (const KinNHost*)ud;
  const double* u   = N_VGetArrayPointer(uv);
  double* a         = SUNBlockDenseMatrix_Oracle_Data(J);
  for (long c = 0; c < H->o->ncells; c++) { kft_jac(H->o->variant, u + 6 * c, a + 36 * c); }
  return 0;
}

static void hn_stats(void* kmem, int flag, double* st)
{
  long nni = 0, nfe = 0, nje = 0, nfeD = 0, nbcf = 0, nbk = 0;
  KINGetNumNonlinSolvIters(kmem, &nni);
  KINGetNumFuncEvals(kmem, &nfe);
  KINGetNumJacEvals(kmem, &nje);
  KINGetNumLinFuncEvals(kmem, &nfeD);
  KINGetNumBetaCondFails(kmem, &nbcf);
  KINGetNumBacktrackOps(kmem, &nbk);
  st[0] = flag;
  st[1] = (double)nni;
  st[2] = (double)nfe;
  st[3] = (double)nje;
  st[4] = (double)nfeD;
  st[5] = (double)nbcf;
  st[6] = (double)nbk;
}

/* shared set-up: vectors, constraints, KINSOL memory, matrix and solver */
typedef struct
{
  SUNContext ctx;
  N_Vector u, s, c;
  SUNMatrix J;
  SUNLinearSolver LS;
  void* kmem;
} KinNRun;

static void hn_free(KinNRun* R)
{
  if (R->kmem) { KINFree(&R->kmem); }
  if (R->LS) { SUNLinSolFree(R->LS); }
  if (R->J) { SUNMatDestroy(R->J); }
  N_VDestroy(R->u);
  N_VDestroy(R->s);
  N_VDestroy(R->c);
  if (R->ctx) { SUNContext_Free(&R->ctx); }
}

static int hn_setup(const SUNAMD_KinNewtonOptions* o, KinNHost* H, KinNRun* R)
{
  const long n = 6 * o->ncells;
  if (SUNContext_Create(SUN_COMM_NULL, &R->ctx) != 0) { return -1; }
  R->u = N_VNew_Serial(n, R->ctx);
  R->s = N_VNew_Serial(n, R->ctx);
  R->c = N_VNew_Serial(n, R->ctx);
  if (!R->u || !R->s || !R->c) { return -1; }
  N_VConst(1.0, R->s);
  double* cd = N_VGetArrayPointer(R->c);
  for (long c = 0; c < o->ncells; c++) { kft_constraints(cd + 6 * c); }
  R->kmem = KINCreate(R->ctx);
  if (!R->kmem) { return -1; }
  int r = KINSetUserData(R->kmem, H);
  r |= KINSetConstraints(R->kmem, R->c);
  r |= KINSetFuncNormTol(R->kmem, o->fnormtol);
  r |= KINSetScaledStepTol(R->kmem, o->scsteptol);
  r |= KINSetNumMaxIters(R->kmem, o->maxiter);
  r |= KINInit(R->kmem, hn_f, R->u);
  if (o->jac == 0)
  {
    R->J  = SUNDenseMatrix(n, n, R->ctx);
    R->LS = R->J ? SUNLinSol_Dense(R->u, R->J, R->ctx) : NULL;
  }
  else
  {
    R->J  = SUNBlockDenseMatrix_Oracle(o->ncells, 6, R->ctx);
    R->LS = R->J ? SUNLinSol_BlockDense_Oracle(R->u, R->J, R->ctx) : NULL;
  }
  if (!R->LS) { return -1; }
  r |= KINSetLinearSolver(R->kmem, R->LS, R->J);
  if (o->jac != 0) { r |= KINSetJacFn(R->kmem, hn_jac); }
  return r == 0 ? 0 : -1;
}

int oracle_kin_newton_solve(const SUNAMD_KinNewtonOptions* o, double* st, double* uout)
{
  if (!o || !st || o->ncells < 1 || (o->variant != 0 && o->variant != 1)) { return -100; }
  KinNHost H = {o};
  KinNRun R  = {0};
  int rc     = -100;
  if (hn_setup(o, &H, &R) == 0)
  {
    double* u = N_VGetArrayPointer(R.u);
    for (long c = 0; c < o->ncells; c++) { kft_u0(c, o->guess, o->perturb, u + 6 * c); }
    int r = KINSetMaxSetupCalls(R.kmem, o->msbset);
    if (r == 0)
    {
      struct timespec t0, t1;
      clock_gettime(CLOCK_MONOTONIC, &t0);
      rc = KINSol(R.kmem, R.u, o->strategy, R.s, R.s);
      clock_gettime(CLOCK_MONOTONIC, &t1);
      hn_stats(R.kmem, rc, st);
      st[7] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
      if (uout)
      {
        for (long i = 0; i < 6 * o->ncells; i++) { uout[i] = u[i]; }
      }
    }
  }
  hn_free(&R);
  return rc;
}

/* The reference example's whole run (kinFerTron_dns.c:150-260): one KINSOL
   memory, guesses 1 and 2, each solved by exact Newton, exact Newton with
   line search, modified Newton and modified Newton with line search, in that
   order (options set by one solve persist into the next, as in the example).
   x: 8 x 2 solutions; st: 8 x SUNAMD_KINN_STATS. */
int oracle_kin_fertron_example(double* x, double* st)
{
  SUNAMD_KinNewtonOptions o = {1, 0, 1, 0, KIN_NONE, 1, 1e-5, 1e-5, 0, 0, 0};
  KinNHost H                = {&o};
  KinNRun R                 = {0};
  int rc                    = 0;
  if (hn_setup(&o, &H, &R) != 0) { rc = -100; }
  const int glstr[4] = {KIN_NONE, KIN_LINESEARCH, KIN_NONE, KIN_LINESEARCH};
  const long mset[4] = {1, 1, 0, 0};
  for (int g = 0; g < 2 && rc == 0; g++)
  {
    for (int v = 0; v < 4 && rc == 0; v++)
    {
      const int run = 4 * g + v;
      double* u     = N_VGetArrayPointer(R.u);
      kft_u0(0, g + 1, 0, u);
      if (KINSetMaxSetupCalls(R.kmem, mset[v]) != 0)
      {
        rc = -100;
        break;
      }
      const int f = KINSol(R.kmem, R.u, glstr[v], R.s, R.s);
      hn_stats(R.kmem, f, st + SUNAMD_KINN_STATS * run);
      st[SUNAMD_KINN_STATS * run + 7] = 0.0;
      x[2 * run]                      = u[0];
      x[2 * run + 1]                  = u[1];
      if (f < 0) { rc = f; }
    }
  }
  hn_free(&R);
  return rc;
}

/* one cell's function and analytic Jacobian block (checks of the cell
   arithmetic against the math library and finite differences) */
void oracle_kft_f(int variant, const double* u, double* f) { kft_f(variant, u, f); }
void oracle_kft_jac(int variant, const double* u, double* J) { kft_jac(variant, u, J); }
