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
