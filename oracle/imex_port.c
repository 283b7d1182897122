/* -----------------------------------------------------------------------------
 * CPU ORACLE (test infrastructure): the IMEX advection-diffusion-reaction
 * problem (include/sunamd/imex_adr.h) on the serial vector -- fe, fi, J_fi v
 * and the per-cell block-Jacobi preconditioner with the cell arithmetic of
 * sundials_amd/csrc/problems/adr_cell.h -- integrated by the ARKStep
 * restatement (core/arkode_ark.c) with the SPGMR restatement, the CPU twin
 * of SUNAMD_AdrSolve_Mi355.
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "adr_cell.h"
#include "oracle.h"
#include "sunamd/arkode_sunamd.h"
#include "sunamd/imex_adr.h"
#include "sunamd/sunlinsol_spgmr.h"

typedef struct
{
  const SUNAMD_Adr* p;
  double ax, ay, az, dd;
  double* J;   /* 9 ncells */
  double* LU;  /* 9 ncells */
  int* perm;   /* 3 ncells */
} AdrHost;

static int h_fe(sunrealtype t, N_Vector yv, N_Vector fv, void* ud)
{
  (void)t;
  const AdrHost* H = (const AdrHost*)ud;
  const int n      = H->p->n;
  const long nn    = (long)n * n;
  const double* y  = N_VGetArrayPointer(yv);
  double* f        = N_VGetArrayPointer(fv);
  for (long c = 0; c < (long)H->p->ncells; c++)
  {
    const int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / nn);
    const long xm = c + (i == 0 ? n - 1 : -1), xp = c + (i == n - 1 ? 1 - n : 1);
    const long ym = c + (j == 0 ? (long)(n - 1) * n : -(long)n);
    const long yp = c + (j == n - 1 ? -(long)(n - 1) * n : (long)n);
    const long zm = c + (k == 0 ? (long)(n - 1) * nn : -nn);
    const long zp = c + (k == n - 1 ? -(long)(n - 1) * nn : nn);
    for (int s = 0; s < 3; s++)
    {
      f[3 * c + s] = adr_fe1(y[3 * c + s], y[3 * xm + s], y[3 * xp + s], y[3 * ym + s],
                             y[3 * yp + s], y[3 * zm + s], y[3 * zp + s], H->ax, H->ay, H->az,
                             H->dd);
    }
  }
  return 0;
}

static int h_fi(sunrealtype t, N_Vector yv, N_Vector fv, void* ud)
{
  (void)t;
  const AdrHost* H = (const AdrHost*)ud;
  const double* y  = N_VGetArrayPointer(yv);
  double* f        = N_VGetArrayPointer(fv);
  for (long c = 0; c < (long)H->p->ncells; c++)
  {
    adr_fi(y + 3 * c, H->p->A, H->p->B, H->p->kappa[c], f + 3 * c);
  }
  return 0;
}

static int h_jv(N_Vector vv, N_Vector Jvv, sunrealtype t, N_Vector yv, N_Vector fy, void* ud,
                N_Vector tmp)
{
  (void)t;
  (void)fy;
  (void)tmp;
  const AdrHost* H = (const AdrHost*)ud;
  const double* y  = N_VGetArrayPointer(yv);
  const double* v  = N_VGetArrayPointer(vv);
  double* r        = N_VGetArrayPointer(Jvv);
  double J[9];
  for (long c = 0; c < (long)H->p->ncells; c++)
  {
    adr_jac(y + 3 * c, H->p->kappa[c], J);
    adr_jv(J, v + 3 * c, r + 3 * c);
  }
  return 0;
}

static int h_psetup(sunrealtype t, N_Vector yv, N_Vector fy, sunbooleantype jok,
                    sunbooleantype* jcurPtr, sunrealtype gamma, void* ud)
{
  (void)t;
  (void)fy;
  AdrHost* H      = (AdrHost*)ud;
  const double* y = N_VGetArrayPointer(yv);
  int fail        = 0;
  *jcurPtr        = jok ? SUNFALSE : SUNTRUE;
  for (long c = 0; c < (long)H->p->ncells; c++)
  {
    if (!jok) { adr_jac(y + 3 * c, H->p->kappa[c], H->J + 9 * c); }
    if (adr_plu(H->J + 9 * c, gamma, H->LU + 9 * c, H->perm + 3 * c)) { fail = 1; }
  }
  return fail ? 1 : 0;
}

static int h_psolve(sunrealtype t, N_Vector yv, N_Vector fy, N_Vector rv, N_Vector zv,
                    sunrealtype gamma, sunrealtype delta, int lr, void* ud)
{
  (void)t;
  (void)yv;
  (void)fy;
  (void)gamma;
  (void)delta;
  (void)lr;
  const AdrHost* H = (const AdrHost*)ud;
  const double* r  = N_VGetArrayPointer(rv);
  double* z        = N_VGetArrayPointer(zv);
  for (long c = 0; c < (long)H->p->ncells; c++)
  {
    adr_psolve(H->LU + 9 * c, H->perm + 3 * c, r + 3 * c, z + 3 * c);
  }
  return 0;
}

int oracle_adr_solve(const SUNAMD_Adr* p, const SUNAMD_AdrOptions* o, double* stats,
                     double* yout)
{
  if (p == NULL || o == NULL || stats == NULL) { return -1; }
  const long nc = (long)p->ncells;
  AdrHost H;
  const double dx = 1.0 / p->n;
  H.p             = p;
  H.ax            = p->cx / dx;
  H.ay            = p->cy / dx;
  H.az            = p->cz / dx;
  H.dd            = p->d / (dx * dx);
  H.J             = (double*)malloc(sizeof(double) * 9 * (size_t)nc);
  H.LU            = (double*)malloc(sizeof(double) * 9 * (size_t)nc);
  H.perm          = (int*)malloc(sizeof(int) * 3 * (size_t)nc);
  SUNContext ctx  = NULL;
  int rc          = -1;
  if (H.J && H.LU && H.perm && SUNContext_Create(SUN_COMM_NULL, &ctx) == 0)
  {
    N_Vector y = N_VNew_Serial(3 * nc, ctx);
    memcpy(N_VGetArrayPointer(y), p->y0, sizeof(double) * 3 * (size_t)nc);
    void* ark         = ARKStepCreate(h_fe, h_fi, 0.0, y, ctx);
    SUNLinearSolver L = SUNLinSol_SPGMR(y, SUN_PREC_LEFT, o->maxl, ctx);
    int ok            = (ark != NULL) && (L != NULL);
    if (ok && o->gstype == SUN_CLASSICAL_GS) { ok = SUNLinSol_SPGMRSetGSType(L, SUN_CLASSICAL_GS) == 0; }
    if (ok)
    {
      const char* cname = (o->controller == 1) ? "PI" : (o->controller == 2) ? "PID" : "I";
      ok = ARKodeSStolerances(ark, o->rtol, o->atol) == 0 && ARKodeSetUserData(ark, &H) == 0 &&
           ARKodeSetOrder(ark, 3) == 0 && ARKodeSetLinearSolver(ark, L, NULL) == 0 &&
           ARKodeSetPreconditioner(ark, h_psetup, h_psolve) == 0 &&
           ARKodeSetJacTimes(ark, NULL, h_jv) == 0 && ARKodeSetStopTime(ark, o->tf) == 0 &&
           ARKodeSetMaxNumSteps(ark, o->mxsteps) == 0 &&
           ARKodeSetAdaptControllerByName(ark, cname) == 0;
    }
    if (ok)
    {
      sunrealtype tret = 0.0;
      struct timespec t0, t1;
      clock_gettime(CLOCK_MONOTONIC, &t0);
      rc = ARKodeEvolve(ark, o->tf, y, &tret, ARK_NORMAL);
      clock_gettime(CLOCK_MONOTONIC, &t1);
      long v[14] = {0};
      ARKodeGetNumSteps(ark, &v[0]);
      ARKodeGetNumStepAttempts(ark, &v[1]);
      ARKodeGetNumErrTestFails(ark, &v[2]);
      ARKodeGetNumConvFails(ark, &v[3]);
      ARKStepGetNumRhsEvals(ark, &v[4], &v[5]);
      ARKodeGetNumNonlinSolvIters(ark, &v[6]);
      ARKodeGetNumNonlinSolvConvFails(ark, &v[7]);
      ARKodeGetNumLinSolvSetups(ark, &v[8]);
      ARKodeGetNumLinIters(ark, &v[9]);
      ARKodeGetNumPrecEvals(ark, &v[10]);
      ARKodeGetNumPrecSolves(ark, &v[11]);
      ARKodeGetNumJtimesEvals(ark, &v[12]);
      ARKodeGetNumLinConvFails(ark, &v[13]);
      for (int q = 0; q < 14; q++) { stats[q] = (double)v[q]; }
      stats[14] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
      sunrealtype hl = 0.0;
      ARKodeGetLastStep(ark, &hl);
      stats[15] = hl;
      if (rc == ARK_TSTOP_RETURN) { rc = 0; }
      if (yout) { memcpy(yout, N_VGetArrayPointer(y), sizeof(double) * 3 * (size_t)nc); }
    }
    if (ark) { ARKodeFree(&ark); }
    if (L) { SUNLinSolFree(L); }
    N_VDestroy(y);
    SUNContext_Free(&ctx);
  }
  free(H.J);
  free(H.LU);
  free(H.perm);
  return rc;
}

/* ---------------------------------------------------------------------------
 * Dahlquist test problem y' = lam_e y + lam_i y (fe = lam_e y explicit,
 * fi = lam_i y implicit), N = 1, SPGMR without preconditioning: fixed steps
 * (h > 0) or adaptive (h = 0) from y(0) = 1 to tf.  stats as in imex_adr.h.
 * ---------------------------------------------------------------------------*/
typedef struct
{
  double le, li;
  int pmode;   /* user stage predictor: 0 none, 1 keeps zpred, 2 asks for a
                  retry on its first call, 3 fails, 4 replaces zpred by the
                  exact solution at the stage time */
  long pcalls; /* stage-predictor calls */
} Dahl;

static int d_fe(sunrealtype t, N_Vector y, N_Vector f, void* ud)
{
  (void)t;
  N_VGetArrayPointer(f)[0] = ((const Dahl*)ud)->le * N_VGetArrayPointer(y)[0];
  return 0;
}

static int d_fi(sunrealtype t, N_Vector y, N_Vector f, void* ud)
{
  (void)t;
  N_VGetArrayPointer(f)[0] = ((const Dahl*)ud)->li * N_VGetArrayPointer(y)[0];
  return 0;
}

static int d_jv(N_Vector v, N_Vector Jv, sunrealtype t, N_Vector y, N_Vector fy, void* ud,
                N_Vector tmp)
{
  (void)t;
  (void)y;
  (void)fy;
  (void)tmp;
  N_VGetArrayPointer(Jv)[0] = ((const Dahl*)ud)->li * N_VGetArrayPointer(v)[0];
  return 0;
}

static int d_predict(sunrealtype t, N_Vector zpred, void* ud)
{
  Dahl* D = (Dahl*)ud;
  D->pcalls++;
  switch (D->pmode)
  {
  case 2: return (D->pcalls == 1) ? 1 : 0;
  case 3: return -1;
  case 4: N_VGetArrayPointer(zpred)[0] = exp((D->le + D->li) * t); return 0;
  default: return 0;
  }
}

static int dahl_run(double lam_e, double lam_i, double h, double tf, double rtol, double atol,
                    const char* controller, int flags, int order, int pmode, double* stats,
                    double* yout, long* pcalls)
{
  Dahl D         = {lam_e, lam_i, pmode, 0};
  SUNContext ctx = NULL;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) { return -1; }
  N_Vector y                = N_VNew_Serial(1, ctx);
  N_VGetArrayPointer(y)[0]  = 1.0;
  void* ark                 = ARKStepCreate(d_fe, d_fi, 0.0, y, ctx);
  SUNLinearSolver L         = SUNLinSol_SPGMR(y, SUN_PREC_NONE, 1, ctx);
  int rc                    = -1;
  int ok = ark && L && ARKodeSStolerances(ark, rtol, atol) == 0 &&
           ARKodeSetUserData(ark, &D) == 0 && ARKodeSetLinearSolver(ark, L, NULL) == 0 &&
           ARKodeSetJacTimes(ark, NULL, d_jv) == 0 && ARKodeSetStopTime(ark, tf) == 0 &&
           ARKodeSetMaxNumSteps(ark, 1000000) == 0 && ARKodeSetOrder(ark, order) == 0;
  if (ok && h > 0.0) { ok = ARKodeSetFixedStep(ark, h) == 0; }
  if (ok && controller) { ok = ARKodeSetAdaptControllerByName(ark, controller) == 0; }
  if (ok && (flags & 1)) { ok = ARKodeSetLinear(ark, 0) == 0; }
  if (ok && (flags & 2)) { ok = ARKodeSetAutonomous(ark, SUNTRUE) == 0; }
  if (ok && (flags & 4)) { ok = ARKodeSetInterpolantType(ark, ARK_INTERP_NONE) == 0; }
  if (ok && pmode > 0) { ok = ARKodeSetStagePredictFn(ark, d_predict) == 0; }
  if (ok)
  {
    sunrealtype tret = 0.0;
    rc               = ARKodeEvolve(ark, tf, y, &tret, ARK_NORMAL);
    long v[14]       = {0};
    ARKodeGetNumSteps(ark, &v[0]);
    ARKodeGetNumStepAttempts(ark, &v[1]);
    ARKodeGetNumErrTestFails(ark, &v[2]);
    ARKodeGetNumConvFails(ark, &v[3]);
    ARKStepGetNumRhsEvals(ark, &v[4], &v[5]);
    ARKodeGetNumNonlinSolvIters(ark, &v[6]);
    ARKodeGetNumNonlinSolvConvFails(ark, &v[7]);
    ARKodeGetNumLinSolvSetups(ark, &v[8]);
    ARKodeGetNumLinIters(ark, &v[9]);
    for (int q = 0; q < 14; q++) { stats[q] = (double)v[q]; }
    stats[14] = tret;
    sunrealtype hl = 0.0;
    ARKodeGetLastStep(ark, &hl);
    stats[15] = hl;
    if (rc == ARK_TSTOP_RETURN) { rc = 0; }
    yout[0] = N_VGetArrayPointer(y)[0];
  }
  if (pcalls) { *pcalls = D.pcalls; }
  if (ark) { ARKodeFree(&ark); }
  if (L) { SUNLinSolFree(L); }
  N_VDestroy(y);
  SUNContext_Free(&ctx);
  return rc;
}

/* flags: 1 linearly implicit (not time dependent), 2 autonomous, 4 no
   interpolant (else Hermite); order selects the pair (3 or 4) */
int oracle_ark_dahlquist(double lam_e, double lam_i, double h, double tf, double rtol,
                         double atol, const char* controller, int flags, int order,
                         double* stats, double* yout)
{
  return dahl_run(lam_e, lam_i, h, tf, rtol, atol, controller, flags, order, 0, stats, yout,
                  NULL);
}

/* the adaptive Dahlquist run (default controller, Hermite interpolant) with
   a user stage predictor of the given mode (Dahl.pmode); *pcalls receives
   its number of calls */
int oracle_ark_dahlquist_predict(double lam_e, double lam_i, double tf, double rtol,
                                 double atol, int order, int pmode, double* stats, double* yout,
                                 long* pcalls)
{
  return dahl_run(lam_e, lam_i, 0.0, tf, rtol, atol, NULL, 0, order, pmode, stats, yout,
                  pcalls);
}

/* ---------------------------------------------------------------------------
 * Kvaerno-Prothero-Robinson problem (test/unit_tests/problems/kpr.hpp) as the
 * reference's ARKStep logging test drives it (test_logging_arkode_arkstep.cpp,
 * arguments "2 1 0": IMEX split, Newton, SPGMR without preconditioning,
 * maxl default, difference-quotient Jv, rtol 1e-6, atol 1e-10, three calls in
 * ARK_ONE_STEP mode).  u' = a g1 + b g2 + rho'/(2u), v' = c g1 + d g2 +
 * sigma'/(2v) with g1 = (-1 + u^2 - rho)/(2u), g2 = (-2 + v^2 - sigma)/(2v),
 * rho = cos(t)/2, sigma = cos(20 t); (a, b, c, d) = (-2, 1/2, 1/2, -1); the
 * rho'/sigma' terms are explicit, the rest implicit.  Exact solution
 * u = sqrt(1 + rho), v = sqrt(2 + sigma).
 * out[3k..3k+2] = (t, u, v) after k calls (k = 0 the initial state);
 * stats: steps, attempts, error-test fails, NLS step fails, h0, last h,
 * current h, fe, fi, NLS iters, NLS fails, LS setups, LS RHS evals, LS iters,
 * LS fails, Jv evals, preconditioner setups, preconditioner solves.
 * ---------------------------------------------------------------------------*/
static const double KPR_ABCD[4] = {-2.0, 0.5, 0.5, -1.0};

static double kpr_rho(double t) { return 0.5 * cos(t); }
static double kpr_rho_p(double t) { return -0.5 * sin(t); }
static double kpr_sig(double t) { return cos(20.0 * t); }
static double kpr_sig_p(double t) { return -20.0 * sin(20.0 * t); }

static int kpr_fe(sunrealtype t, N_Vector y, N_Vector f, void* ud)
{
  (void)ud;
  const double* Y = N_VGetArrayPointer(y);
  double* F       = N_VGetArrayPointer(f);
  F[0]            = kpr_rho_p(t) / (2.0 * Y[0]);
  F[1]            = kpr_sig_p(t) / (2.0 * Y[1]);
  return 0;
}

static int kpr_fi(sunrealtype t, N_Vector y, N_Vector f, void* ud)
{
  const double* k = (const double*)ud;
  const double* Y = N_VGetArrayPointer(y);
  double* F       = N_VGetArrayPointer(f);
  const double u = Y[0], v = Y[1];
  const double g1 = (-1.0 + u * u - kpr_rho(t)) / (2.0 * u);
  const double g2 = (-2.0 + v * v - kpr_sig(t)) / (2.0 * v);
  F[0]            = k[0] * g1 + k[1] * g2;
  F[1]            = k[2] * g1 + k[3] * g2;
  return 0;
}

int oracle_ark_kpr(int ncalls, double* out, double* stats)
{
  SUNContext ctx = NULL;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) { return -1; }
  N_Vector y = N_VNew_Serial(2, ctx);
  double* Y  = N_VGetArrayPointer(y);
  Y[0]       = sqrt(1.0 + kpr_rho(0.0));
  Y[1]       = sqrt(2.0 + kpr_sig(0.0));
  out[0]     = 0.0;
  out[1]     = Y[0];
  out[2]     = Y[1];
  void* ark         = ARKStepCreate(kpr_fe, kpr_fi, 0.0, y, ctx);
  SUNLinearSolver L = SUNLinSol_SPGMR(y, SUN_PREC_NONE, 0, ctx);
  int rc            = -1;
  int ok = ark && L && ARKodeSetUserData(ark, (void*)KPR_ABCD) == 0 &&
           ARKodeSStolerances(ark, 1.0e-6, 1.0e-10) == 0 &&
           ARKodeSetLinearSolver(ark, L, NULL) == 0;
  if (ok)
  {
    double tret = 0.0, tout = 1.0;
    for (int k = 1; k <= ncalls; k++)
    {
      rc = ARKodeEvolve(ark, tout, y, &tret, ARK_ONE_STEP);
      if (rc < 0) { break; }
      out[3 * k]     = tret;
      out[3 * k + 1] = Y[0];
      out[3 * k + 2] = Y[1];
      tout += 1.0;
    }
    long v[15] = {0};
    double h[3] = {0.0, 0.0, 0.0};
    ARKodeGetNumSteps(ark, &v[0]);
    ARKodeGetNumStepAttempts(ark, &v[1]);
    ARKodeGetNumErrTestFails(ark, &v[2]);
    ARKodeGetNumConvFails(ark, &v[3]);
    ARKodeGetActualInitStep(ark, &h[0]);
    ARKodeGetLastStep(ark, &h[1]);
    ARKodeGetCurrentStep(ark, &h[2]);
    ARKStepGetNumRhsEvals(ark, &v[4], &v[5]);
    ARKodeGetNumNonlinSolvIters(ark, &v[6]);
    ARKodeGetNumNonlinSolvConvFails(ark, &v[7]);
    ARKodeGetNumLinSolvSetups(ark, &v[8]);
    ARKodeGetNumLinRhsEvals(ark, &v[9]);
    ARKodeGetNumLinIters(ark, &v[10]);
    ARKodeGetNumLinConvFails(ark, &v[11]);
    ARKodeGetNumJtimesEvals(ark, &v[12]);
    ARKodeGetNumPrecEvals(ark, &v[13]);
    ARKodeGetNumPrecSolves(ark, &v[14]);
    const double sv[18] = {v[0], v[1], v[2], v[3], h[0], h[1], h[2], v[4], v[5],
                           v[6], v[7], v[8], v[9], v[10], v[11], v[12], v[13], v[14]};
    memcpy(stats, sv, sizeof(sv));
  }
  if (ark) { ARKodeFree(&ark); }
  if (L) { SUNLinSolFree(L); }
  N_VDestroy(y);
  SUNContext_Free(&ctx);
  return rc;
}
