/* -----------------------------------------------------------------------------
 * CPU ORACLE (test infrastructure): problem callbacks and integration drivers
 * that run the integrator on the serial vector + dense / block-dense LU.
 *
 *  - Robertson kinetics driver with the setup of the reference example
 *    examples/cvode/serial/cvRoberts_dns.c (y0 = (1,0,0), rtol 1e-4,
 *    atol (1e-8, 1e-14, 1e-6), outputs at 0.4 * 10^k, k < 12, dense LU with an
 *    analytic Jacobian); rootfinding is not used, which does not change the
 *    step sequence.  Its counters are checked against the example's
 *    published .out / _stats.csv.
 *  - mass-action drivers: all cells as one coupled system (serial vector,
 *    block-dense LU), and every cell as its own independent integration
 *    (the per-system form), optionally spread over OpenMP threads.
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "sunamd/cvode_sunamd.h"
#include "sunamd/batched_bdf.h"
#include "sunamd/tridiag_mi355.h"
#include "sunamd/massaction.h"

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- Robertson ------------------------------------------------------------ */
static int rob_f(sunrealtype t, N_Vector y, N_Vector ydot, void* ud)
{
  (void)t;
  (void)ud;
  sunrealtype* yv = N_VGetArrayPointer(y);
  sunrealtype* dv = N_VGetArrayPointer(ydot);
  sunrealtype y1 = yv[0], y2 = yv[1], y3 = yv[2];
  sunrealtype yd1 = -0.04 * y1 + 1.0e4 * y2 * y3;
  sunrealtype yd3 = 3.0e7 * y2 * y2;
  dv[0]           = yd1;
  dv[2]           = yd3;
  dv[1]           = -yd1 - yd3;
  return 0;
}

static int rob_jac(sunrealtype t, N_Vector y, N_Vector fy, SUNMatrix J, void* ud,
                   N_Vector t1, N_Vector t2, N_Vector t3)
{
  (void)t;
  (void)fy;
  (void)ud;
  (void)t1;
  (void)t2;
  (void)t3;
  sunrealtype* yv  = N_VGetArrayPointer(y);
  sunrealtype** c  = SUNDenseMatrix_Cols(J);
  sunrealtype y2 = yv[1], y3 = yv[2];
  c[0][0] = -0.04;
  c[1][0] = 1.0e4 * y3;
  c[2][0] = 1.0e4 * y2;
  c[0][1] = 0.04;
  c[1][1] = -1.0e4 * y3 - 6.0e7 * y2;
  c[2][1] = -1.0e4 * y2;
  c[0][2] = 0.0;
  c[1][2] = 6.0e7 * y2;
  c[2][2] = 0.0;
  return 0;
}

/* per output k: out[k*20 + ...] = t, y1, y2, y3, tcur, nst, netf, ncfn, h0,
   hlast, hcur, qlast, qcur, nfe, nni, nnf, nsetups, nje, flag, 0 */
int oracle_roberts_run(int nout, double* out)
{
  SUNContext ctx;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) { return -1; }
  N_Vector y      = N_VNew_Serial(3, ctx);
  N_Vector abstol = N_VNew_Serial(3, ctx);
  sunrealtype* yv = N_VGetArrayPointer(y);
  sunrealtype* av = N_VGetArrayPointer(abstol);
  yv[0] = 1.0;
  yv[1] = 0.0;
  yv[2] = 0.0;
  av[0] = 1.0e-8;
  av[1] = 1.0e-14;
  av[2] = 1.0e-6;
  void* cv = CVodeCreate(CV_BDF, ctx);
  int rc   = CVodeInit(cv, rob_f, 0.0, y);
  rc |= CVodeSVtolerances(cv, 1.0e-4, abstol);
  SUNMatrix A       = SUNDenseMatrix(3, 3, ctx);
  SUNLinearSolver L = SUNLinSol_Dense(y, A, ctx);
  rc |= CVodeSetLinearSolver(cv, L, A);
  rc |= CVodeSetJacFn(cv, rob_jac);
  if (rc) { return -2; }
  sunrealtype tout = 0.4, t = 0.0;
  for (int k = 0; k < nout; k++)
  {
    int flag = CVode(cv, tout, y, &t, CV_NORMAL);
    double* o = out + 20 * k;
    long counts[8];
    int orders[2];
    double steps[4];
    CVodeGetAllStats(cv, counts, orders, steps);
    o[0]  = t;
    o[1]  = yv[0];
    o[2]  = yv[1];
    o[3]  = yv[2];
    o[4]  = steps[3];
    o[5]  = (double)counts[0];
    o[6]  = (double)counts[3];
    o[7]  = (double)counts[6];
    o[8]  = steps[0];
    o[9]  = steps[1];
    o[10] = steps[2];
    o[11] = orders[0];
    o[12] = orders[1];
    o[13] = (double)counts[1];
    o[14] = (double)counts[4];
    o[15] = (double)counts[5];
    o[16] = (double)counts[2];
    o[17] = (double)counts[7];
    o[18] = flag;
    o[19] = 0.0;
    if (flag < 0) { break; }
    tout *= 10.0;
  }
  CVodeFree(&cv);
  SUNLinSolFree(L);
  SUNMatDestroy(A);
  N_VDestroy(y);
  N_VDestroy(abstol);
  SUNContext_Free(&ctx);
  return 0;
}

/* ---- mass action ---------------------------------------------------------- */
static int ma_f(sunrealtype t, N_Vector y, N_Vector ydot, void* ud)
{
  (void)t;
  const SUNAMD_MassAction* p = (const SUNAMD_MassAction*)ud;
  SUNAMD_MassActionRhsHost(p, N_VGetArrayPointer(y), N_VGetArrayPointer(ydot), 0, p->ncells);
  return 0;
}

static int ma_jac(sunrealtype t, N_Vector y, N_Vector fy, SUNMatrix J, void* ud,
                  N_Vector t1, N_Vector t2, N_Vector t3)
{
  (void)t;
  (void)fy;
  (void)t1;
  (void)t2;
  (void)t3;
  const SUNAMD_MassAction* p = (const SUNAMD_MassAction*)ud;
  SUNAMD_MassActionJacHost(p, N_VGetArrayPointer(y), SUNBlockDenseMatrix_Oracle_Data(J), 0,
                           p->ncells);
  return 0;
}

static void fill_stats(void* cv, double* st)
{
  long counts[8];
  int orders[2];
  double steps[4];
  CVodeGetAllStats(cv, counts, orders, steps);
  for (int i = 0; i < 8; i++) { st[i] = (double)counts[i]; }
  st[8]  = orders[0];
  st[9]  = orders[1];
  st[10] = steps[0];
  st[11] = steps[1];
  st[12] = steps[2];
  st[13] = steps[3];
}

/* stats layout (16 doubles): nst nfe nsetups netf nni nnf ncfn nje qlast qcur
   h0 hlast hcur tcur flag wall_s */
int oracle_massaction_solve_coupled(const SUNAMD_MassAction* p, double rtol, double atol,
                                    double tout, double* yout, double* stats)
{
  SUNContext ctx;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) { return -1; }
  sunindextype N = (sunindextype)p->ncells * p->m;
  N_Vector y     = N_VNew_Serial(N, ctx);
  memcpy(N_VGetArrayPointer(y), p->y0, sizeof(double) * (size_t)N);
  SUNMatrix A       = SUNBlockDenseMatrix_Oracle(p->ncells, p->m, ctx);
  SUNLinearSolver L = SUNLinSol_BlockDense_Oracle(y, A, ctx);
  void* cv          = CVodeCreate(CV_BDF, ctx);
  int rc            = CVodeInit(cv, ma_f, 0.0, y);
  rc |= CVodeSStolerances(cv, rtol, atol);
  rc |= CVodeSetUserData(cv, (void*)p);
  rc |= CVodeSetLinearSolver(cv, L, A);
  rc |= CVodeSetJacFn(cv, ma_jac);
  rc |= CVodeSetMaxNumSteps(cv, 1000000);
  if (rc) { return -2; }
  sunrealtype t = 0.0;
  double t0     = now_s();
  int flag      = CVode(cv, tout, y, &t, CV_NORMAL);
  double t1     = now_s();
  if (stats)
  {
    fill_stats(cv, stats);
    stats[14] = flag;
    stats[15] = t1 - t0;
  }
  if (yout) { memcpy(yout, N_VGetArrayPointer(y), sizeof(double) * (size_t)N); }
  CVodeFree(&cv);
  SUNLinSolFree(L);
  SUNMatDestroy(A);
  N_VDestroy(y);
  SUNContext_Free(&ctx);
  return flag < 0 ? flag : 0;
}

/* every cell integrated independently; stats is ncells x 16 (layout above,
   wall_s per cell); returns total wall time in *wall (may be NULL) */
int oracle_massaction_solve_per_cell(const SUNAMD_MassAction* p, double rtol, double atol,
                                     double tout, int nthreads, double* yout, double* stats,
                                     double* wall)
{
  const int m = p->m, R = p->R;
  int nfail   = 0;
  double t0   = now_s();
#ifdef _OPENMP
  if (nthreads < 1) { nthreads = 1; }
#pragma omp parallel num_threads(nthreads) reduction(+ : nfail)
#endif
  {
    SUNContext ctx;
    SUNContext_Create(SUN_COMM_NULL, &ctx);
    N_Vector y        = N_VNew_Serial(m, ctx);
    SUNMatrix A       = SUNBlockDenseMatrix_Oracle(1, m, ctx);
    SUNLinearSolver L = SUNLinSol_BlockDense_Oracle(y, A, ctx);
    void* cv          = CVodeCreate(CV_BDF, ctx);
    SUNAMD_MassAction view = *p;
    view.ncells            = 1;
    int rc = CVodeInit(cv, ma_f, 0.0, y);
    rc |= CVodeSStolerances(cv, rtol, atol);
    rc |= CVodeSetUserData(cv, &view);
    rc |= CVodeSetLinearSolver(cv, L, A);
    rc |= CVodeSetJacFn(cv, ma_jac);
    rc |= CVodeSetMaxNumSteps(cv, 1000000);
    if (rc) { nfail++; }
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int64_t c = 0; c < p->ncells; c++)
    {
      view.kf = p->kf + c * R;
      view.kb = p->kb + c * R;
      memcpy(N_VGetArrayPointer(y), p->y0 + c * m, sizeof(double) * (size_t)m);
      CVodeReInit(cv, 0.0, y);
      sunrealtype t = 0.0;
      double c0     = now_s();
      int flag      = CVode(cv, tout, y, &t, CV_NORMAL);
      double c1     = now_s();
      if (flag < 0) { nfail++; }
      if (stats)
      {
        fill_stats(cv, stats + 16 * c);
        stats[16 * c + 14] = flag;
        stats[16 * c + 15] = c1 - c0;
      }
      if (yout) { memcpy(yout + c * m, N_VGetArrayPointer(y), sizeof(double) * (size_t)m); }
    }
    CVodeFree(&cv);
    SUNLinSolFree(L);
    SUNMatDestroy(A);
    N_VDestroy(y);
    SUNContext_Free(&ctx);
  }
  if (wall) { *wall = now_s() - t0; }
  return nfail ? -1 : 0;
}

/* ---- uneven batch of independent systems ----------------------------------
   Each bucket's systems go through oracle_massaction_solve_per_cell (one
   integration per system, its own CVodeReInit) and the results are scattered
   back to system order: y has batch->ytotal entries, counters 10 per system
   (nst nfe nsetups netf nni nnf ncfn nje qlast flag). */
int oracle_batch_solve(const SUNAMD_Batch* batch, double rtol, double atol, double tout,
                       int nthreads, double* y, long int* counters, double* wall)
{
  double t0 = now_s();
  int rc    = 0;
  for (int b = 0; b < batch->nbuckets; b++)
  {
    const SUNAMD_MassAction* p = batch->bucket[b];
    if (p->ncells <= 0) { continue; }
    double* yb = (double*)malloc(sizeof(double) * (size_t)(p->ncells * p->m));
    double* sb = (double*)malloc(sizeof(double) * 16 * (size_t)p->ncells);
    if (!yb || !sb)
    {
      free(yb);
      free(sb);
      return -1;
    }
    if (oracle_massaction_solve_per_cell(p, rtol, atol, tout, nthreads, yb, sb, NULL) != 0)
    {
      rc = -1;
    }
    for (int64_t s = 0; s < batch->nsys; s++)
    {
      if (batch->sys_bucket[s] != b) { continue; }
      int64_t c = batch->sys_cell[s];
      if (y) { memcpy(y + batch->sys_yoff[s], yb + c * p->m, sizeof(double) * (size_t)p->m); }
      if (counters)
      {
        long int* o = counters + 10 * s;
        for (int k = 0; k < 8; k++) { o[k] = (long int)sb[16 * c + k]; }
        o[8] = (long int)sb[16 * c + 8];
        o[9] = (long int)sb[16 * c + 14];
      }
    }
    free(yb);
    free(sb);
  }
  if (wall) { *wall = now_s() - t0; }
  return rc;
}

/* ---- SPGMR test problem (test_sunlinsol_spgmr_serial.c:433-503) ------------
   A = S1^{-1} tridiag(-1, 5, -1) S2 and a Jacobi preconditioner, on serial
   vectors, in the test's own operation order. */
int oracle_tridiag_atimes(void* data, N_Vector v_vec, N_Vector z_vec)
{
  const SUNAMD_TridiagData* P = (const SUNAMD_TridiagData*)data;
  const sunrealtype* v        = N_VGetArrayPointer(v_vec);
  sunrealtype* z              = N_VGetArrayPointer(z_vec);
  const sunrealtype* s1       = N_VGetArrayPointer(P->s1);
  const sunrealtype* s2       = N_VGetArrayPointer(P->s2);
  const sunindextype N        = P->N;
  if (N < 2) { return -1; }
  z[0] = (5.0 * v[0] * s2[0] - v[1] * s2[1]) / s1[0];
  for (sunindextype i = 1; i < N - 1; i++)
  {
    z[i] = (-v[i - 1] * s2[i - 1] + 5.0 * v[i] * s2[i] - v[i + 1] * s2[i + 1]) / s1[i];
  }
  z[N - 1] = (-v[N - 2] * s2[N - 2] + 5.0 * v[N - 1] * s2[N - 1]) / s1[N - 1];
  return 0;
}

int oracle_tridiag_psetup(void* data)
{
  (void)data;
  return 0;
}

int oracle_tridiag_psolve(void* data, N_Vector r_vec, N_Vector z_vec, sunrealtype tol, int lr)
{
  (void)tol;
  (void)lr;
  const SUNAMD_TridiagData* P = (const SUNAMD_TridiagData*)data;
  const sunrealtype* r        = N_VGetArrayPointer(r_vec);
  sunrealtype* z              = N_VGetArrayPointer(z_vec);
  const sunrealtype* d        = N_VGetArrayPointer(P->d);
  for (sunindextype i = 0; i < P->N; i++) { z[i] = r[i] / d[i]; }
  return 0;
}

/* cvRoberts_dns's statistics output (examples/cvode/serial/cvRoberts_dns.c:
   198-229): after every CVode return CVodePrintAllStats in CSV format to
   csv_path (one line per return: the example's file, whose lines the
   driver separates), and at the end the TABLE format to table_path */
int oracle_roberts_stats(int nout, const char* csv_path, const char* table_path)
{
  SUNContext ctx;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) { return -1; }
  N_Vector y      = N_VNew_Serial(3, ctx);
  N_Vector abstol = N_VNew_Serial(3, ctx);
  sunrealtype* yv = N_VGetArrayPointer(y);
  sunrealtype* av = N_VGetArrayPointer(abstol);
  yv[0] = 1.0;
  yv[1] = 0.0;
  yv[2] = 0.0;
  av[0] = 1.0e-8;
  av[1] = 1.0e-14;
  av[2] = 1.0e-6;
  void* cv = CVodeCreate(CV_BDF, ctx);
  int rc   = CVodeInit(cv, rob_f, 0.0, y);
  rc |= CVodeSVtolerances(cv, 1.0e-4, abstol);
  SUNMatrix A       = SUNDenseMatrix(3, 3, ctx);
  SUNLinearSolver L = SUNLinSol_Dense(y, A, ctx);
  rc |= CVodeSetLinearSolver(cv, L, A);
  rc |= CVodeSetJacFn(cv, rob_jac);
  FILE* csv = fopen(csv_path, "w");
  if (rc || !csv) { rc = -2; }
  sunrealtype tout = 0.4, t = 0.0;
  for (int k = 0; k < nout && rc == 0; k++)
  {
    int flag = CVode(cv, tout, y, &t, CV_NORMAL);
    if (flag < 0) { rc = flag; }
    if (CVodePrintAllStats(cv, csv, SUN_OUTPUTFORMAT_CSV) != CV_SUCCESS) { rc = -3; }
    fprintf(csv, "\n");
    tout *= 10.0;
  }
  if (csv) { fclose(csv); }
  FILE* tab = fopen(table_path, "w");
  if (!tab || CVodePrintAllStats(cv, tab, SUN_OUTPUTFORMAT_TABLE) != CV_SUCCESS) { rc = -4; }
  if (tab) { fclose(tab); }
  CVodeFree(&cv);
  SUNLinSolFree(L);
  SUNMatDestroy(A);
  N_VDestroy(y);
  N_VDestroy(abstol);
  SUNContext_Free(&ctx);
  return rc;
}
