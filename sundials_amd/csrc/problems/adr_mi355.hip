// -----------------------------------------------------------------------------
// Device side of the IMEX advection-diffusion-reaction problem (imex_adr.h):
// ARKStep callbacks fe, fi, J_fi v and the per-cell block-Jacobi
// preconditioner on MI355X vectors, one thread per cell, launched on the
// output vector's stream; and the solve driver (ARKStep IMEX + SPGMR).
// Cell arithmetic comes from adr_cell.h, shared with the CPU twin.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../core/profiler_internal.h"
#include "adr_cell.h"
#include "sunamd/arkode_sunamd.h"
#include "sunamd/imex_adr.h"
#include "sunamd/nvector_mi355.h"
#include "sunamd/sunamd_profiler.h"
#include "sunamd/sundials_core.h"
#include "sunamd/sunlinsol_spgmr.h"

namespace
{

struct AdrDev
{
  int n;
  long ncells;
  double ax, ay, az, dd, A, B;
  double* kappa;  // ncells
  double* J;      // 9 ncells: reaction Jacobian at the last refresh
  double* LU;     // 9 ncells (FP64 factors)
  float* LU32;    // 9 ncells (FP32 factors)
  int* perm;      // 3 ncells
  int* flag;      // device flag: a zero pivot
  int prec_fp32;
};

constexpr int TPB = 256;

inline unsigned blocks(long n) { return (unsigned)((n + TPB - 1) / TPB); }

__global__ void k_fe(AdrDev P, const double* __restrict__ y, double* __restrict__ f)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.ncells) { return; }
  const int n = P.n;
  const int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((long)n * n));
  const long nn = (long)n * n;
  const long xm = c + (i == 0 ? n - 1 : -1), xp = c + (i == n - 1 ? 1 - n : 1);
  const long ym = c + (j == 0 ? (long)(n - 1) * n : -(long)n);
  const long yp = c + (j == n - 1 ? -(long)(n - 1) * n : (long)n);
  const long zm = c + (k == 0 ? (long)(n - 1) * nn : -nn);
  const long zp = c + (k == n - 1 ? -(long)(n - 1) * nn : nn);
#pragma unroll
  for (int s = 0; s < 3; s++)
  {
    f[3 * c + s] = adr_fe1(y[3 * c + s], y[3 * xm + s], y[3 * xp + s], y[3 * ym + s],
                           y[3 * yp + s], y[3 * zm + s], y[3 * zp + s], P.ax, P.ay, P.az, P.dd);
  }
}

__global__ void k_fi(AdrDev P, const double* __restrict__ y, double* __restrict__ f)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.ncells) { return; }
  double yc[3] = {y[3 * c], y[3 * c + 1], y[3 * c + 2]}, fc[3];
  adr_fi(yc, P.A, P.B, P.kappa[c], fc);
  f[3 * c]     = fc[0];
  f[3 * c + 1] = fc[1];
  f[3 * c + 2] = fc[2];
}

__global__ void k_jv(AdrDev P, const double* __restrict__ y, const double* __restrict__ v,
                     double* __restrict__ r)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.ncells) { return; }
  double yc[3] = {y[3 * c], y[3 * c + 1], y[3 * c + 2]};
  double vc[3] = {v[3 * c], v[3 * c + 1], v[3 * c + 2]};
  double J[9], rc[3];
  adr_jac(yc, P.kappa[c], J);
  adr_jv(J, vc, rc);
  r[3 * c]     = rc[0];
  r[3 * c + 1] = rc[1];
  r[3 * c + 2] = rc[2];
}

__global__ void k_psetup(AdrDev P, const double* __restrict__ y, int refresh, double gamma)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.ncells) { return; }
  double J[9];
  if (refresh)
  {
    double yc[3] = {y[3 * c], y[3 * c + 1], y[3 * c + 2]};
    adr_jac(yc, P.kappa[c], J);
#pragma unroll
    for (int q = 0; q < 9; q++) { P.J[9 * c + q] = J[q]; }
  }
  else
  {
#pragma unroll
    for (int q = 0; q < 9; q++) { J[q] = P.J[9 * c + q]; }
  }
  double LU[9];
  int perm[3];
  const int fail = adr_plu(J, gamma, LU, perm);
  if (fail) { atomicOr(P.flag, 1); }
  if (P.prec_fp32)
  {
#pragma unroll
    for (int q = 0; q < 9; q++) { P.LU32[9 * c + q] = (float)LU[q]; }
  }
  else
  {
#pragma unroll
    for (int q = 0; q < 9; q++) { P.LU[9 * c + q] = LU[q]; }
  }
  P.perm[3 * c]     = perm[0];
  P.perm[3 * c + 1] = perm[1];
  P.perm[3 * c + 2] = perm[2];
}

__global__ void k_psolve(AdrDev P, const double* __restrict__ r, double* __restrict__ z)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.ncells) { return; }
  double LU[9];
  if (P.prec_fp32)
  {
#pragma unroll
    for (int q = 0; q < 9; q++) { LU[q] = (double)P.LU32[9 * c + q]; }
  }
  else
  {
#pragma unroll
    for (int q = 0; q < 9; q++) { LU[q] = P.LU[9 * c + q]; }
  }
  const int perm[3] = {P.perm[3 * c], P.perm[3 * c + 1], P.perm[3 * c + 2]};
  double rc[3] = {r[3 * c], r[3 * c + 1], r[3 * c + 2]}, zc[3];
  adr_psolve(LU, perm, rc, zc);
  z[3 * c]     = zc[0];
  z[3 * c + 1] = zc[1];
  z[3 * c + 2] = zc[2];
}

int launch_check(const char* what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
  {
    fprintf(stderr, "[sundials_amd adr] %s: %s\n", what, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int cb_fe(sunrealtype t, N_Vector y, N_Vector f, void* ud)
{
  (void)t;
  const AdrDev* P = (const AdrDev*)ud;
  hipStream_t s   = (hipStream_t)N_VGetStream_Mi355(f);
  sunamd_prof_begin(SUNAMD_K_ADR_FE, s);
  k_fe<<<blocks(P->ncells), TPB, 0, s>>>(*P, N_VGetDeviceArrayPointer(y),
                                         N_VGetDeviceArrayPointer(f));
  sunamd_prof_end(SUNAMD_K_ADR_FE, s);
  return launch_check("fe");
}

int cb_fi(sunrealtype t, N_Vector y, N_Vector f, void* ud)
{
  (void)t;
  const AdrDev* P = (const AdrDev*)ud;
  hipStream_t s   = (hipStream_t)N_VGetStream_Mi355(f);
  k_fi<<<blocks(P->ncells), TPB, 0, s>>>(*P, N_VGetDeviceArrayPointer(y),
                                         N_VGetDeviceArrayPointer(f));
  return launch_check("fi");
}

int cb_jv(N_Vector v, N_Vector Jv, sunrealtype t, N_Vector y, N_Vector fy, void* ud,
          N_Vector tmp)
{
  (void)t;
  (void)fy;
  (void)tmp;
  const AdrDev* P = (const AdrDev*)ud;
  hipStream_t s   = (hipStream_t)N_VGetStream_Mi355(Jv);
  k_jv<<<blocks(P->ncells), TPB, 0, s>>>(*P, N_VGetDeviceArrayPointer(y),
                                         N_VGetDeviceArrayPointer(v),
                                         N_VGetDeviceArrayPointer(Jv));
  return launch_check("jtimes");
}

int cb_psetup(sunrealtype t, N_Vector y, N_Vector fy, sunbooleantype jok,
              sunbooleantype* jcurPtr, sunrealtype gamma, void* ud)
{
  (void)t;
  (void)fy;
  AdrDev* P     = (AdrDev*)ud;
  hipStream_t s = (hipStream_t)N_VGetStream_Mi355(y);
  const int refresh = jok ? 0 : 1;
  *jcurPtr          = refresh ? SUNTRUE : SUNFALSE;
  int zero          = 0;
  if (hipMemcpyAsync(P->flag, &zero, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
  {
    return -1;
  }
  k_psetup<<<blocks(P->ncells), TPB, 0, s>>>(*P, N_VGetDeviceArrayPointer(y), refresh, gamma);
  if (launch_check("psetup")) { return -1; }
  int fail = 0;
  if (hipMemcpyAsync(&fail, P->flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
  {
    return -1;
  }
  return fail ? 1 : 0;
}

int cb_psolve(sunrealtype t, N_Vector y, N_Vector fy, N_Vector r, N_Vector z, sunrealtype gamma,
              sunrealtype delta, int lr, void* ud)
{
  (void)t;
  (void)y;
  (void)fy;
  (void)gamma;
  (void)delta;
  (void)lr;
  const AdrDev* P = (const AdrDev*)ud;
  hipStream_t s   = (hipStream_t)N_VGetStream_Mi355(z);
  k_psolve<<<blocks(P->ncells), TPB, 0, s>>>(*P, N_VGetDeviceArrayPointer(r),
                                             N_VGetDeviceArrayPointer(z));
  return launch_check("psolve");
}

void dev_free(AdrDev& P)
{
  if (P.kappa) { (void)hipFree(P.kappa); }
  if (P.J) { (void)hipFree(P.J); }
  if (P.LU) { (void)hipFree(P.LU); }
  if (P.LU32) { (void)hipFree(P.LU32); }
  if (P.perm) { (void)hipFree(P.perm); }
  if (P.flag) { (void)hipFree(P.flag); }
  memset(&P, 0, sizeof(P));
}

} // namespace

// integrate the problem over [0, tf] with ARKStep IMEX (ARK3(2)4L[2]SA),
// SPGMR (left block-Jacobi preconditioning, no restarts) on MI355X vectors;
// stats as in imex_adr.h, yout (3 ncells, host) the final state
extern "C" int SUNAMD_AdrSolve_Mi355(const SUNAMD_Adr* p, const SUNAMD_AdrOptions* o,
                                     double* stats, double* yout)
{
  if (p == NULL || o == NULL || stats == NULL) { return -1; }
  AdrDev P;
  memset(&P, 0, sizeof(P));
  P.n         = p->n;
  P.ncells    = (long)p->ncells;
  const double dx = 1.0 / p->n;
  P.ax        = p->cx / dx;
  P.ay        = p->cy / dx;
  P.az        = p->cz / dx;
  P.dd        = p->d / (dx * dx);
  P.A         = p->A;
  P.B         = p->B;
  P.prec_fp32 = o->prec_fp32;
  const size_t nc = (size_t)P.ncells;
  bool ok = hipMalloc(&P.kappa, sizeof(double) * nc) == hipSuccess &&
            hipMalloc(&P.J, sizeof(double) * 9 * nc) == hipSuccess &&
            hipMalloc(&P.perm, sizeof(int) * 3 * nc) == hipSuccess &&
            hipMalloc(&P.flag, sizeof(int)) == hipSuccess;
  if (ok)
  {
    ok = o->prec_fp32 ? hipMalloc(&P.LU32, sizeof(float) * 9 * nc) == hipSuccess
                      : hipMalloc(&P.LU, sizeof(double) * 9 * nc) == hipSuccess;
  }
  if (ok) { ok = hipMemcpy(P.kappa, p->kappa, sizeof(double) * nc, hipMemcpyHostToDevice) == hipSuccess; }
  SUNContext ctx    = NULL;
  N_Vector y        = NULL;
  SUNLinearSolver L = NULL;
  void* ark         = NULL;
  int rc            = -1;
  if (ok) { ok = SUNContext_Create(SUN_COMM_NULL, &ctx) == 0; }
  if (ok) { ok = (y = N_VNew_Mi355((sunindextype)(3 * nc), ctx)) != NULL; }
  if (ok)
  {
    ok = N_VSetReductionOrder_Mi355(y, o->sequential ? SUNAMD_REDUCE_SEQUENTIAL
                                                     : SUNAMD_REDUCE_TREE) == 0;
  }
  if (ok) { ok = N_VCopyFromHostArray_Mi355(y, p->y0) == 0; }
  if (ok) { ok = (ark = ARKStepCreate(cb_fe, cb_fi, 0.0, y, ctx)) != NULL; }
  if (ok) { ok = (L = SUNLinSol_SPGMR(y, SUN_PREC_LEFT, o->maxl, ctx)) != NULL; }
  if (ok && o->gstype == SUN_CLASSICAL_GS)
  {
    ok = SUNLinSol_SPGMRSetGSType(L, SUN_CLASSICAL_GS) == 0 &&
         SUNLinSol_SPGMRSetClassicalGSFn(L, SUNAMD_ClassicalGS_Mi355) == 0;
  }
  if (ok)
  {
    const char* cname = (o->controller == 1) ? "PI" : (o->controller == 2) ? "PID" : "I";
    ok = ARKodeSStolerances(ark, o->rtol, o->atol) == 0 && ARKodeSetUserData(ark, &P) == 0 &&
         ARKodeSetOrder(ark, 3) == 0 && ARKodeSetLinearSolver(ark, L, NULL) == 0 &&
         ARKodeSetPreconditioner(ark, cb_psetup, cb_psolve) == 0 &&
         ARKodeSetJacTimes(ark, NULL, cb_jv) == 0 && ARKodeSetStopTime(ark, o->tf) == 0 &&
         ARKodeSetMaxNumSteps(ark, o->mxsteps) == 0 &&
         ARKodeSetAdaptControllerByName(ark, cname) == 0;
  }
  if (ok)
  {
    sunrealtype tret = 0.0;
    (void)hipDeviceSynchronize();
    auto t0 = std::chrono::steady_clock::now();
    rc      = ARKodeEvolve(ark, o->tf, y, &tret, ARK_NORMAL);
    (void)N_VSynchronize_Mi355(y);
    auto t1 = std::chrono::steady_clock::now();
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
    stats[14] = std::chrono::duration<double>(t1 - t0).count();
    sunrealtype hl = 0.0;
    ARKodeGetLastStep(ark, &hl);
    stats[15] = hl;
    if (rc == ARK_TSTOP_RETURN) { rc = 0; }
    if (yout != NULL && N_VCopyToHostArray_Mi355(y, yout) != 0) { rc = -1; }
  }
  if (ark) { ARKodeFree(&ark); }
  if (L) { SUNLinSolFree(L); }
  if (y) { N_VDestroy(y); }
  if (ctx) { SUNContext_Free(&ctx); }
  dev_free(P);
  return ok ? rc : -1;
}
