// -----------------------------------------------------------------------------
// Device side of the Newton problem of kin_newton.h: the cell functions and
// the 6 x 6 Jacobian blocks evaluated one thread per cell on the vector's
// stream, and the solve driver (KINSol with KIN_NONE or KIN_LINESEARCH on
// MI355X vectors, Newton systems solved by the batched block LU).  Cell
// arithmetic comes from kin_fertron.h, shared with the CPU twin.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <chrono>

#include "kin_fertron.h"
#include "sunamd/kin_newton.h"
#include "sunamd/kinsol_sunamd.h"
#include "sunamd/nvector_mi355.h"
#include "sunamd/sundials_core.h"
#include "sunamd/sunmatrix_mi355block.h"

namespace
{

constexpr int TPB = 256;

inline unsigned blocks(long n) { return (unsigned)((n + TPB - 1) / TPB); }

__global__ void k_kft_f(const double* __restrict__ u, double* __restrict__ f, long ncells)
{
  const long c = (long)blockIdx.x * TPB + threadIdx.x;
  if (c >= ncells) { return; }
  double uc[KFT_NEQ], fc[KFT_NEQ];
  for (int i = 0; i < KFT_NEQ; i++) { uc[i] = u[KFT_NEQ * c + i]; }
  kft_f(1, uc, fc);
  for (int i = 0; i < KFT_NEQ; i++) { f[KFT_NEQ * c + i] = fc[i]; }
}

__global__ void k_kft_jac(const double* __restrict__ u, double* __restrict__ J, long ncells)
{
  const long c = (long)blockIdx.x * TPB + threadIdx.x;
  if (c >= ncells) { return; }
  double uc[KFT_NEQ], Jc[KFT_NEQ * KFT_NEQ];
  for (int i = 0; i < KFT_NEQ; i++) { uc[i] = u[KFT_NEQ * c + i]; }
  kft_jac(1, uc, Jc);
  for (int i = 0; i < KFT_NEQ * KFT_NEQ; i++) { J[KFT_NEQ * KFT_NEQ * c + i] = Jc[i]; }
}

__global__ void k_kft_init(int guess, int perturb, double* __restrict__ u,
                           double* __restrict__ cons, long ncells)
{
  const long c = (long)blockIdx.x * TPB + threadIdx.x;
  if (c >= ncells) { return; }
  double uc[KFT_NEQ], cc[KFT_NEQ];
  kft_u0(c, guess, perturb, uc);
  kft_constraints(cc);
  for (int i = 0; i < KFT_NEQ; i++)
  {
    u[KFT_NEQ * c + i]    = uc[i];
    cons[KFT_NEQ * c + i] = cc[i];
  }
}

struct KinNDev
{
  long ncells;
};

int cb_f(N_Vector u, N_Vector f, void* ud)
{
  const KinNDev* D = (const KinNDev*)ud;
  hipStream_t s    = (hipStream_t)N_VGetStream_Mi355(f);
  k_kft_f<<<blocks(D->ncells), TPB, 0, s>>>(N_VGetDeviceArrayPointer(u),
                                            N_VGetDeviceArrayPointer(f), D->ncells);
  return hipPeekAtLastError() == hipSuccess ? 0 : -1;
}

int cb_jac(N_Vector u, N_Vector fu, SUNMatrix J, void* ud, N_Vector t1, N_Vector t2)
{
  (void)fu; (void)t1; (void)t2;
  const KinNDev* D = (const KinNDev*)ud;
  if (SUNMatrix_Mi355BlockDense_NumBlocks(J) != D->ncells ||
      SUNMatrix_Mi355BlockDense_BlockSize(J) != KFT_NEQ)
  {
    return -1;
  }
  hipStream_t s = (hipStream_t)N_VGetStream_Mi355(u);
  k_kft_jac<<<blocks(D->ncells), TPB, 0, s>>>(N_VGetDeviceArrayPointer(u),
                                              SUNMatrix_Mi355BlockDense_DeviceData(J),
                                              D->ncells);
  return hipPeekAtLastError() == hipSuccess ? 0 : -1;
}

} // namespace

extern "C" int SUNAMD_KinNewtonSolve_Mi355(const SUNAMD_KinNewtonOptions* o, double* stats,
                                           double* uout)
{
  constexpr int SETUP_FAIL = -100;
  if (!o || !stats || o->ncells < 1 || o->variant != 1 || o->jac != 1 ||
      (o->guess != 1 && o->guess != 2) ||
      (o->strategy != KIN_NONE && o->strategy != KIN_LINESEARCH))
  {
    return SETUP_FAIL;
  }
  SUNContext ctx = NULL;
  N_Vector u = NULL, scale = NULL, cons = NULL;
  SUNMatrix J        = NULL;
  SUNLinearSolver LS = NULL;
  KinNDev D          = {o->ncells};
  void* kmem         = NULL;
  int rc             = SETUP_FAIL;
  const sunindextype n = (sunindextype)(KFT_NEQ * o->ncells);
  bool ok              = SUNContext_Create(SUN_COMM_NULL, &ctx) == 0;
  if (ok) { ok = (u = N_VNew_Mi355(n, ctx)) != NULL; }
  if (ok)
  {
    ok = N_VSetReductionOrder_Mi355(u, o->sequential ? SUNAMD_REDUCE_SEQUENTIAL
                                                     : SUNAMD_REDUCE_TREE) == 0;
  }
  if (ok) { ok = (scale = N_VClone(u)) != NULL && (cons = N_VClone(u)) != NULL; }
  if (ok)
  {
    hipStream_t s = (hipStream_t)N_VGetStream_Mi355(u);
    k_kft_init<<<blocks(o->ncells), TPB, 0, s>>>(o->guess, o->perturb,
                                                 N_VGetDeviceArrayPointer(u),
                                                 N_VGetDeviceArrayPointer(cons), o->ncells);
    ok = hipPeekAtLastError() == hipSuccess;
    N_VConst(1.0, scale);
  }
  if (ok)
  {
    ok = (J = SUNMatrix_Mi355BlockDense(o->ncells, KFT_NEQ, N_VGetStream_Mi355(u), ctx)) != NULL;
  }
  if (ok) { ok = (LS = SUNLinSol_Mi355BlockLU(u, J, SUNAMD_LU_FP64, ctx)) != NULL; }
  if (ok) { ok = (kmem = KINCreate(ctx)) != NULL; }
  if (ok)
  {
    int r = KINSetUserData(kmem, &D);
    r |= KINSetConstraints(kmem, cons);
    r |= KINSetFuncNormTol(kmem, o->fnormtol);
    r |= KINSetScaledStepTol(kmem, o->scsteptol);
    r |= KINSetNumMaxIters(kmem, o->maxiter);
    r |= KINSetMaxSetupCalls(kmem, o->msbset);
    r |= KINInit(kmem, cb_f, u);
    r |= KINSetLinearSolver(kmem, LS, J);
    r |= KINSetJacFn(kmem, cb_jac);
    if (o->fused) { r |= KINSetUseFusedKernels(kmem, SUNTRUE); }
    ok = r == 0;
  }
  if (ok)
  {
    (void)N_VSynchronize_Mi355(u);
    auto t0 = std::chrono::steady_clock::now();
    rc      = KINSol(kmem, u, o->strategy, scale, scale);
    (void)N_VSynchronize_Mi355(u);
    auto t1 = std::chrono::steady_clock::now();
    long nni = 0, nfe = 0, nje = 0, nfeD = 0, nbcf = 0, nbk = 0;
    KINGetNumNonlinSolvIters(kmem, &nni);
    KINGetNumFuncEvals(kmem, &nfe);
    KINGetNumJacEvals(kmem, &nje);
    KINGetNumLinFuncEvals(kmem, &nfeD);
    KINGetNumBetaCondFails(kmem, &nbcf);
    KINGetNumBacktrackOps(kmem, &nbk);
    stats[0] = rc;
    stats[1] = (double)nni;
    stats[2] = (double)nfe;
    stats[3] = (double)nje;
    stats[4] = (double)nfeD;
    stats[5] = (double)nbcf;
    stats[6] = (double)nbk;
    stats[7] = std::chrono::duration<double>(t1 - t0).count();
    if (hipGetLastError() != hipSuccess) { ok = false; }
    if (uout != NULL && N_VCopyToHostArray_Mi355(u, uout) != 0) { ok = false; }
  }
  if (kmem) { KINFree(&kmem); }
  if (LS) { SUNLinSolFree(LS); }
  if (J) { SUNMatDestroy(J); }
  if (cons) { N_VDestroy(cons); }
  if (scale) { N_VDestroy(scale); }
  if (u) { N_VDestroy(u); }
  if (ctx) { SUNContext_Free(&ctx); }
  return ok ? rc : SETUP_FAIL;
}
