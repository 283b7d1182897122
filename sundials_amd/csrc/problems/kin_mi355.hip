// -----------------------------------------------------------------------------
// Device side of the fixed-point problem of kin_fp.h: the map G evaluated one
// thread per cell on the vector's stream, the example's damping and depth
// functions through the vector ops, and the solve driver (KINSol with
// KIN_FP, Anderson acceleration on MI355X vectors).  Cell arithmetic comes
// from kin_cell.h, shared with the CPU twin.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>

#include "kin_cell.h"
#include "sunamd/kin_fp.h"
#include "sunamd/kinsol_sunamd.h"
#include "sunamd/nvector_mi355.h"
#include "sunamd/sundials_core.h"

namespace
{

constexpr int TPB = 256;

inline unsigned blocks(long n) { return (unsigned)((n + TPB - 1) / TPB); }

__global__ void k_kin_g(int variant, const double* __restrict__ u, double* __restrict__ g,
                        long ncells)
{
  const long c = (long)blockIdx.x * TPB + threadIdx.x;
  if (c >= ncells) { return; }
  double uc[3] = {u[3 * c], u[3 * c + 1], u[3 * c + 2]};
  double gc[3];
  kin_cell_g(variant, uc, gc);
  g[3 * c]     = gc[0];
  g[3 * c + 1] = gc[1];
  g[3 * c + 2] = gc[2];
}

__global__ void k_kin_u0(int perturb, double* __restrict__ u, long ncells)
{
  const long c = (long)blockIdx.x * TPB + threadIdx.x;
  if (c >= ncells) { return; }
  double uc[3];
  kin_cell_u0(c, perturb, uc);
  u[3 * c]     = uc[0];
  u[3 * c + 1] = uc[1];
  u[3 * c + 2] = uc[2];
}

struct KinDev
{
  const SUNAMD_KinOptions* o;
  N_Vector tmp;
};

int cb_g(N_Vector u, N_Vector g, void* ud)
{
  const KinDev* D = (const KinDev*)ud;
  const long nc   = D->o->ncells;
  hipStream_t s   = (hipStream_t)N_VGetStream_Mi355(g);
  k_kin_g<<<blocks(nc), TPB, 0, s>>>(D->o->variant, N_VGetDeviceArrayPointer(u),
                                     N_VGetDeviceArrayPointer(g), nc);
  return hipPeekAtLastError() == hipSuccess ? 0 : -1;
}

int cb_damping(long int iter, N_Vector u, N_Vector g, sunrealtype* qt_fn, long int depth,
               void* ud, sunrealtype* beta)
{
  (void)iter;
  const KinDev* D = (const KinDev*)ud;
  if (depth == 0)
  {
    *beta = 0.5;
    return 0;
  }
  double q2 = 0.0;
  for (long int i = 0; i < depth; i++) { q2 += qt_fn[i] * qt_fn[i]; }
  N_VLinearSum(1.0, g, -1.0, u, D->tmp);
  const double f2   = N_VDotProd(D->tmp, D->tmp);
  const double arg  = 1.0 - q2 / f2;
  const double gain = (arg <= 0.0) ? 0.0 : sqrt(arg);
  *beta             = 0.9 - 0.5 * gain;
  return 0;
}

int cb_depth(long int iter, N_Vector u, N_Vector g, N_Vector f, N_Vector* df, sunrealtype* R,
             long int depth, void* ud, long int* new_depth, sunbooleantype* remove)
{
  (void)u; (void)g; (void)f; (void)df; (void)R; (void)ud; (void)remove;
  *new_depth = (iter < 2) ? 1 : depth;
  return 0;
}

} // namespace

extern "C" int SUNAMD_KinSolve_Mi355(const SUNAMD_KinOptions* o, double* stats, double* uout)
{
  if (!o || !stats || o->ncells < 1 || (o->variant != 0 && o->variant != 1)) { return -1; }
  SUNContext ctx = NULL;
  N_Vector u = NULL, scale = NULL;
  KinDev D   = {o, NULL};
  void* kmem = NULL;
  int rc     = -1;
  bool ok    = SUNContext_Create(SUN_COMM_NULL, &ctx) == 0;
  if (ok) { ok = (u = N_VNew_Mi355((sunindextype)(3 * o->ncells), ctx)) != NULL; }
  if (ok)
  {
    ok = N_VSetReductionOrder_Mi355(u, o->sequential ? SUNAMD_REDUCE_SEQUENTIAL
                                                     : SUNAMD_REDUCE_TREE) == 0;
  }
  if (ok) { ok = (scale = N_VClone(u)) != NULL && (D.tmp = N_VClone(u)) != NULL; }
  if (ok)
  {
    hipStream_t s = (hipStream_t)N_VGetStream_Mi355(u);
    k_kin_u0<<<blocks(o->ncells), TPB, 0, s>>>(o->perturb, N_VGetDeviceArrayPointer(u),
                                               o->ncells);
    ok = hipPeekAtLastError() == hipSuccess;
    N_VConst(1.0, scale);
  }
  if (ok) { ok = (kmem = KINCreate(ctx)) != NULL; }
  if (ok)
  {
    int r = KINSetUserData(kmem, &D);
    r |= KINSetMAA(kmem, o->m_aa);
    r |= KINSetOrthAA(kmem, o->orth_aa);
    r |= KINInit(kmem, cb_g, u);
    r |= KINSetFuncNormTol(kmem, o->tol);
    r |= KINSetNumMaxIters(kmem, o->maxiter);
    if (o->m_aa == 0) { r |= KINSetDamping(kmem, o->damping_fp); }
    else
    {
      r |= KINSetDampingAA(kmem, o->damping_aa);
      r |= KINSetDelayAA(kmem, o->delay_aa);
    }
    if (o->damping_fn) { r |= KINSetDampingFn(kmem, cb_damping); }
    if (o->depth_fn) { r |= KINSetDepthFn(kmem, cb_depth); }
    if (o->fused) { r |= KINSetUseFusedKernels(kmem, SUNTRUE); }
    ok = r == 0;
  }
  if (ok)
  {
    (void)N_VSynchronize_Mi355(u);
    auto t0 = std::chrono::steady_clock::now();
    rc      = KINSol(kmem, u, KIN_FP, scale, scale);
    (void)N_VSynchronize_Mi355(u);
    auto t1 = std::chrono::steady_clock::now();
    long nni = 0, nfe = 0;
    double fnorm = 0.0;
    KINGetNumNonlinSolvIters(kmem, &nni);
    KINGetNumFuncEvals(kmem, &nfe);
    KINGetFuncNorm(kmem, &fnorm);
    stats[0] = rc;
    stats[1] = (double)nni;
    stats[2] = (double)nfe;
    stats[3] = fnorm;
    stats[4] = std::chrono::duration<double>(t1 - t0).count();
    if (uout != NULL && N_VCopyToHostArray_Mi355(u, uout) != 0) { rc = -1; }
  }
  if (kmem) { KINFree(&kmem); }
  if (D.tmp) { N_VDestroy(D.tmp); }
  if (scale) { N_VDestroy(scale); }
  if (u) { N_VDestroy(u); }
  if (ctx) { SUNContext_Free(&ctx); }
  return ok ? rc : -1;
}
