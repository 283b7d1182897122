// -----------------------------------------------------------------------------
// Mass-action right-hand side and block Jacobian on the MI355X, and the
// coupled-system driver (every cell in one N_Vector, block-diagonal Jacobian,
// BDF + modified Newton + batched block LU).
//
// RHS: one lane per cell; the workgroup stages its cells' species through LDS
// with coalesced loads (row stride m+1 to spread LDS banks), accumulates f in
// LDS in the CPU's reaction order, and writes f back coalesced.  Rate
// constants are stored reaction-major (r*ncells + cell) so each reaction's
// rates are one coalesced load across the workgroup.
// Jacobian: one group of pow2ceil(m) lanes per cell, lane = row; the block is
// assembled in LDS (each lane only touches its own row, in the CPU's
// reaction / target / column order) and streamed out column by column.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../batched/coupled_device.h"
#include "../core/profiler_internal.h"
#include "sunamd/cvode_sunamd.h"
#include "sunamd/sunamd_profiler.h"
#include "sunamd/massaction_mi355.h"
#include "sunamd/nvector_mi355.h"
#include "sunamd/sunmatrix_mi355block.h"

struct SUNAMD_MassActionDev_
{
  int m, R;
  long ncells;
  int* idx;   // 4R
  double* kf; // R*ncells, reaction-major
  double* kb;
  hipStream_t stream;
};

static int hipfail(hipError_t e, const char* what)
{
  fprintf(stderr, "[sundials_amd massaction] %s failed: %s\n", what, hipGetErrorString(e));
  fflush(stderr);
  return -1;
}

// cells per workgroup (one lane per cell): 64, or 32 above 32 species so the
// two LDS tiles stay within 64 KiB
static inline int rhs_cpb(int m) { return m <= 32 ? 64 : 32; }

__global__ void __launch_bounds__(64) k_ma_rhs(const double* __restrict__ y,
                                               double* __restrict__ f,
                                               const int* __restrict__ idx,
                                               const double* __restrict__ kf,
                                               const double* __restrict__ kb, int m, int R,
                                               long ncells)
{
  extern __shared__ double sm[];
  const int cpb    = blockDim.x;
  const int ld     = m + 1;
  double* ys       = sm;
  double* fs       = sm + cpb * ld;
  const long cell0 = (long)blockIdx.x * cpb;
  const long rem   = ncells - cell0;
  const int ncb    = rem < cpb ? (int)rem : cpb;
  const double* yb = y + cell0 * m;
  for (int e = threadIdx.x; e < ncb * m; e += cpb)
  {
    int c                  = e / m;
    ys[c * ld + e - c * m] = yb[e];
  }
  __syncthreads();
  const int c = threadIdx.x;
  if (c < ncb)
  {
    const double* yc = ys + c * ld;
    double* fc       = fs + c * ld;
    for (int s = 0; s < m; s++) { fc[s] = 0.0; }
    const long cell = cell0 + c;
    for (int r = 0; r < R; r++)
    {
      const int a = idx[4 * r], b = idx[4 * r + 1], cc = idx[4 * r + 2], d = idx[4 * r + 3];
      double wf = (kf[(long)r * ncells + cell] * yc[a]) * yc[b];
      double wb = (kb[(long)r * ncells + cell] * yc[cc]) * yc[d];
      double w  = wf - wb;
      fc[a]     = fc[a] - w;
      fc[b]     = fc[b] - w;
      fc[cc]    = fc[cc] + w;
      fc[d]     = fc[d] + w;
    }
  }
  __syncthreads();
  double* fb = f + cell0 * m;
  for (int e = threadIdx.x; e < ncb * m; e += cpb)
  {
    int cc = e / m;
    fb[e]  = fs[cc * ld + e - cc * m];
  }
}

__global__ void __launch_bounds__(64) k_ma_jac(const double* __restrict__ y,
                                               double* __restrict__ J,
                                               const int* __restrict__ idx,
                                               const double* __restrict__ kf,
                                               const double* __restrict__ kb, int m, int G,
                                               int R, long ncells)
{
  extern __shared__ double sj[];
  const int lane   = threadIdx.x;
  const int grp    = lane / G;
  const int i      = lane % G;
  const long cell  = (long)blockIdx.x * (64 / G) + grp;
  const bool live  = cell < ncells;
  const bool valid = live && i < m;
  double* Jl       = sj + (long)grp * m * m;
  if (live)
  {
    for (int e = i; e < m * m; e += G) { Jl[e] = 0.0; }
  }
  __syncthreads();
  if (valid)
  {
    const double* yc = y + cell * m;
    for (int r = 0; r < R; r++)
    {
      int ix[4] = {idx[4 * r], idx[4 * r + 1], idx[4 * r + 2], idx[4 * r + 3]};
      if (ix[0] != i && ix[1] != i && ix[2] != i && ix[3] != i) { continue; }
      double kfv   = kf[(long)r * ncells + cell];
      double kbv   = kb[(long)r * ncells + cell];
      double dw[4] = {kfv * yc[ix[1]], kfv * yc[ix[0]], -(kbv * yc[ix[3]]), -(kbv * yc[ix[2]])};
      for (int t = 0; t < 4; t++)
      {
        if (ix[t] != i) { continue; }
        for (int q = 0; q < 4; q++)
        {
          int e = ix[q] * m + i;
          Jl[e] = (t < 2) ? (Jl[e] - dw[q]) : (Jl[e] + dw[q]);
        }
      }
    }
  }
  __syncthreads();
  if (live)
  {
    double* Jc = J + cell * (long)m * m;
    for (int e = i; e < m * m; e += G) { Jc[e] = Jl[e]; }
  }
}

static int pow2ceil_h(int m)
{
  int g = 1;
  while (g < m) { g <<= 1; }
  return g;
}

extern "C" SUNAMD_MassActionDev* SUNAMD_MassActionToDevice(const SUNAMD_MassAction* p,
                                                           void* stream)
{
  if (!p || p->m > 64) { return NULL; }
  SUNAMD_MassActionDev* d = (SUNAMD_MassActionDev*)calloc(1, sizeof(SUNAMD_MassActionDev));
  if (!d) { return NULL; }
  d->m      = p->m;
  d->R      = p->R;
  d->ncells = (long)p->ncells;
  d->stream = (hipStream_t)stream;
  size_t nr = (size_t)p->R * (size_t)p->ncells;
  double* t = (double*)malloc(sizeof(double) * nr);
  hipError_t e = hipMalloc((void**)&d->idx, sizeof(int) * 4 * (size_t)p->R);
  if (e == hipSuccess) { e = hipMalloc((void**)&d->kf, sizeof(double) * nr); }
  if (e == hipSuccess) { e = hipMalloc((void**)&d->kb, sizeof(double) * nr); }
  if (e == hipSuccess && t)
  {
    e = hipMemcpy(d->idx, p->idx, sizeof(int) * 4 * (size_t)p->R, hipMemcpyHostToDevice);
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++)
    {
      const double* src = pass ? p->kb : p->kf;
      for (long c = 0; c < d->ncells; c++)
      {
        for (int r = 0; r < p->R; r++) { t[(size_t)r * d->ncells + c] = src[c * p->R + r]; }
      }
      e = hipMemcpy(pass ? d->kb : d->kf, t, sizeof(double) * nr, hipMemcpyHostToDevice);
    }
  }
  free(t);
  if (e != hipSuccess || !t)
  {
    if (e != hipSuccess) { hipfail(e, "problem upload"); }
    SUNAMD_MassActionDevFree(d);
    return NULL;
  }
  return d;
}

extern "C" void SUNAMD_MassActionDevFree(SUNAMD_MassActionDev* d)
{
  if (!d) { return; }
  if (d->idx) { (void)hipFree(d->idx); }
  if (d->kf) { (void)hipFree(d->kf); }
  if (d->kb) { (void)hipFree(d->kb); }
  free(d);
}

extern "C" int SUNAMD_MassActionRhsDevArrays(SUNAMD_MassActionDev* d, const double* dy,
                                             double* df)
{
  const int cpb = rhs_cpb(d->m);
  long g        = (d->ncells + cpb - 1) / cpb;
  size_t shm    = sizeof(double) * 2 * cpb * (size_t)(d->m + 1);
  sunamd_prof_begin(SUNAMD_K_RHS, d->stream);
  k_ma_rhs<<<(unsigned)g, cpb, shm, d->stream>>>(dy, df, d->idx, d->kf, d->kb, d->m, d->R,
                                                 d->ncells);
  sunamd_prof_end(SUNAMD_K_RHS, d->stream);
  hipError_t e = hipGetLastError();
  return (e == hipSuccess) ? 0 : hipfail(e, "rhs launch");
}

extern "C" int SUNAMD_MassActionJacDevArrays(SUNAMD_MassActionDev* d, const double* dy,
                                             double* dJ)
{
  int G      = pow2ceil_h(d->m);
  int cpw    = 64 / G;
  long g     = (d->ncells + cpw - 1) / cpw;
  size_t shm = sizeof(double) * (size_t)cpw * d->m * d->m;
  k_ma_jac<<<(unsigned)g, 64, shm, d->stream>>>(dy, dJ, d->idx, d->kf, d->kb, d->m, G, d->R,
                                                d->ncells);
  hipError_t e = hipGetLastError();
  return (e == hipSuccess) ? 0 : hipfail(e, "jac launch");
}

extern "C" int SUNAMD_MassActionRhsDev(sunrealtype t, N_Vector y, N_Vector ydot,
                                       void* user_data)
{
  (void)t;
  SUNAMD_MassActionDev* d = (SUNAMD_MassActionDev*)user_data;
  return SUNAMD_MassActionRhsDevArrays(d, N_VGetDeviceArrayPointer(y),
                                       N_VGetDeviceArrayPointer(ydot));
}

extern "C" int SUNAMD_MassActionJacDev(sunrealtype t, N_Vector y, N_Vector fy, SUNMatrix J,
                                       void* user_data, N_Vector tmp1, N_Vector tmp2,
                                       N_Vector tmp3)
{
  (void)t;
  (void)fy;
  (void)tmp1;
  (void)tmp2;
  (void)tmp3;
  SUNAMD_MassActionDev* d = (SUNAMD_MassActionDev*)user_data;
  return SUNAMD_MassActionJacDevArrays(d, N_VGetDeviceArrayPointer(y),
                                       SUNMatrix_Mi355BlockDense_DeviceData(J));
}

struct SUNAMD_CoupledRun_
{
  SUNContext ctx;
  sunindextype N;
  N_Vector y, y0;
  SUNAMD_MassActionDev* dev;
  SUNMatrix A;
  SUNLinearSolver LS;
  void* cv;
  // the device-stepped integrator (coupled_device.hip) and what it needs
  int devstep;
  void* dstate;
  double rtol, atol;
  int lu_precision, reduction_order;
};

extern "C" int SUNAMD_CoupledSetDeviceStepping_Mi355(SUNAMD_CoupledRun* r, int onoff)
{
  if (!r) { return -1; }
  r->devstep = onoff ? 1 : 0;
  return 0;
}

extern "C" int SUNAMD_CoupledSetFusedKernels_Mi355(SUNAMD_CoupledRun* r, int onoff)
{
  if (!r || !r->cv) { return -1; }
  return CVodeSetUseIntegratorFusedKernels(r->cv, onoff ? SUNTRUE : SUNFALSE);
}

extern "C" void SUNAMD_CoupledFree_Mi355(SUNAMD_CoupledRun* r)
{
  if (!r) { return; }
  if (r->dstate) { sunamd_coupled_device_free(r->dstate); }
  if (r->cv) { CVodeFree(&r->cv); }
  if (r->LS) { SUNLinSolFree(r->LS); }
  if (r->A) { SUNMatDestroy(r->A); }
  if (r->dev) { SUNAMD_MassActionDevFree(r->dev); }
  if (r->y0) { N_VDestroy(r->y0); }
  if (r->y) { N_VDestroy(r->y); }
  if (r->ctx) { SUNContext_Free(&r->ctx); }
  free(r);
}

extern "C" SUNAMD_CoupledRun* SUNAMD_CoupledCreate_Mi355(const SUNAMD_MassAction* p,
                                                         double rtol, double atol,
                                                         int lu_precision,
                                                         int reduction_order)
{
  SUNAMD_CoupledRun* r = (SUNAMD_CoupledRun*)calloc(1, sizeof(SUNAMD_CoupledRun));
  if (!r) { return NULL; }
  bool ok = SUNContext_Create(SUN_COMM_NULL, &r->ctx) == 0;
  r->N    = (sunindextype)p->ncells * p->m;
  r->rtol = rtol;
  r->atol = atol;
  r->lu_precision    = lu_precision;
  r->reduction_order = reduction_order;
  if (ok) { ok = (r->y = N_VNew_Mi355(r->N, r->ctx)) != NULL; }
  if (ok) { ok = N_VSetReductionOrder_Mi355(r->y, reduction_order) == 0; }
  if (ok) { ok = (r->y0 = N_VClone(r->y)) != NULL; }
  if (ok) { ok = N_VCopyFromHostArray_Mi355(r->y0, p->y0) == 0; }
  if (ok) { N_VScale(1.0, r->y0, r->y); }
  if (ok) { ok = (r->dev = SUNAMD_MassActionToDevice(p, N_VGetStream_Mi355(r->y))) != NULL; }
  if (ok)
  {
    ok = (r->A = SUNMatrix_Mi355BlockDense(p->ncells, p->m, N_VGetStream_Mi355(r->y),
                                           r->ctx)) != NULL;
  }
  if (ok) { ok = (r->LS = SUNLinSol_Mi355BlockLU(r->y, r->A, lu_precision, r->ctx)) != NULL; }
  if (ok) { ok = (r->cv = CVodeCreate(CV_BDF, r->ctx)) != NULL; }
  if (ok) { ok = CVodeInit(r->cv, SUNAMD_MassActionRhsDev, 0.0, r->y) == 0; }
  if (ok) { ok = CVodeSStolerances(r->cv, rtol, atol) == 0; }
  if (ok) { ok = CVodeSetUserData(r->cv, r->dev) == 0; }
  if (ok) { ok = CVodeSetLinearSolver(r->cv, r->LS, r->A) == 0; }
  if (ok) { ok = CVodeSetJacFn(r->cv, SUNAMD_MassActionJacDev) == 0; }
  if (ok) { ok = CVodeSetMaxNumSteps(r->cv, 1000000) == 0; }
  if (ok) { ok = hipDeviceSynchronize() == hipSuccess; }
  if (!ok)
  {
    SUNAMD_CoupledFree_Mi355(r);
    return NULL;
  }
  return r;
}

static int coupled_run_device(SUNAMD_CoupledRun* r, double tout, SUNAMD_SolveStats* st)
{
  if (hipDeviceSynchronize() != hipSuccess) { return -1; }
  CoupledDeviceJob job;
  job.m            = r->dev->m;
  job.R            = r->dev->R;
  job.ncells       = r->dev->ncells;
  job.idx          = r->dev->idx;
  job.kf_rm        = r->dev->kf;
  job.kb_rm        = r->dev->kb;
  job.y0           = N_VGetDeviceArrayPointer(r->y0);
  job.yout         = N_VGetDeviceArrayPointer(r->y);
  job.lu           = SUNMatrix_Mi355BlockDense_DeviceData(r->A);
  job.piv          = SUNLinSol_Mi355BlockLU_DevicePivots(r->LS);
  job.rtol         = r->rtol;
  job.atol         = r->atol;
  job.tout         = tout;
  job.mxstep       = 1000000;
  job.lu_precision = r->lu_precision;
  job.sequential   = r->reduction_order == SUNAMD_REDUCE_SEQUENTIAL;
  job.stream       = (hipStream_t)N_VGetStream_Mi355(r->y);
  SUNAMD_SolveStats s;
  memset(&s, 0, sizeof(s));
  auto t0  = std::chrono::steady_clock::now();
  int flag = sunamd_coupled_device_run(job, &r->dstate, &s);
  if (hipDeviceSynchronize() != hipSuccess) { return -1; }
  auto t1 = std::chrono::steady_clock::now();
  s.wall_s = std::chrono::duration<double>(t1 - t0).count();
  s.flag   = flag;
  if (st) { *st = s; }
  return (flag < 0) ? flag : 0;
}

extern "C" int SUNAMD_CoupledRun_Mi355(SUNAMD_CoupledRun* r, double tout, SUNAMD_SolveStats* st)
{
  sunrealtype t = 0.0;
  if (r->devstep) { return coupled_run_device(r, tout, st); }
  if (hipDeviceSynchronize() != hipSuccess) { return -1; }
  long l0 = N_VGetNumLaunches_Mi355(r->y);
  auto t0 = std::chrono::steady_clock::now();
  N_VScale(1.0, r->y0, r->y);
  int flag = CVodeReInit(r->cv, 0.0, r->y);
  if (flag == 0) { flag = CVode(r->cv, tout, r->y, &t, CV_NORMAL); }
  if (hipDeviceSynchronize() != hipSuccess) { return -1; }
  auto t1 = std::chrono::steady_clock::now();
  if (st)
  {
    long counts[8];
    int orders[2];
    double steps[4];
    CVodeGetAllStats(r->cv, counts, orders, steps);
    st->nst      = counts[0];
    st->nfe      = counts[1];
    st->nsetups  = counts[2];
    st->netf     = counts[3];
    st->nni      = counts[4];
    st->nnf      = counts[5];
    st->ncfn     = counts[6];
    st->nje      = counts[7];
    st->qlast    = orders[0];
    st->qcur     = orders[1];
    st->hinused  = steps[0];
    st->hlast    = steps[1];
    st->hcur     = steps[2];
    st->tcur     = steps[3];
    st->wall_s   = std::chrono::duration<double>(t1 - t0).count();
    st->launches = N_VGetNumLaunches_Mi355(r->y) - l0;
    st->flag     = flag;
  }
  return (flag < 0) ? flag : 0;
}

extern "C" int SUNAMD_CoupledGetState_Mi355(SUNAMD_CoupledRun* r, double* yout)
{
  return N_VCopyToHostArray_Mi355(r->y, yout) == 0 ? 0 : -1;
}

extern "C" int SUNAMD_MassActionSolveCoupled_Mi355(const SUNAMD_MassAction* p, double rtol,
                                                   double atol, double tout,
                                                   int lu_precision, int reduction_order,
                                                   double* yout,
                                                   SUNAMD_SolveStats* st)
{
  SUNAMD_CoupledRun* r =
    SUNAMD_CoupledCreate_Mi355(p, rtol, atol, lu_precision, reduction_order);
  if (!r) { return -1; }
  int rc = SUNAMD_CoupledRun_Mi355(r, tout, st);
  if (rc == 0 && yout) { rc = SUNAMD_CoupledGetState_Mi355(r, yout); }
  SUNAMD_CoupledFree_Mi355(r);
  return rc;
}
