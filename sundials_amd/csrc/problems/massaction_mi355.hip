// -----------------------------------------------------------------------------
// Mass-action right-hand side and block Jacobian on the MI355X, and the
// coupled-system driver (every cell in one N_Vector, block-diagonal Jacobian,
// BDF + modified Newton + batched block LU).
//
// RHS and Jacobian: one group of G = pow2ceil(m) lanes per cell (lane =
// species, or Jacobian row), 64 / G cells per wavefront, one-wavefront
// workgroups striding over the cells.  The RHS evaluates each reaction's net
// rate once (reactions spread over the group's lanes, into LDS) and each
// species then adds its terms in reaction / role order -- the order in which
// the serial evaluation updates f_i -- so f is bit-identical to it.  The
// Jacobian row of each lane is assembled in LDS from that lane's terms only,
// in the serial evaluation's reaction / target / column order, and the block
// is streamed out column by column.  Rate constants are stored cell-major
// (cell * R + r), so a cell's rates are one coalesced read by its group; the
// species, rates and results move as whole rows of the cell.
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
  int* soff;  // m+1: species s's terms are sent[soff[s] .. soff[s+1])
  int* sent;  // 4R term codes 4r + role (role 0, 1: reactants; 2, 3: products)
  double* kf; // ncells*R, cell-major
  double* kb;
  hipStream_t stream;
  unsigned rhs_grid, jac_grid; // launch sizes, chosen at the first call
};

static int hipfail(hipError_t e, const char* what)
{
  fprintf(stderr, "[sundials_amd massaction] %s failed: %s\n", what, hipGetErrorString(e));
  fflush(stderr);
  return -1;
}

// LDS of one workgroup: per group its species (m doubles), then the
// reaction table, the term offsets and the term codes; doubles counted in
// pairs so the int4 reads of the table are 16-byte aligned
static inline size_t ma_table_ints(int m, int R) { return (size_t)(8 * R + ((m + 2) & ~1)); }

__device__ __forceinline__ void ma_stage_tables(const int* __restrict__ idx,
                                                const int* __restrict__ soff,
                                                const int* __restrict__ sent, int* s_idx,
                                                int* s_soff, int* s_sent, int m, int R)
{
  for (int t = threadIdx.x; t < 4 * R; t += blockDim.x)
  {
    s_idx[t]  = idx[t];
    s_sent[t] = sent[t];
  }
  for (int t = threadIdx.x; t <= m; t += blockDim.x) { s_soff[t] = soff[t]; }
}

// f = RHS(y): per group, y (m) and the net rates w (R) in LDS
__global__ void __launch_bounds__(64) k_ma_rhs(const double* __restrict__ y,
                                               double* __restrict__ f,
                                               const int* __restrict__ idx,
                                               const int* __restrict__ soff,
                                               const int* __restrict__ sent,
                                               const double* __restrict__ kf,
                                               const double* __restrict__ kb, int m, int G,
                                               int R, long ncells)
{
  extern __shared__ __align__(16) double sm[];
  const int gpw   = 64 / G;
  const long per  = (long)((m + R + 1) & ~1);
  const int grp   = threadIdx.x / G;
  const int i     = threadIdx.x % G;
  double* ys      = sm + grp * per;
  double* ws      = ys + m;
  int* s_idx      = (int*)(sm + gpw * per);
  int* s_sent     = s_idx + 4 * R;
  int* s_soff     = s_sent + 4 * R;
  ma_stage_tables(idx, soff, sent, s_idx, s_soff, s_sent, m, R);
  __syncthreads();
  const int e0    = (i < m) ? s_soff[i] : 0;
  const int e1    = (i < m) ? s_soff[i + 1] : 0;
  const int4* ix4 = (const int4*)s_idx;
  // the loop bound is the workgroup's, so every lane meets every barrier
  for (long c0 = (long)blockIdx.x * gpw; c0 < ncells; c0 += (long)gridDim.x * gpw)
  {
    const long cell = c0 + grp;
    const bool live = cell < ncells;
    if (live && i < m) { ys[i] = y[cell * m + i]; }
    __syncthreads();
    if (live)
    {
      const double* kfc = kf + cell * R;
      const double* kbc = kb + cell * R;
      for (int r = i; r < R; r += G)
      {
        const int4 ix   = ix4[r];
        const double wf = (kfc[r] * ys[ix.x]) * ys[ix.y];
        const double wb = (kbc[r] * ys[ix.z]) * ys[ix.w];
        ws[r]           = wf - wb;
      }
    }
    __syncthreads();
    if (live && i < m)
    {
      double acc = 0.0;
      for (int e = e0; e < e1; e++)
      {
        const int code = s_sent[e];
        const double w = ws[code >> 2];
        acc            = ((code & 3) < 2) ? acc - w : acc + w;
      }
      f[cell * m + i] = acc;
    }
    __syncthreads();
  }
}

// J = df/dy, block column-major: per group, the block (m*m), y (m) and the
// cell's rates (2R) in LDS
__global__ void __launch_bounds__(64) k_ma_jac(const double* __restrict__ y,
                                               double* __restrict__ J,
                                               const int* __restrict__ idx,
                                               const int* __restrict__ soff,
                                               const int* __restrict__ sent,
                                               const double* __restrict__ kf,
                                               const double* __restrict__ kb, int m, int G,
                                               int R, long ncells)
{
  extern __shared__ __align__(16) double sj[];
  const int gpw   = 64 / G;
  const int mm    = m * m;
  const long per  = (long)((mm + m + 2 * R + 1) & ~1);
  const int grp   = threadIdx.x / G;
  const int i     = threadIdx.x % G;
  double* Jl      = sj + grp * per;
  double* ys      = Jl + mm;
  double* rk      = ys + m;
  int* s_idx      = (int*)(sj + gpw * per);
  int* s_sent     = s_idx + 4 * R;
  int* s_soff     = s_sent + 4 * R;
  ma_stage_tables(idx, soff, sent, s_idx, s_soff, s_sent, m, R);
  __syncthreads();
  const int e0    = (i < m) ? s_soff[i] : 0;
  const int e1    = (i < m) ? s_soff[i + 1] : 0;
  const int4* ix4 = (const int4*)s_idx;
  for (long c0 = (long)blockIdx.x * gpw; c0 < ncells; c0 += (long)gridDim.x * gpw)
  {
    const long cell = c0 + grp;
    const bool live = cell < ncells;
    if (live)
    {
      for (int e = i; e < mm; e += G) { Jl[e] = 0.0; }
      if (i < m) { ys[i] = y[cell * m + i]; }
      for (int r = i; r < R; r += G)
      {
        rk[r]     = kf[cell * R + r];
        rk[R + r] = kb[cell * R + r];
      }
    }
    __syncthreads();
    if (live && i < m)
    {
      // row i: the terms in which species i is a target, each adding the
      // reaction's partial derivatives to the columns of its four species
      for (int e = e0; e < e1; e++)
      {
        const int code  = s_sent[e];
        const int r     = code >> 2;
        const bool loss = (code & 3) < 2;
        const int4 ix   = ix4[r];
        const double kfv = rk[r], kbv = rk[R + r];
        const int col[4]   = {ix.x, ix.y, ix.z, ix.w};
        const double dw[4] = {kfv * ys[ix.y], kfv * ys[ix.x], -(kbv * ys[ix.w]),
                              -(kbv * ys[ix.z])};
#pragma unroll
        for (int q = 0; q < 4; q++)
        {
          const int a = col[q] * m + i;
          Jl[a]       = loss ? (Jl[a] - dw[q]) : (Jl[a] + dw[q]);
        }
      }
    }
    __syncthreads();
    if (live)
    {
      double* Jc = J + cell * (long)mm;
      for (int e = i; e < mm; e += G) { Jc[e] = Jl[e]; }
    }
    __syncthreads();
  }
}

static int pow2ceil_h(int m)
{
  int g = 1;
  while (g < m) { g <<= 1; }
  return g;
}

// dynamic LDS of the RHS and Jacobian workgroups
static size_t ma_rhs_lds(int m, int R)
{
  const int gpw = 64 / pow2ceil_h(m);
  return sizeof(double) * (size_t)gpw * (size_t)((m + R + 1) & ~1) +
         sizeof(int) * ma_table_ints(m, R);
}

static size_t ma_jac_lds(int m, int R)
{
  const int gpw = 64 / pow2ceil_h(m);
  return sizeof(double) * (size_t)gpw * (size_t)((m * m + m + 2 * R + 1) & ~1) +
         sizeof(int) * ma_table_ints(m, R);
}

extern "C" SUNAMD_MassActionDev* SUNAMD_MassActionToDevice(const SUNAMD_MassAction* p,
                                                           void* stream)
{
  if (!p || p->m < 1 || p->m > 64 || p->R < 1) { return NULL; }
  if (ma_jac_lds(p->m, p->R) > 65536 || ma_rhs_lds(p->m, p->R) > 65536) { return NULL; }
  for (int t = 0; t < 4 * p->R; t++)
  {
    if (p->idx[t] < 0 || p->idx[t] >= p->m) { return NULL; }
  }
  SUNAMD_MassActionDev* d = (SUNAMD_MassActionDev*)calloc(1, sizeof(SUNAMD_MassActionDev));
  if (!d) { return NULL; }
  d->m      = p->m;
  d->R      = p->R;
  d->ncells = (long)p->ncells;
  d->stream = (hipStream_t)stream;
  const size_t nr = (size_t)p->R * (size_t)p->ncells;
  // species s's terms in reaction order, roles a b c d within a reaction:
  // the order in which the serial RHS and Jacobian accumulate them
  int* off  = (int*)calloc((size_t)p->m + 1, sizeof(int));
  int* ent  = (int*)malloc(sizeof(int) * 4 * (size_t)p->R);
  int* fill = (int*)malloc(sizeof(int) * (size_t)p->m);
  hipError_t e = (off && ent && fill) ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess)
  {
    for (int t = 0; t < 4 * p->R; t++) { off[p->idx[t] + 1]++; }
    for (int q = 0; q < p->m; q++) { off[q + 1] += off[q]; }
    memcpy(fill, off, sizeof(int) * (size_t)p->m);
    for (int t = 0; t < 4 * p->R; t++) { ent[fill[p->idx[t]]++] = t; }
  }
  if (e == hipSuccess) { e = hipMalloc((void**)&d->idx, sizeof(int) * 4 * (size_t)p->R); }
  if (e == hipSuccess) { e = hipMalloc((void**)&d->soff, sizeof(int) * ((size_t)p->m + 1)); }
  if (e == hipSuccess) { e = hipMalloc((void**)&d->sent, sizeof(int) * 4 * (size_t)p->R); }
  if (e == hipSuccess) { e = hipMalloc((void**)&d->kf, sizeof(double) * nr); }
  if (e == hipSuccess) { e = hipMalloc((void**)&d->kb, sizeof(double) * nr); }
  if (e == hipSuccess)
  {
    e = hipMemcpy(d->idx, p->idx, sizeof(int) * 4 * (size_t)p->R, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess)
  {
    e = hipMemcpy(d->soff, off, sizeof(int) * ((size_t)p->m + 1), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess)
  {
    e = hipMemcpy(d->sent, ent, sizeof(int) * 4 * (size_t)p->R, hipMemcpyHostToDevice);
  }
  // the host problem keeps its rates cell-major already
  if (e == hipSuccess) { e = hipMemcpy(d->kf, p->kf, sizeof(double) * nr, hipMemcpyHostToDevice); }
  if (e == hipSuccess) { e = hipMemcpy(d->kb, p->kb, sizeof(double) * nr, hipMemcpyHostToDevice); }
  free(off);
  free(ent);
  free(fill);
  if (e != hipSuccess)
  {
    hipfail(e, "problem upload");
    SUNAMD_MassActionDevFree(d);
    return NULL;
  }
  return d;
}

extern "C" void SUNAMD_MassActionDevFree(SUNAMD_MassActionDev* d)
{
  if (!d) { return; }
  void* ps[] = {d->idx, d->soff, d->sent, d->kf, d->kb};
  for (void* q : ps)
  {
    if (q) { (void)hipFree(q); }
  }
  free(d);
}

// workgroups for a grid-stride launch: enough to fill every CU at the
// kernel's occupancy, no more than the cells need
template <typename K>
static unsigned ma_grid(K kern, size_t shm, long ncells, int gpw)
{
  static int cus = 0;
  if (cus == 0)
  {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
    {
      cus = 256;
    }
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, shm) != hipSuccess ||
      per_cu < 1)
  {
    per_cu = 1;
  }
  const long need = (ncells + gpw - 1) / gpw;
  const long full = 4L * per_cu * cus;
  return (unsigned)(need < full ? need : full);
}

extern "C" int SUNAMD_MassActionRhsDevArrays(SUNAMD_MassActionDev* d, const double* dy,
                                             double* df)
{
  const int G   = pow2ceil_h(d->m);
  const int gpw = 64 / G;
  const size_t shm = ma_rhs_lds(d->m, d->R);
  if (d->ncells < 1) { return 0; }
  if (d->rhs_grid == 0) { d->rhs_grid = ma_grid(k_ma_rhs, shm, d->ncells, gpw); }
  const unsigned g = d->rhs_grid;
  sunamd_prof_begin(SUNAMD_K_RHS, d->stream);
  k_ma_rhs<<<g, 64, shm, d->stream>>>(dy, df, d->idx, d->soff, d->sent, d->kf, d->kb, d->m, G,
                                      d->R, d->ncells);
  sunamd_prof_end(SUNAMD_K_RHS, d->stream);
  hipError_t e = hipGetLastError();
  return (e == hipSuccess) ? 0 : hipfail(e, "rhs launch");
}

extern "C" int SUNAMD_MassActionJacDevArrays(SUNAMD_MassActionDev* d, const double* dy,
                                             double* dJ)
{
  const int G   = pow2ceil_h(d->m);
  const int gpw = 64 / G;
  const size_t shm = ma_jac_lds(d->m, d->R);
  if (d->ncells < 1) { return 0; }
  if (d->jac_grid == 0) { d->jac_grid = ma_grid(k_ma_jac, shm, d->ncells, gpw); }
  const unsigned g = d->jac_grid;
  k_ma_jac<<<g, 64, shm, d->stream>>>(dy, dJ, d->idx, d->soff, d->sent, d->kf, d->kb, d->m, G,
                                      d->R, d->ncells);
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
  job.kf           = r->dev->kf;
  job.kb           = r->dev->kb;
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
