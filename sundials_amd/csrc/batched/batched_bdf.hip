// -----------------------------------------------------------------------------
// Batched per-system BDF on the MI355X (host part): uploads a batch of size
// buckets, launches one persistent kernel per bucket on its own stream
// (largest systems first) and gathers results back into system order.  The
// integrator itself is in bdf_device.h; each (m, precision) kernel is compiled
// in one of the bdf_inst_*.hip units.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "sunamd/batched_bdf.h"
#include "sunamd/sunmatrix_mi355block.h"

#include "bdf_args.h"
#include "sunamd/learned_predictor.h"

using sunamd_bdf::BucketArgs;
using sunamd_bdf::pow2ceil;

static constexpr long MXSTEP_DEFAULT = 500; // CVODE's default step limit (cvode_impl.h)

// -----------------------------------------------------------------------------
// host side
// -----------------------------------------------------------------------------
struct BucketDev
{
  int m, R;
  long ncells;
  int* idx;
  int *soff, *sent; // per-species term lists (m+1 offsets, 4R entries)
  double *kf, *kb, *y0, *yout;
  long* stats;
  unsigned long long* next;
  hipStream_t stream;
  hipEvent_t ev0, ev1; // bracket the bucket's launch on its stream
  double last_ms;      // kernel time of the last run
};

struct SUNAMD_BatchedBDF_
{
  int nb;
  BucketDev b[SUNAMD_BATCH_MAXB];
  double rtol, atol;
  long mxstep;
  int prec;
  int num_cus;
  // system-order mapping
  long nsys;
  int* sys_bucket;
  long* sys_cell;
  long* sys_yoff;
  long ytotal;
  // learned initial iterate and training log
  const void* mlp;      // device weight block (SUNAMD_MLPDeviceWeights) or NULL
  int learned_mask;     // bit k: bucket k uses it
  int waves[SUNAMD_BATCH_MAXB]; // register budget of bucket k's kernel (1, 2 or 4 waves/SIMD)
  int grid_cap[SUNAMD_BATCH_MAXB]; // bucket k's launch: at most this many workgroups per CU
  float *log_x, *log_t, *log_m;
  long log_cap;
  int log_stride;
  unsigned long long* dcount; // [0] log rows, [1 + 2k], [2 + 2k]: bucket k tried / kept
  unsigned long long* dphase; // phase clock build: SUNAMD_BDF_PHASE_VALUES per bucket
};

static int hfail(hipError_t e, const char* what)
{
  fprintf(stderr, "[sundials_amd batched_bdf] %s failed: %s\n", what, hipGetErrorString(e));
  fflush(stderr);
  return -1;
}

#define BDF_SIZES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(10) X(12) X(16) X(24) X(32) X(48) X(64)

template <typename T>
static bool launch_bucket(const SUNAMD_BatchedBDF* B, int k, double tout, int grid)
{
  const BucketDev& d  = B->b[k];
  const double rtol   = B->rtol;
  const double atol   = B->atol;
  const long mxstep   = B->mxstep;
  BucketArgs A;
  memset(&A, 0, sizeof(A));
  if (B->mlp && (B->learned_mask >> k) & 1)
  {
    A.mlp_w = (const unsigned short*)B->mlp;
    A.mlp_b = (const float*)((const unsigned short*)B->mlp + SUNAMD_MLP_WEIGHTS_BF16);
  }
  if (B->log_x && B->log_cap > 0)
  {
    A.log_x      = B->log_x;
    A.log_t      = B->log_t;
    A.log_m      = B->log_m;
    A.log_next   = B->dcount;
    A.log_cap    = B->log_cap;
    A.log_stride = B->log_stride > 0 ? B->log_stride : 1;
  }
  A.lstats = B->dcount + 1 + 2 * k;
  A.phase  = B->dphase ? B->dphase + (size_t)k * SUNAMD_BDF_PHASE_VALUES : nullptr;
  {
    const int gpw = 64 / pow2ceil(d.m);
    if (sunamd_bdf::bucket_lds_bytes(d.m, d.R, gpw, A.mlp_w != nullptr, A.log_x != nullptr,
                                     (int)sizeof(T)) > 60000)
    {
      fprintf(stderr, "[sundials_amd batched_bdf] bucket m=%d R=%d needs too much LDS\n", d.m,
              d.R);
      return false;
    }
  }
  A.m      = d.m;
  A.R      = d.R;
  A.ncells = d.ncells;
  A.idx    = d.idx;
  A.soff   = d.soff;
  A.sent   = d.sent;
  A.kf     = d.kf;
  A.kb     = d.kb;
  A.y0     = d.y0;
  A.yout   = d.yout;
  A.stats  = d.stats;
  A.next   = d.next;
  A.rtol   = rtol;
  A.atol   = atol;
  A.tout   = tout;
  A.mxstep = mxstep;
  switch (d.m)
  {
#define BDF_CASE(MM)                                                             \
  case MM: sunamd_bdf::launch<MM, T>(A, grid, d.stream, B->waves[k]); return true;
    BDF_SIZES(BDF_CASE)
#undef BDF_CASE
  default: return false;
  }
}

extern "C" SUNAMD_BatchedBDF* SUNAMD_BatchedBDFCreate(const SUNAMD_Batch* batch, double rtol,
                                                      double atol, int lu_precision)
{
  if (!batch || batch->nbuckets < 1) { return NULL; }
  SUNAMD_BatchedBDF* B = (SUNAMD_BatchedBDF*)calloc(1, sizeof(SUNAMD_BatchedBDF));
  if (!B) { return NULL; }
  B->nb   = batch->nbuckets;
  B->rtol = rtol;
  B->atol = atol;
  B->prec = lu_precision;
  B->mxstep = MXSTEP_DEFAULT;
  for (int k = 0; k < SUNAMD_BATCH_MAXB; k++)
  {
    B->waves[k]    = 1;
    B->grid_cap[k] = 16;
  }
  if (hipMalloc((void**)&B->dcount, sizeof(unsigned long long) * (1 + 2 * SUNAMD_BATCH_MAXB)) !=
        hipSuccess ||
      hipMemset(B->dcount, 0, sizeof(unsigned long long) * (1 + 2 * SUNAMD_BATCH_MAXB)) != hipSuccess)
  {
    free(B);
    return NULL;
  }
#ifdef SUNAMD_BDF_PHASES
  {
    const size_t pb = sizeof(unsigned long long) * SUNAMD_BATCH_MAXB * SUNAMD_BDF_PHASE_VALUES;
    if (hipMalloc((void**)&B->dphase, pb) != hipSuccess)
    {
      (void)hipFree(B->dcount);
      free(B);
      return NULL;
    }
  }
#endif
  int dev = 0;
  hipDeviceProp_t prop;
  B->num_cus = 256;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
  {
    B->num_cus = prop.multiProcessorCount;
  }
  hipError_t e = hipSuccess;
  for (int k = 0; k < B->nb && e == hipSuccess; k++)
  {
    const SUNAMD_MassAction* p = batch->bucket[k];
    BucketDev& d               = B->b[k];
    d.m                        = p->m;
    d.R                        = p->R;
    d.ncells                   = (long)p->ncells;
    if (d.m > 64) { e = hipErrorInvalidValue; break; }
    long nc = d.ncells > 0 ? d.ncells : 1;
    e       = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking);
    if (e == hipSuccess) { e = hipEventCreate(&d.ev0); }
    if (e == hipSuccess) { e = hipEventCreate(&d.ev1); }
    if (e == hipSuccess) { e = hipMalloc((void**)&d.idx, sizeof(int) * 4 * d.R); }
    if (e == hipSuccess) { e = hipMalloc((void**)&d.kf, sizeof(double) * nc * d.R); }
    if (e == hipSuccess) { e = hipMalloc((void**)&d.kb, sizeof(double) * nc * d.R); }
    if (e == hipSuccess) { e = hipMalloc((void**)&d.y0, sizeof(double) * nc * d.m); }
    if (e == hipSuccess) { e = hipMalloc((void**)&d.yout, sizeof(double) * nc * d.m); }
    if (e == hipSuccess) { e = hipMalloc((void**)&d.stats, sizeof(long) * nc * SUNAMD_BS_COUNT); }
    if (e == hipSuccess) { e = hipMalloc((void**)&d.next, sizeof(unsigned long long)); }
    if (e == hipSuccess)
    {
      e = hipMemcpy(d.idx, p->idx, sizeof(int) * 4 * d.R, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
    {
      // species s's RHS / Jacobian-row terms in reaction order, roles a b c d
      // within a reaction: the order the host evaluation accumulates them in
      int* off = (int*)calloc((size_t)d.m + 1, sizeof(int));
      int* ent = (int*)malloc(sizeof(int) * 4 * (size_t)d.R);
      if (!off || !ent) { e = hipErrorOutOfMemory; }
      else
      {
        for (int r = 0; r < d.R; r++)
        {
          for (int t = 0; t < 4; t++) { off[p->idx[4 * r + t] + 1]++; }
        }
        for (int q = 0; q < d.m; q++) { off[q + 1] += off[q]; }
        int* fill = (int*)malloc(sizeof(int) * (size_t)d.m);
        if (!fill) { e = hipErrorOutOfMemory; }
        else
        {
          memcpy(fill, off, sizeof(int) * (size_t)d.m);
          for (int r = 0; r < d.R; r++)
          {
            for (int t = 0; t < 4; t++) { ent[fill[p->idx[4 * r + t]]++] = 4 * r + t; }
          }
          free(fill);
          e = hipMalloc((void**)&d.soff, sizeof(int) * ((size_t)d.m + 1));
          if (e == hipSuccess) { e = hipMalloc((void**)&d.sent, sizeof(int) * 4 * (size_t)d.R); }
          if (e == hipSuccess)
          {
            e = hipMemcpy(d.soff, off, sizeof(int) * ((size_t)d.m + 1), hipMemcpyHostToDevice);
          }
          if (e == hipSuccess)
          {
            e = hipMemcpy(d.sent, ent, sizeof(int) * 4 * (size_t)d.R, hipMemcpyHostToDevice);
          }
        }
      }
      free(off);
      free(ent);
    }
    if (e == hipSuccess && d.ncells > 0)
    {
      e = hipMemcpy(d.kf, p->kf, sizeof(double) * d.ncells * d.R, hipMemcpyHostToDevice);
      if (e == hipSuccess)
      {
        e = hipMemcpy(d.kb, p->kb, sizeof(double) * d.ncells * d.R, hipMemcpyHostToDevice);
      }
      if (e == hipSuccess)
      {
        e = hipMemcpy(d.y0, p->y0, sizeof(double) * d.ncells * d.m, hipMemcpyHostToDevice);
      }
    }
  }
  if (e == hipSuccess)
  {
    B->nsys       = (long)batch->nsys;
    B->ytotal     = (long)batch->ytotal;
    B->sys_bucket = (int*)malloc(sizeof(int) * B->nsys);
    B->sys_cell   = (long*)malloc(sizeof(long) * B->nsys);
    B->sys_yoff   = (long*)malloc(sizeof(long) * B->nsys);
    if (!B->sys_bucket || !B->sys_cell || !B->sys_yoff) { e = hipErrorOutOfMemory; }
    else
    {
      for (long s = 0; s < B->nsys; s++)
      {
        B->sys_bucket[s] = batch->sys_bucket[s];
        B->sys_cell[s]   = (long)batch->sys_cell[s];
        B->sys_yoff[s]   = (long)batch->sys_yoff[s];
      }
    }
  }
  if (e != hipSuccess)
  {
    hfail(e, "batched BDF setup");
    SUNAMD_BatchedBDFFree(B);
    return NULL;
  }
  return B;
}

extern "C" int SUNAMD_BatchedBDFSetMaxNumSteps(SUNAMD_BatchedBDF* B, long int mxsteps)
{
  if (!B) { return -1; }
  B->mxstep = (mxsteps == 0) ? MXSTEP_DEFAULT : mxsteps;
  return 0;
}

extern "C" int SUNAMD_BatchedBDFRun(SUNAMD_BatchedBDF* B, double tout, double* wall_s)
{
  if (hipDeviceSynchronize() != hipSuccess) { return -1; }
  auto t0 = std::chrono::steady_clock::now();
  // largest systems first: they have the longest per-system latency
  int order[SUNAMD_BATCH_MAXB];
  for (int k = 0; k < B->nb; k++) { order[k] = k; }
  for (int a = 0; a < B->nb; a++)
  {
    for (int c = a + 1; c < B->nb; c++)
    {
      if (B->b[order[c]].m > B->b[order[a]].m)
      {
        int t    = order[a];
        order[a] = order[c];
        order[c] = t;
      }
    }
  }
  // learned-guess counters restart with every run (the log keeps filling)
  {
    hipError_t e = hipMemset(B->dcount + 1, 0, sizeof(unsigned long long) * 2 * SUNAMD_BATCH_MAXB);
    if (e != hipSuccess) { return hfail(e, "counter reset"); }
    if (B->dphase)
    {
      e = hipMemset(B->dphase, 0,
                    sizeof(unsigned long long) * SUNAMD_BATCH_MAXB * SUNAMD_BDF_PHASE_VALUES);
      if (e != hipSuccess) { return hfail(e, "phase reset"); }
    }
  }
  for (int n = 0; n < B->nb; n++)
  {
    BucketDev& d = B->b[order[n]];
    if (d.ncells <= 0) { continue; }
    hipError_t e = hipMemsetAsync(d.next, 0, sizeof(unsigned long long), d.stream);
    if (e != hipSuccess) { return hfail(e, "queue reset"); }
    const int gpw = 64 / pow2ceil(d.m);
    long want     = (d.ncells + gpw - 1) / gpw;
    long cap      = (long)B->num_cus * B->grid_cap[order[n]];
    int grid      = (int)(want < cap ? want : cap);
    e = hipEventRecord(d.ev0, d.stream);
    if (e != hipSuccess) { return hfail(e, "event record"); }
    bool ok = (B->prec == SUNAMD_LU_FP32) ? launch_bucket<float>(B, order[n], tout, grid)
                                          : launch_bucket<double>(B, order[n], tout, grid);
    e = hipEventRecord(d.ev1, d.stream);
    if (e != hipSuccess) { return hfail(e, "event record"); }
    if (!ok)
    {
      fprintf(stderr, "[sundials_amd batched_bdf] no kernel for m = %d\n", d.m);
      return -1;
    }
    e = hipGetLastError();
    if (e != hipSuccess) { return hfail(e, "batched BDF launch"); }
  }
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) { return hfail(e, "batched BDF run"); }
  auto t1 = std::chrono::steady_clock::now();
  for (int k = 0; k < B->nb; k++)
  {
    BucketDev& d = B->b[k];
    d.last_ms    = 0.0;
    if (d.ncells <= 0) { continue; }
    float ms = 0.0f;
    e        = hipEventElapsedTime(&ms, d.ev0, d.ev1);
    if (e != hipSuccess) { return hfail(e, "event timing"); }
    d.last_ms = ms;
  }
  if (wall_s) { *wall_s = std::chrono::duration<double>(t1 - t0).count(); }
  return 0;
}

extern "C" int SUNAMD_BatchedBDFGetResults(SUNAMD_BatchedBDF* B, double* y, long int* counters)
{
  int rc = 0;
  for (int k = 0; k < B->nb; k++)
  {
    BucketDev& d = B->b[k];
    if (d.ncells <= 0) { continue; }
    double* hy = (double*)malloc(sizeof(double) * d.ncells * d.m);
    long* hs   = (long*)malloc(sizeof(long) * d.ncells * SUNAMD_BS_COUNT);
    if (!hy || !hs)
    {
      free(hy);
      free(hs);
      return -1;
    }
    hipError_t e = hipMemcpy(hy, d.yout, sizeof(double) * d.ncells * d.m, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
    {
      e = hipMemcpy(hs, d.stats, sizeof(long) * d.ncells * SUNAMD_BS_COUNT,
                    hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess)
    {
      free(hy);
      free(hs);
      return hfail(e, "result copy");
    }
    for (long s = 0; s < B->nsys; s++)
    {
      if (B->sys_bucket[s] != k) { continue; }
      long c = B->sys_cell[s];
      if (y) { memcpy(y + B->sys_yoff[s], hy + c * d.m, sizeof(double) * d.m); }
      if (counters)
      {
        memcpy(counters + s * SUNAMD_BS_COUNT, hs + c * SUNAMD_BS_COUNT,
               sizeof(long) * SUNAMD_BS_COUNT);
        if (hs[c * SUNAMD_BS_COUNT + SUNAMD_BS_FLAG] < 0) { rc = -1; }
      }
    }
    free(hy);
    free(hs);
  }
  return rc;
}

extern "C" int SUNAMD_BatchedBDFSetOccupancy(SUNAMD_BatchedBDF* B, const int* waves)
{
  if (!B || !waves) { return -1; }
  for (int k = 0; k < B->nb; k++)
  {
    if (waves[k] < 1 || waves[k] > 4) { return -1; }
  }
  for (int k = 0; k < B->nb; k++) { B->waves[k] = waves[k]; }
  return 0;
}

extern "C" int SUNAMD_BatchedBDFSetGridCap(SUNAMD_BatchedBDF* B, const int* per_cu)
{
  if (!B || !per_cu) { return -1; }
  for (int k = 0; k < B->nb; k++)
  {
    if (per_cu[k] < 1 || per_cu[k] > 64) { return -1; }
  }
  for (int k = 0; k < B->nb; k++) { B->grid_cap[k] = per_cu[k]; }
  return 0;
}

extern "C" int SUNAMD_BatchedBDFSetLearnedPredictor(SUNAMD_BatchedBDF* B, const void* weights,
                                                    int bucket_mask)
{
  if (!B) { return -1; }
  B->mlp          = weights;
  B->learned_mask = bucket_mask;
  return 0;
}

extern "C" int SUNAMD_BatchedBDFSetTrainingLog(SUNAMD_BatchedBDF* B, float* x, float* t, float* mask,
                                               long capacity, int stride)
{
  if (!B || capacity < 0 || stride < 1) { return -1; }
  if (capacity > 0 && (!x || !t || !mask)) { return -1; }
  B->log_x      = x;
  B->log_t      = t;
  B->log_m      = mask;
  B->log_cap    = capacity;
  B->log_stride = stride;
  hipError_t e  = hipMemset(B->dcount, 0, sizeof(unsigned long long));
  return e == hipSuccess ? 0 : hfail(e, "log reset");
}

extern "C" long int SUNAMD_BatchedBDFGetTrainingLogCount(SUNAMD_BatchedBDF* B)
{
  unsigned long long n = 0;
  if (!B || hipMemcpy(&n, B->dcount, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess) { return -1; }
  return (long)(n < (unsigned long long)B->log_cap ? n : (unsigned long long)B->log_cap);
}

extern "C" int SUNAMD_BatchedBDFGetLearnedStats(SUNAMD_BatchedBDF* B, long int* tried,
                                                long int* kept)
{
  unsigned long long c[1 + 2 * SUNAMD_BATCH_MAXB];
  if (!B || hipMemcpy(c, B->dcount, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) { return -1; }
  long tt = 0, kk = 0;
  for (int k = 0; k < B->nb; k++)
  {
    tt += (long)c[1 + 2 * k];
    kk += (long)c[2 + 2 * k];
  }
  if (tried) { *tried = tt; }
  if (kept) { *kept = kk; }
  return 0;
}

extern "C" int SUNAMD_BatchedBDFGetLearnedStatsByBucket(SUNAMD_BatchedBDF* B, long int* tried,
                                                        long int* kept)
{
  unsigned long long c[1 + 2 * SUNAMD_BATCH_MAXB];
  if (!B || !tried || !kept ||
      hipMemcpy(c, B->dcount, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
  {
    return -1;
  }
  for (int k = 0; k < B->nb; k++)
  {
    tried[k] = (long)c[1 + 2 * k];
    kept[k]  = (long)c[2 + 2 * k];
  }
  return B->nb;
}

extern "C" int SUNAMD_BatchedBDFGetKernelTimes(SUNAMD_BatchedBDF* B, double* ms)
{
  if (!B || !ms) { return -1; }
  for (int k = 0; k < B->nb; k++) { ms[k] = B->b[k].last_ms; }
  return B->nb;
}

extern "C" int SUNAMD_BatchedBDFGetPhaseTicks(SUNAMD_BatchedBDF* B, unsigned long long* out)
{
  if (!B || !out || !B->dphase) { return -1; }
  hipError_t e = hipMemcpy(out, B->dphase,
                           sizeof(unsigned long long) * (size_t)B->nb * SUNAMD_BDF_PHASE_VALUES,
                           hipMemcpyDeviceToHost);
  return e == hipSuccess ? B->nb : hfail(e, "phase copy");
}

extern "C" void SUNAMD_BatchedBDFFree(SUNAMD_BatchedBDF* B)
{
  if (!B) { return; }
  (void)hipDeviceSynchronize();
  for (int k = 0; k < B->nb; k++)
  {
    BucketDev& d = B->b[k];
    if (d.idx) { (void)hipFree(d.idx); }
    if (d.soff) { (void)hipFree(d.soff); }
    if (d.sent) { (void)hipFree(d.sent); }
    if (d.kf) { (void)hipFree(d.kf); }
    if (d.kb) { (void)hipFree(d.kb); }
    if (d.y0) { (void)hipFree(d.y0); }
    if (d.yout) { (void)hipFree(d.yout); }
    if (d.stats) { (void)hipFree(d.stats); }
    if (d.next) { (void)hipFree(d.next); }
    if (d.ev0) { (void)hipEventDestroy(d.ev0); }
    if (d.ev1) { (void)hipEventDestroy(d.ev1); }
    if (d.stream) { (void)hipStreamDestroy(d.stream); }
  }
  if (B->dcount) { (void)hipFree(B->dcount); }
  if (B->dphase) { (void)hipFree(B->dphase); }
  free(B->sys_bucket);
  free(B->sys_cell);
  free(B->sys_yoff);
  free(B);
}

// -----------------------------------------------------------------------------
// Allocation-free integration of one size bucket
// -----------------------------------------------------------------------------
// The caller owns every input and output buffer; the only device memory beyond
// them is this file-scope array: the work-queue ticket, the validation flag,
// the reaction table as validated, and the per-species term lists the kernel
// reads.  A one-workgroup kernel fills it on the caller's stream, so a call
// touches no host memory after its argument checks.  Each device holds its own
// copy; calls that overlap on one device share it and race.
namespace
{
constexpr int NA_NEXT  = 0;                            // 2 ints: u64 work-queue ticket
constexpr int NA_BAD   = 2;                            // 1 int: table failed validation
constexpr int NA_SOFF  = 64;                           // up to 72 ints: m+1 offsets
constexpr int NA_IDX   = 136;                          // 4R ints: reaction table (validated)
constexpr int NA_SENT  = NA_IDX + 4 * SUNAMD_NOALLOC_MAX_REACTIONS; // 4R ints: term lists
constexpr int NA_INTS  = NA_SENT + 4 * SUNAMD_NOALLOC_MAX_REACTIONS;
constexpr int MAX_DEVS = 64;
} // namespace

static __device__ __align__(256) int scratch[NA_INTS];

static std::mutex g_scratch_mu;
static int* g_scratch_ptr[MAX_DEVS];

// device address of `scratch` on the current device (looked up once per device)
static int* scratch_address()
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
  {
    hfail(e, "hipGetDevice");
    return nullptr;
  }
  if (dev < 0 || dev >= MAX_DEVS) { return nullptr; }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr)
  {
    void* p = nullptr;
    e       = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
    {
      hfail(e, "hipGetSymbolAddress(scratch)");
      return nullptr;
    }
    g_scratch_ptr[dev] = (int*)p;
  }
  return g_scratch_ptr[dev];
}

// Validates the reaction table (every index in [0, m)), copies it into the
// scratch, and builds species s's term list (4*reaction + role, in reaction
// and role order: the order the host integrator accumulates f_s and row s of
// the Jacobian in).  An invalid table leaves zeros in the copy and sets the
// flag, which makes the integrator return every system unintegrated.
__global__ void __launch_bounds__(64) k_noalloc_prepare(const int* __restrict__ idx, int m, int R,
                                                        int* __restrict__ sc)
{
  __shared__ int s_idx[4 * SUNAMD_NOALLOC_MAX_REACTIONS];
  __shared__ int s_sent[4 * SUNAMD_NOALLOC_MAX_REACTIONS];
  __shared__ int s_off[72];
  __shared__ int s_bad;
  const int t = threadIdx.x;
  if (t == 0) { s_bad = 0; }
  __syncthreads();
  for (int e = t; e < 4 * R; e += 64)
  {
    const int v = idx[e];
    const bool ok = (v >= 0) && (v < m);
    s_idx[e]      = ok ? v : 0;
    if (!ok) { atomicOr(&s_bad, 1); }
  }
  __syncthreads();
  if (t == 0)
  {
    for (int s = 0; s <= m; s++) { s_off[s] = 0; }
    if (s_bad == 0)
    {
      for (int e = 0; e < 4 * R; e++) { s_off[s_idx[e] + 1]++; }
      for (int s = 0; s < m; s++) { s_off[s + 1] += s_off[s]; }
      // s_sent doubles as the fill cursor's home: cursors in s_off shifted back
      for (int e = 0; e < 4 * R; e++)
      {
        const int s        = s_idx[e];
        s_sent[s_off[s]++] = e;
      }
      for (int s = m; s > 0; s--) { s_off[s] = s_off[s - 1]; }
      s_off[0] = 0;
    }
    else
    {
      for (int e = 0; e < 4 * R; e++) { s_sent[e] = 0; }
    }
    sc[NA_NEXT]     = 0;
    sc[NA_NEXT + 1] = 0;
    sc[NA_BAD]      = s_bad;
  }
  __syncthreads();
  for (int e = t; e < 4 * R; e += 64)
  {
    sc[NA_IDX + e]  = s_idx[e];
    sc[NA_SENT + e] = s_sent[e];
  }
  for (int s = t; s <= m; s += 64) { sc[NA_SOFF + s] = s_off[s]; }
}

extern "C" int SUNAMD_BatchedBDFSolveBucketNoAlloc(int m, int R, int64_t ncells, const int* idx,
                                                    const double* kf, const double* kb,
                                                    const double* y0, double rtol, double atol,
                                                    double tout, long int mxsteps,
                                                    int lu_precision, double* yout,
                                                    long int* counters, void* stream)
{
  static const int sizes[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16, 24, 32, 48, 64};
  bool known = false;
  for (int s : sizes) { known = known || (s == m); }
  if (!known || R < 1 || R > SUNAMD_NOALLOC_MAX_REACTIONS || ncells < 0 ||
      (lu_precision != SUNAMD_LU_FP64 && lu_precision != SUNAMD_LU_FP32))
  {
    fprintf(stderr,
            "[sundials_amd batched_bdf] no-alloc solve: unsupported m = %d / R = %d (1..%d) / "
            "ncells = %lld / precision %d\n",
            m, R, SUNAMD_NOALLOC_MAX_REACTIONS, (long long)ncells, lu_precision);
    return -1;
  }
  const int gpw = 64 / pow2ceil(m);
  if (sunamd_bdf::bucket_lds_bytes(m, R, gpw, false, false,
                                   lu_precision == SUNAMD_LU_FP32 ? 4 : 8) > 60000)
  {
    fprintf(stderr, "[sundials_amd batched_bdf] no-alloc solve: m=%d R=%d needs too much LDS\n",
            m, R);
    return -1;
  }
  if (ncells == 0) { return 0; }
  if (!idx || !kf || !kb || !y0 || !yout || !counters) { return -1; }
  int* sc = scratch_address();
  if (sc == nullptr) { return -2; }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
    {
      cus = v;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  k_noalloc_prepare<<<1, 64, 0, st>>>(idx, m, R, sc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { return hfail(e, "no-alloc prepare launch") - 1; }

  BucketArgs A;
  memset(&A, 0, sizeof(A));
  A.m      = m;
  A.R      = R;
  A.ncells = (long)ncells;
  A.idx    = sc + NA_IDX;
  A.soff   = sc + NA_SOFF;
  A.sent   = sc + NA_SENT;
  A.kf     = kf;
  A.kb     = kb;
  A.y0     = y0;
  A.yout   = yout;
  A.stats  = counters;
  A.next   = (unsigned long long*)(sc + NA_NEXT);
  A.bad    = sc + NA_BAD;
  A.rtol   = rtol;
  A.atol   = atol;
  A.tout   = tout;
  A.mxstep = (mxsteps == 0) ? MXSTEP_DEFAULT : mxsteps;
  const long want = (A.ncells + gpw - 1) / gpw;
  const long cap  = (long)cus * 16;
  const int grid  = (int)(want < cap ? want : cap);
  // two wavefronts per SIMD: the register budget the autotuner picks for
  // every bucket of the benchmark batch
  switch (m)
  {
#define BDF_CASE(MM)                                                                     \
  case MM:                                                                               \
    if (lu_precision == SUNAMD_LU_FP32) { sunamd_bdf::launch<MM, float>(A, grid, st, 2); } \
    else { sunamd_bdf::launch<MM, double>(A, grid, st, 2); }                             \
    break;
    BDF_SIZES(BDF_CASE)
#undef BDF_CASE
  default: return -1;
  }
  e = hipGetLastError();
  if (e != hipSuccess) { return hfail(e, "no-alloc batched BDF launch") - 1; }
  return 0;
}
