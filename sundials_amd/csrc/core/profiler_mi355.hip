// -----------------------------------------------------------------------------
// HIP-event kernel timers and the N_Vector microbenchmark.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "profiler_internal.h"
#include "sunamd/nvector_mi355.h"
#include "sunamd/sunamd_profiler.h"

namespace
{
struct Family
{
  std::vector<hipEvent_t> start, stop; // pairs recorded since the last read
  std::vector<hipEvent_t> pool;
  double total_ms = 0.0;
  long count      = 0;
};

bool g_on = false;
Family g_fam[SUNAMD_K_COUNT];

hipEvent_t take(Family& f)
{
  if (!f.pool.empty())
  {
    hipEvent_t e = f.pool.back();
    f.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) { return nullptr; }
  return e;
}

void drain(Family& f)
{
  for (size_t i = 0; i < f.start.size(); i++)
  {
    float ms = 0.0f;
    if (hipEventSynchronize(f.stop[i]) == hipSuccess &&
        hipEventElapsedTime(&ms, f.start[i], f.stop[i]) == hipSuccess)
    {
      f.total_ms += ms;
      f.count++;
    }
    f.pool.push_back(f.start[i]);
    f.pool.push_back(f.stop[i]);
  }
  f.start.clear();
  f.stop.clear();
}
} // namespace

void sunamd_prof_begin(int fam, hipStream_t s)
{
  if (!g_on) { return; }
  Family& f = g_fam[fam];
  if (f.start.size() >= 4096) { drain(f); }
  hipEvent_t e = take(f);
  if (!e) { return; }
  (void)hipEventRecord(e, s);
  f.start.push_back(e);
}

void sunamd_prof_end(int fam, hipStream_t s)
{
  if (!g_on) { return; }
  Family& f = g_fam[fam];
  if (f.stop.size() >= f.start.size()) { return; }
  hipEvent_t e = take(f);
  if (!e)
  {
    f.pool.push_back(f.start.back());
    f.start.pop_back();
    return;
  }
  (void)hipEventRecord(e, s);
  f.stop.push_back(e);
}

extern "C" void SUNAMD_ProfilerEnable(int on) { g_on = (on != 0); }

extern "C" void SUNAMD_ProfilerReset(void)
{
  for (auto& f : g_fam)
  {
    drain(f);
    f.total_ms = 0.0;
    f.count    = 0;
  }
}

extern "C" int SUNAMD_ProfilerGet(int fam, double* total_ms, long int* count)
{
  if (fam < 0 || fam >= SUNAMD_K_COUNT) { return -1; }
  drain(g_fam[fam]);
  *total_ms = g_fam[fam].total_ms;
  *count    = g_fam[fam].count;
  return 0;
}

// ----------------------------------------------------------------------------
// microbenchmark
// ----------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long mix64d(unsigned long long z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__global__ void k_fill_uniform(double* __restrict__ v, long n, unsigned long long seed, double lo,
                               double hi)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    unsigned long long z = mix64d(seed * 0x9E3779B97F4A7C15ULL + mix64d((unsigned long long)i));
    double u             = ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    v[i]                 = lo + (hi - lo) * u;
  }
}

// Copy ceiling: z = x in the vector kernels' chunked form (each lane's U
// 128-bit non-temporal loads issued before its stores), B threads per
// workgroup, at most cap workgroups per CU.
template <int U, int B>
__global__ void __launch_bounds__(B) k_copy_chunks(const double* __restrict__ x,
                                                   double* __restrict__ z, long np)
{
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2* __restrict__ xs = reinterpret_cast<const d2*>(x);
  d2* __restrict__ zs       = reinterpret_cast<d2*>(z);
  const long chunk          = (long)U * B;
  for (long c = (long)blockIdx.x * chunk; c < np; c += (long)gridDim.x * chunk)
  {
    const long i0 = c + threadIdx.x;
    if (c + chunk <= np)
    {
      d2 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) { v[u] = __builtin_nontemporal_load(xs + i0 + (long)u * B); }
#pragma unroll
      for (int u = 0; u < U; u++) { __builtin_nontemporal_store(v[u], zs + i0 + (long)u * B); }
    }
    else
    {
      for (long i = i0; i < np; i += B) { __builtin_nontemporal_store(__builtin_nontemporal_load(xs + i), zs + i); }
    }
  }
}

template <int U, int B>
static void launch_copy(const double* x, double* z, long n, int cap_per_cu, hipStream_t s)
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
  {
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  }
  const long np = n / 2, chunk = (long)U * B;
  long g        = (np + chunk - 1) / chunk;
  const long cap = (long)cus * cap_per_cu;
  if (g > cap) { g = cap; }
  if (g < 1) { g = 1; }
  k_copy_chunks<U, B><<<(int)g, B, 0, s>>>(x, z, np);
}

extern "C" int SUNAMD_BenchVectorOps(sunindextype n, int nvec, int reps, double* out_ms)
{
  if (n <= 0 || (n & 1) || nvec < 1 || nvec > 32 || reps < 1) { return -1; }
  SUNContext ctx = NULL;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) { return -1; }
  int rc = -1;
  std::vector<N_Vector> Y;
  N_Vector x = N_VNew_Mi355(n, ctx);
  N_Vector y = x ? N_VClone(x) : NULL;
  N_Vector z = x ? N_VClone(x) : NULL;
  N_Vector w = x ? N_VClone(x) : NULL;
  hipStream_t s = x ? (hipStream_t)N_VGetStream_Mi355(x) : 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double d[32];
  if (!x || !y || !z || !w) { goto done; }
  for (int k = 0; k < nvec; k++)
  {
    N_Vector v = N_VClone(x);
    if (!v) { goto done; }
    Y.push_back(v);
  }
  {
    auto fill = [&](N_Vector v, unsigned long long seed, double lo, double hi) {
      k_fill_uniform<<<4096, 256, 0, s>>>(N_VGetDeviceArrayPointer(v), (long)n, seed, lo, hi);
    };
    fill(x, 1, -1.0, 1.0);
    fill(y, 2, -1.0, 1.0);
    fill(w, 3, 1e-3, 1.0);
    for (int k = 0; k < nvec; k++) { fill(Y[k], 10 + k, -1.0, 1.0); }
    if (hipStreamSynchronize(s) != hipSuccess) { goto done; }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { goto done; }
    double* xd = N_VGetDeviceArrayPointer(x);
    double* zd = N_VGetDeviceArrayPointer(z);
    for (int op = 0; op < 8; op++)
    {
      auto run = [&]() {
        switch (op)
        {
        case 0: N_VLinearSum(0.3, x, 0.7, y, z); break;
        case 1: (void)N_VWrmsNorm(x, w); break;
        case 2: (void)N_VDotProdMulti(nvec, x, Y.data(), d); break;
        case 3: N_VLinearSum(1.0, x, 1.0, y, y); break;
        case 4: (void)N_VDotProd(x, y); break;
        case 5: launch_copy<8, 256>(xd, zd, (long)n, 32, s); break;
        case 6: launch_copy<8, 1024>(xd, zd, (long)n, 8, s); break;
        case 7: (void)hipMemcpyAsync(zd, xd, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s); break;
        }
      };
      run(); // warm-up
      if (hipStreamSynchronize(s) != hipSuccess) { goto done; }
      // reductions synchronise inside each call; the event pair brackets the
      // whole timed sequence on the vector stream
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps; r++) { run(); }
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) { goto done; }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      out_ms[op] = (double)ms / reps;
    }
    rc = 0;
  }
done:
  if (e0) { (void)hipEventDestroy(e0); }
  if (e1) { (void)hipEventDestroy(e1); }
  for (auto v : Y) { N_VDestroy(v); }
  if (x) { N_VDestroy(x); }
  if (y) { N_VDestroy(y); }
  if (z) { N_VDestroy(z); }
  if (w) { N_VDestroy(w); }
  SUNContext_Free(&ctx);
  return rc;
}
