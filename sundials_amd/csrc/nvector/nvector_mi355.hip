// -----------------------------------------------------------------------------
// N_Vector for MI355X (gfx950): every op is a hand-written HIP kernel.
//
// Ops-table semantics follow the reference (include/sundials/sundials_nvector.h
// :101-195; per-op formulas from src/nvector/serial/nvector_serial.c:387-1592).
// Design, replacing the reference HIP vector (src/nvector/hip/nvector_hip.hip.cpp,
// VectorKernels.hip.hpp, VectorArrayKernels.hip.hpp):
//  * streaming ops move 128-bit (2 x FP64) per lane per access, non-temporal,
//    in contiguous per-workgroup chunks with several loads in flight per
//    lane before the first use; grid sized to the device;
//  * LinearSum keeps the serial vector's rounding: its nine special cases
//    collapse to three per-element formulas, a*x+b*y, a*(x+y), a*(x-y)
//    (nvector_serial.c:387-480 / :1600-1775), with no fused multiply-add;
//  * reductions: per-lane sequential accumulation, 64-wide wavefront
//    shuffle tree, LDS across wavefronts, per-workgroup partials; a
//    one-workgroup finish kernel folds them in workgroup order and writes the
//    scalar(s) straight into pinned host memory -- one stream synchronisation
//    per reduction, deterministic results;
//  * DotProdMulti reads x once for all nvec dot products (the serial vector
//    re-reads x per product, nvector_serial.c:1019-1024) and spreads the work
//    over the whole device (the reference uses one workgroup per product);
//  * LinearCombination / ScaleAddMulti write each output once; vector
//    pointers and coefficients travel as kernel arguments (no per-call
//    hipMalloc / hipMemcpy of pointer arrays).
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../core/profiler_internal.h"
#include "sunamd/nvector_mi355.h"
#include "sunamd/sunamd_profiler.h"
#include "sunamd/sundials_iterative.h"

#define NV_MAX_MULTI 32      // vectors per fused call handled in one kernel
#define NV_BLOCK 256         // threads per workgroup
#define NV_MAX_RED_BLOCKS 2048

typedef double dbl2 __attribute__((ext_vector_type(2)));

// ----------------------------------------------------------------------------
// execution context shared by a vector and all of its clones
// ----------------------------------------------------------------------------
struct Mi355Exec
{
  int refcount;
  int device;
  int num_cus;
  hipStream_t stream;
  double* d_partials;     // NV_MAX_RED_BLOCKS * (NV_MAX_MULTI + 1)
  double* h_result;       // pinned, mapped: NV_MAX_MULTI + 1 scalars
  double* d_gs;           // Gram-Schmidt coefficients on the device (lazy, GS_MAX + 2)
  double* h_gs;           // pinned copy of them (lazy)
  long int launches;
  hipError_t err;         // first HIP failure on this context since it was last read
  bool own_stream;
  bool is_default;
};

struct Mi355Content
{
  sunindextype length;
  sunbooleantype own_host;
  sunbooleantype own_dev;
  sunrealtype* host;
  sunrealtype* dev;
  Mi355Exec* exec;
  int red_order; // SUNAMD_REDUCE_TREE or SUNAMD_REDUCE_SEQUENTIAL, inherited by clones
};

#define CONTENT(v) ((Mi355Content*)((v)->content))
#define NVLEN(v)   (CONTENT(v)->length)
#define DDATA(v)   (CONTENT(v)->dev)
#define EXEC(v)    (CONTENT(v)->exec)

// HIP failures are reported on stderr and through return codes; the library
// never writes into the SUNContext, whose layout belongs to the SUNDIALS core
// it is linked with.
static void hip_fail(SUNContext ctx, hipError_t e, const char* what, int line)
{
  (void)ctx;
  fprintf(stderr, "[sundials_amd nvector_mi355] %s failed (line %d): %s\n", what, line,
          hipGetErrorString(e));
  fflush(stderr);
}

#define HIPCHK(ctx, call)                                             \
  do {                                                                \
    hipError_t e_ = (call);                                           \
    if (e_ != hipSuccess) { hip_fail((ctx), e_, #call, __LINE__); }   \
  } while (0)

// the same, also kept as the context's sticky error: void ops cannot return
// it, so the next reduction returns NaN, the next op returning SUNErrCode
// returns SUN_ERR_EXT_FAIL, and N_VGetHipError_Mi355 reports it
static inline void exec_record(Mi355Exec* ex, hipError_t e)
{
  if (ex && e != hipSuccess && ex->err == hipSuccess) { ex->err = e; }
}

#define HIPCHK_EX(v, call)                                            \
  do {                                                                \
    hipError_t e_ = (call);                                           \
    if (e_ != hipSuccess)                                             \
    {                                                                 \
      hip_fail((v)->sunctx, e_, #call, __LINE__);                     \
      exec_record(EXEC(v), e_);                                       \
    }                                                                 \
  } while (0)

#define LAUNCHCHK(v)                                                  \
  do {                                                                \
    EXEC(v)->launches++;                                              \
    hipError_t e_ = hipGetLastError();                                \
    if (e_ != hipSuccess)                                             \
    {                                                                 \
      hip_fail((v)->sunctx, e_, "kernel launch", __LINE__);           \
      exec_record(EXEC(v), e_);                                       \
    }                                                                 \
  } while (0)

static inline SUNErrCode exec_status(const Mi355Exec* ex)
{
  return (ex->err == hipSuccess) ? SUN_SUCCESS : SUN_ERR_EXT_FAIL;
}

// One execution context per device is shared by every vector created without
// an explicit stream, so that all vector, matrix and solver work issued by one
// host thread is ordered on a single stream (as the reference HIP vector's
// default stream orders it).  A caller-supplied stream gets its own context.
static Mi355Exec* g_default_exec[64];

static Mi355Exec* exec_create(SUNContext ctx, hipStream_t user_stream, bool own_stream)
{
  Mi355Exec* ex = (Mi355Exec*)calloc(1, sizeof(Mi355Exec));
  if (!ex) { return NULL; }
  ex->refcount   = 1;
  ex->own_stream = own_stream;
  HIPCHK(ctx, hipGetDevice(&ex->device));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ex->device) == hipSuccess)
  {
    ex->num_cus = prop.multiProcessorCount;
  }
  if (ex->num_cus <= 0) { ex->num_cus = 256; }
  hipError_t e = hipSuccess;
  if (own_stream) { e = hipStreamCreateWithFlags(&ex->stream, hipStreamNonBlocking); }
  else { ex->stream = user_stream; }
  if (e != hipSuccess)
  {
    hip_fail(ctx, e, "hipStreamCreate", __LINE__);
    free(ex);
    return NULL;
  }
  e = hipMalloc((void**)&ex->d_partials,
                sizeof(double) * NV_MAX_RED_BLOCKS * (NV_MAX_MULTI + 1));
  if (e == hipSuccess)
  {
    e = hipHostMalloc((void**)&ex->h_result, sizeof(double) * (NV_MAX_MULTI + 1),
                      hipHostMallocMapped);
  }
  if (e != hipSuccess)
  {
    hip_fail(ctx, e, "exec allocation", __LINE__);
    if (ex->d_partials) { (void)hipFree(ex->d_partials); }
    if (own_stream) { (void)hipStreamDestroy(ex->stream); }
    free(ex);
    return NULL;
  }
  return ex;
}

static Mi355Exec* exec_default(SUNContext ctx)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) { return NULL; }
  if (g_default_exec[dev])
  {
    g_default_exec[dev]->refcount++;
    return g_default_exec[dev];
  }
  Mi355Exec* ex = exec_create(ctx, 0, true);
  if (ex) { ex->is_default = true; }
  g_default_exec[dev] = ex;
  return ex;
}

static void exec_release(Mi355Exec* ex)
{
  if (!ex) { return; }
  if (--ex->refcount > 0) { return; }
  (void)hipStreamSynchronize(ex->stream);
  (void)hipFree(ex->d_partials);
  (void)hipHostFree(ex->h_result);
  if (ex->d_gs) { (void)hipFree(ex->d_gs); }
  if (ex->h_gs) { (void)hipHostFree(ex->h_gs); }
  if (ex->own_stream) { (void)hipStreamDestroy(ex->stream); }
  if (ex->is_default && ex->device >= 0 && ex->device < 64 &&
      g_default_exec[ex->device] == ex)
  {
    g_default_exec[ex->device] = NULL;
  }
  free(ex);
}

// ----------------------------------------------------------------------------
// launch geometry
// ----------------------------------------------------------------------------
// Element pairs a lane handles per chunk for a kernel that issues `loads`
// 128-bit loads per pair (measured sweeps over 6.3M-2^28 elements): 8 for one
// or two input vectors (1 pair per lane per iteration leaves ~20 % of the
// bandwidth unused), 4 for more -- at 2^28 a nine-vector dot product moves
// 6.7 TB/s with 4 pairs per lane against 6.2-6.4 TB/s with 1.
__host__ __device__ constexpr int unroll_for(int loads)
{
  return loads <= 2 ? 8 : 4;
}

// streaming: chunks of U * NV_BLOCK pairs per workgroup, grid capped at 32
// workgroups per CU and chunk-strided beyond that
static inline int stream_grid(const Mi355Exec* ex, sunindextype npairs, int U)
{
  const sunindextype chunk = (sunindextype)U * NV_BLOCK;
  sunindextype g           = (npairs + chunk - 1) / chunk;
  sunindextype cap         = (sunindextype)ex->num_cus * 32;
  if (g > cap) { g = cap; }
  if (g < 1) { g = 1; }
  return (int)g;
}

// The reduction grid depends only on the vector length and the kernel (not on
// the device), so the summation order -- and the rounded result -- is a fixed
// function of n.
static inline int red_grid(const Mi355Exec* ex, sunindextype npairs, int U)
{
  (void)ex;
  const sunindextype chunk = (sunindextype)U * NV_BLOCK;
  sunindextype g           = (npairs + chunk - 1) / chunk;
  if (g > NV_MAX_RED_BLOCKS) { g = NV_MAX_RED_BLOCKS; }
  if (g < 1) { g = 1; }
  return (int)g;
}

// ----------------------------------------------------------------------------
// device helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ dbl2 ld2(const double* __restrict__ p, long i)
{
  return __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(p) + i);
}
__device__ __forceinline__ void st2(double* __restrict__ p, long i, dbl2 v)
{
  __builtin_nontemporal_store(v, reinterpret_cast<dbl2*>(p) + i);
}

enum RedOp
{
  RSUM = 0,
  RMAX = 1,
  RMIN = 2
};

template <int OP>
__device__ __forceinline__ double rcomb(double a, double b)
{
  if (OP == RSUM) { return a + b; }
  if (OP == RMAX) { return (b > a) ? b : a; }
  return (b < a) ? b : a;
}

template <int OP>
__device__ __forceinline__ double wave_reduce(double v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) { v = rcomb<OP>(v, __shfl_xor(v, off, 64)); }
  return v;
}

// reduce NV per-lane values over the workgroup; result valid in thread 0
template <int OP, int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double* lds /* NV*4 */)
{
  const int lane = threadIdx.x & 63;
  const int wid  = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++)
  {
    double r = wave_reduce<OP>(v[k]);
    if (lane == 0) { lds[k * (NV_BLOCK / 64) + wid] = r; }
  }
  __syncthreads();
  if (threadIdx.x == 0)
  {
#pragma unroll
    for (int k = 0; k < NV; k++)
    {
      double r = lds[k * (NV_BLOCK / 64)];
      for (int w = 1; w < NV_BLOCK / 64; w++) { r = rcomb<OP>(r, lds[k * (NV_BLOCK / 64) + w]); }
      v[k] = r;
    }
  }
}

// Each workgroup publishes its NV partials; k_finish, launched next on the
// same stream, folds them in workgroup order and writes the results to `out`
// (pinned host memory or device memory).  Two launches instead of a
// last-workgroup ticket: the kernel boundary orders the partials without a
// device-scope fence and atomic per workgroup, which serialise at the end of
// every reduction.
template <int OP, int NV>
__device__ __forceinline__ void grid_finish(double (&v)[NV], double* partials)
{
  __shared__ double lds[NV * (NV_BLOCK / 64)];
  block_reduce<OP, NV>(v, lds);
  if (threadIdx.x == 0)
  {
#pragma unroll
    for (int k = 0; k < NV; k++) { partials[k * gridDim.x + blockIdx.x] = v[k]; }
  }
}

template <int OP, int NV>
__global__ void __launch_bounds__(NV_BLOCK) k_finish(const double* __restrict__ partials,
                                                     int nblocks, double* out, double identity)
{
  __shared__ double lds[NV * (NV_BLOCK / 64)];
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++)
  {
    acc[k] = identity;
    for (int b = threadIdx.x; b < nblocks; b += NV_BLOCK)
    {
      acc[k] = rcomb<OP>(acc[k], __builtin_nontemporal_load(&partials[k * nblocks + b]));
    }
  }
  block_reduce<OP, NV>(acc, lds);
  if (threadIdx.x == 0)
  {
#pragma unroll
    for (int k = 0; k < NV; k++) { out[k] = acc[k]; }
    __threadfence_system();
  }
}

// ----------------------------------------------------------------------------
// streaming kernels: pairs [0, n/2) vectorised, odd tail element by thread 0
// ----------------------------------------------------------------------------
// Pairs [0, np) in chunks of U * NV_BLOCK consecutive pairs, chunks strided
// over the grid; lane t of a chunk starting at c takes pairs c + t + u *
// NV_BLOCK, u < U.  A full chunk issues all of a lane's loads before the
// first use (U * loads per pair in flight, a wavefront's loads inside one
// contiguous chunk); the last, partial chunk goes pair by pair.  Each lane
// visits its pairs in increasing order, so per-lane accumulations have an
// order fixed by np and the grid.
template <int U, typename Load, typename Use>
__device__ __forceinline__ void chunk_pairs(long np, Load load, Use use)
{
  const long chunk = (long)U * NV_BLOCK;
  for (long c = (long)blockIdx.x * chunk; c < np; c += (long)gridDim.x * chunk)
  {
    const long i0 = c + threadIdx.x;
    if (c + chunk <= np)
    {
      decltype(load(0L)) v[U];
#pragma unroll
      for (int u = 0; u < U; u++) { v[u] = load(i0 + (long)u * NV_BLOCK); }
#pragma unroll
      for (int u = 0; u < U; u++) { use(i0 + (long)u * NV_BLOCK, v[u]); }
    }
    else
    {
      for (long i = i0; i < np; i += NV_BLOCK) { use(i, load(i)); }
    }
  }
}

template <int K>
struct Pairs
{
  dbl2 v[K];
};

#define TAIL(n) if (((n) & 1) && blockIdx.x == 0 && threadIdx.x == 0)

enum LSMode
{
  LS_GENERIC  = 0, // a*x + b*y
  LS_SCALESUM = 1, // a*(x+y)
  LS_SCALEDIF = 2  // a*(x-y)
};

template <int MODE>
__global__ void __launch_bounds__(NV_BLOCK) k_linearsum(double a, const double* __restrict__ x,
                                                        double b, const double* __restrict__ y,
                                                        double* __restrict__ z, long n)
{
  const long np = n >> 1;
  chunk_pairs<unroll_for(2)>(
    np, [&](long i) { return Pairs<2>{{ld2(x, i), ld2(y, i)}}; },
    [&](long i, const Pairs<2>& p) {
      dbl2 r;
      if (MODE == LS_GENERIC) { r = (a * p.v[0]) + (b * p.v[1]); }
      else if (MODE == LS_SCALESUM) { r = a * (p.v[0] + p.v[1]); }
      else { r = a * (p.v[0] - p.v[1]); }
      st2(z, i, r);
    });
  TAIL(n)
  {
    double xv = x[n - 1], yv = y[n - 1];
    if (MODE == LS_GENERIC) { z[n - 1] = (a * xv) + (b * yv); }
    else if (MODE == LS_SCALESUM) { z[n - 1] = a * (xv + yv); }
    else { z[n - 1] = a * (xv - yv); }
  }
}

__global__ void __launch_bounds__(NV_BLOCK) k_const(double c, double* __restrict__ z, long n)
{
  const long np = n >> 1;
  dbl2 cv       = {c, c};
  const long stride = (long)gridDim.x * NV_BLOCK;
  for (long i = (long)blockIdx.x * NV_BLOCK + threadIdx.x; i < np; i += stride) { st2(z, i, cv); }
  TAIL(n) { z[n - 1] = c; }
}

// unary / binary elementwise family
enum EwOp
{
  EW_PROD,
  EW_DIV,
  EW_SCALE,
  EW_ABS,
  EW_INV,
  EW_ADDCONST,
  EW_COMPARE
};

template <int OP>
__device__ __forceinline__ double ew(double x, double y, double c)
{
  if (OP == EW_PROD) { return x * y; }
  if (OP == EW_DIV) { return x / y; }
  if (OP == EW_SCALE) { return c * x; }
  if (OP == EW_ABS) { return fabs(x); }
  if (OP == EW_INV) { return 1.0 / x; }
  if (OP == EW_ADDCONST) { return x + c; }
  return (fabs(x) >= c) ? 1.0 : 0.0;
}

template <int OP>
__global__ void __launch_bounds__(NV_BLOCK) k_ew(const double* __restrict__ x,
                                                 const double* __restrict__ y, double c,
                                                 double* __restrict__ z, long n)
{
  const bool binary = (OP == EW_PROD || OP == EW_DIV);
  const long np     = n >> 1;
  constexpr int NL  = binary ? 2 : 1;
  chunk_pairs<unroll_for(NL)>(
    np,
    [&](long i) {
      Pairs<NL> p;
      p.v[0] = ld2(x, i);
      if (binary) { p.v[NL - 1] = ld2(y, i); }
      return p;
    },
    [&](long i, const Pairs<NL>& p) {
      dbl2 r;
      r.x = ew<OP>(p.v[0].x, p.v[NL - 1].x, c);
      r.y = ew<OP>(p.v[0].y, p.v[NL - 1].y, c);
      st2(z, i, r);
    });
  TAIL(n) { z[n - 1] = ew<OP>(x[n - 1], binary ? y[n - 1] : 0.0, c); }
}

// ----------------------------------------------------------------------------
// reduction kernels
// ----------------------------------------------------------------------------
enum RedKind
{
  RK_DOT,      // sum x*y
  RK_WSQR,     // sum (x*w)^2
  RK_WSQRMASK, // sum (x*w)^2 where id > 0
  RK_MAXABS,   // max |x|
  RK_MIN,      // min x
  RK_L1        // sum |x|
};

template <int K>
struct RedTraits;
template <>
struct RedTraits<RK_DOT>
{
  static const int op = RSUM;
};
template <>
struct RedTraits<RK_WSQR>
{
  static const int op = RSUM;
};
template <>
struct RedTraits<RK_WSQRMASK>
{
  static const int op = RSUM;
};
template <>
struct RedTraits<RK_MAXABS>
{
  static const int op = RMAX;
};
template <>
struct RedTraits<RK_MIN>
{
  static const int op = RMIN;
};
template <>
struct RedTraits<RK_L1>
{
  static const int op = RSUM;
};

template <int K>
__device__ __forceinline__ double red_elem(double x, double y, double id)
{
  if (K == RK_DOT) { return x * y; }
  if (K == RK_WSQR)
  {
    double p = x * y;
    return p * p;
  }
  if (K == RK_WSQRMASK)
  {
    double p = x * y;
    return (id > 0.0) ? p * p : 0.0;
  }
  if (K == RK_MAXABS) { return fabs(x); }
  if (K == RK_MIN) { return x; }
  return fabs(x);
}

// 128-bit loads per element pair of reduction K
template <int K>
__host__ __device__ constexpr int red_loads()
{
  return K == RK_WSQRMASK ? 3 : ((K == RK_DOT || K == RK_WSQR) ? 2 : 1);
}

template <int K>
__global__ void __launch_bounds__(NV_BLOCK) k_reduce(const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     const double* __restrict__ id, long n,
                                                     double identity, double* partials)
{
  const int OP     = RedTraits<K>::op;
  const bool two   = (K == RK_DOT || K == RK_WSQR || K == RK_WSQRMASK);
  const bool three = (K == RK_WSQRMASK);
  double acc[1]    = {identity};
  const long np    = n >> 1;
  chunk_pairs<unroll_for(red_loads<K>())>(
    np,
    [&](long i) {
      Pairs<red_loads<K>()> p;
      p.v[0] = ld2(x, i);
      if (two) { p.v[1 % red_loads<K>()] = ld2(y, i); }
      if (three) { p.v[2 % red_loads<K>()] = ld2(id, i); }
      return p;
    },
    [&](long i, const Pairs<red_loads<K>()>& p) {
      (void)i;
      const dbl2 xv = p.v[0];
      const dbl2 yv = p.v[two ? 1 % red_loads<K>() : 0];
      const dbl2 iv = p.v[three ? 2 % red_loads<K>() : 0];
      acc[0]        = rcomb<OP>(acc[0], red_elem<K>(xv.x, yv.x, iv.x));
      acc[0]        = rcomb<OP>(acc[0], red_elem<K>(xv.y, yv.y, iv.y));
    });
  TAIL(n)
  {
    acc[0] = rcomb<OP>(acc[0], red_elem<K>(x[n - 1], two ? y[n - 1] : 0.0,
                                          three ? id[n - 1] : 0.0));
  }
  grid_finish<OP, 1>(acc, partials);
}

// Left-to-right sums in the serial vector's order (nvector_serial.c:605-761):
// the wavefront stages 64 terms at a time in LDS and lane 0 adds them in index
// order, so the rounded result equals the sequential CPU sum bit for bit.
template <int K>
__global__ void __launch_bounds__(64) k_reduce_seq(const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ id, long n,
                                                   double* out)
{
  const bool two   = (K == RK_DOT || K == RK_WSQR || K == RK_WSQRMASK);
  const bool three = (K == RK_WSQRMASK);
  __shared__ double t[64];
  double acc = 0.0;
  for (long base = 0; base < n; base += 64)
  {
    long i          = base + threadIdx.x;
    t[threadIdx.x]  = (i < n) ? red_elem<K>(x[i], two ? y[i] : 0.0, three ? id[i] : 0.0) : 0.0;
    __syncthreads();
    if (threadIdx.x == 0)
    {
      int cnt = (n - base) < 64 ? (int)(n - base) : 64;
      for (int j = 0; j < cnt; j++)
      {
        if (K == RK_WSQRMASK && !(id[base + j] > 0.0)) { continue; }
        acc = acc + t[j];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = acc; }
}

// fused d_k = x . Y_k, k < NV (x read once)
struct PtrPack
{
  const double* p[NV_MAX_MULTI];
};
struct WPtrPack
{
  double* p[NV_MAX_MULTI];
};
struct CoefPack
{
  double c[NV_MAX_MULTI];
};

// sequential multi-dot (serial order): lane k accumulates d_k over i in order
__global__ void __launch_bounds__(64) k_dotmulti_seq(const double* __restrict__ x, PtrPack Y,
                                                     int nvec, long n, double* out)
{
  __shared__ double xs[64];
  const int k = threadIdx.x;
  double acc  = 0.0;
  for (long base = 0; base < n; base += 64)
  {
    long i          = base + threadIdx.x;
    xs[threadIdx.x] = (i < n) ? x[i] : 0.0;
    __syncthreads();
    if (k < nvec)
    {
      int cnt = (n - base) < 64 ? (int)(n - base) : 64;
      for (int j = 0; j < cnt; j++) { acc = acc + xs[j] * Y.p[k][base + j]; }
    }
    __syncthreads();
  }
  if (k < nvec) { out[k] = acc; }
}

template <int NV>
__global__ void __launch_bounds__(NV_BLOCK) k_dotmulti(const double* __restrict__ x, PtrPack Y,
                                                       long n, double* partials)
{
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) { acc[k] = 0.0; }
  const long np = n >> 1;
  chunk_pairs<unroll_for(NV + 1)>(
    np,
    [&](long i) {
      Pairs<NV + 1> p;
      p.v[0] = ld2(x, i);
#pragma unroll
      for (int k = 0; k < NV; k++) { p.v[k + 1] = ld2(Y.p[k], i); }
      return p;
    },
    [&](long i, const Pairs<NV + 1>& p) {
      (void)i;
#pragma unroll
      for (int k = 0; k < NV; k++)
      {
        acc[k] += p.v[0].x * p.v[k + 1].x;
        acc[k] += p.v[0].y * p.v[k + 1].y;
      }
    });
  TAIL(n)
  {
#pragma unroll
    for (int k = 0; k < NV; k++) { acc[k] += x[n - 1] * Y.p[k][n - 1]; }
  }
  grid_finish<RSUM, NV>(acc, partials);
}

// z = c0*X0; z += c_k*X_k (k = 1..nvec-1): the rounding of the generic-layer
// loop (sundials_nvector.c:549-572), one pass, z written once
template <int NV>
__global__ void __launch_bounds__(NV_BLOCK) k_lincomb(CoefPack c, PtrPack X, double* z, long n)
{
  const long np = n >> 1;
  chunk_pairs<unroll_for(NV)>(
    np,
    [&](long i) {
      Pairs<NV> p;
#pragma unroll
      for (int k = 0; k < NV; k++) { p.v[k] = ld2(X.p[k], i); }
      return p;
    },
    [&](long i, const Pairs<NV>& p) {
      dbl2 acc = c.c[0] * p.v[0];
#pragma unroll
      for (int k = 1; k < NV; k++) { acc = acc + c.c[k] * p.v[k]; }
      st2(z, i, acc);
    });
  TAIL(n)
  {
    double acc = c.c[0] * X.p[0][n - 1];
#pragma unroll
    for (int k = 1; k < NV; k++) { acc = acc + c.c[k] * X.p[k][n - 1]; }
    z[n - 1] = acc;
  }
}

// Z_k = a_k*x + Y_k (k < NV); x read once
template <int NV>
__global__ void __launch_bounds__(NV_BLOCK) k_scaleaddmulti(CoefPack a, const double* __restrict__ x,
                                                            PtrPack Y, WPtrPack Z, long n)
{
  const long np = n >> 1;
  chunk_pairs<unroll_for(NV + 1)>(
    np,
    [&](long i) {
      Pairs<NV + 1> p;
      p.v[0] = ld2(x, i);
#pragma unroll
      for (int k = 0; k < NV; k++) { p.v[k + 1] = ld2(Y.p[k], i); }
      return p;
    },
    [&](long i, const Pairs<NV + 1>& p) {
#pragma unroll
      for (int k = 0; k < NV; k++) { st2(Z.p[k], i, (a.c[k] * p.v[0]) + p.v[k + 1]); }
    });
  TAIL(n)
  {
#pragma unroll
    for (int k = 0; k < NV; k++) { Z.p[k][n - 1] = (a.c[k] * x[n - 1]) + Y.p[k][n - 1]; }
  }
}

// Z_k = c_k * X_k (k < NV)
template <int NV>
__global__ void __launch_bounds__(NV_BLOCK) k_scalearray(CoefPack c, PtrPack X, WPtrPack Z, long n)
{
  const long np = n >> 1;
  chunk_pairs<unroll_for(NV)>(
    np,
    [&](long i) {
      Pairs<NV> p;
#pragma unroll
      for (int k = 0; k < NV; k++) { p.v[k] = ld2(X.p[k], i); }
      return p;
    },
    [&](long i, const Pairs<NV>& p) {
#pragma unroll
      for (int k = 0; k < NV; k++) { st2(Z.p[k], i, c.c[k] * p.v[k]); }
    });
  TAIL(n)
  {
#pragma unroll
    for (int k = 0; k < NV; k++) { Z.p[k][n - 1] = c.c[k] * X.p[k][n - 1]; }
  }
}

// flag/min reductions for InvTest, ConstrMask, MinQuotient
__global__ void __launch_bounds__(NV_BLOCK) k_invtest(const double* __restrict__ x,
                                                      double* __restrict__ z, long n,
                                                      double* partials)
{
  double acc[1] = {0.0}; // max of "zero found" flags
  for (long i = (long)blockIdx.x * NV_BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * NV_BLOCK)
  {
    double xv = x[i];
    if (xv == 0.0) { acc[0] = 1.0; }
    else { z[i] = 1.0 / xv; }
  }
  grid_finish<RMAX, 1>(acc, partials);
}

__global__ void __launch_bounds__(NV_BLOCK) k_constrmask(const double* __restrict__ c,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ m, long n,
                                                         double* partials)
{
  double acc[1] = {0.0};
  for (long i = (long)blockIdx.x * NV_BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * NV_BLOCK)
  {
    double ci = c[i], xi = x[i];
    double mi = 0.0;
    if (ci != 0.0)
    {
      bool test = (fabs(ci) > 1.5 && xi * ci <= 0.0) || (fabs(ci) > 0.5 && xi * ci < 0.0);
      if (test) { mi = acc[0] = 1.0; }
    }
    m[i] = mi;
  }
  grid_finish<RMAX, 1>(acc, partials);
}

__global__ void __launch_bounds__(NV_BLOCK) k_minquotient(const double* __restrict__ num,
                                                          const double* __restrict__ den, long n,
                                                          double* partials)
{
  double acc[1] = {DBL_MAX};
  for (long i = (long)blockIdx.x * NV_BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * NV_BLOCK)
  {
    double d = den[i];
    if (d != 0.0)
    {
      double q = num[i] / d;
      acc[0]   = (q < acc[0]) ? q : acc[0];
    }
  }
  grid_finish<RMIN, 1>(acc, partials);
}

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------
static N_Vector mi_clone_empty(N_Vector w);
static N_Vector mi_clone(N_Vector w);
static void mi_destroy(N_Vector v);

static N_Vector_ID mi_getid(N_Vector v)
{
  (void)v;
  return SUNDIALS_NVEC_CUSTOM;
}
static sunindextype mi_getlength(N_Vector v) { return NVLEN(v); }
static void mi_space(N_Vector v, sunindextype* lrw, sunindextype* liw)
{
  *lrw = NVLEN(v);
  *liw = 2;
}

static sunrealtype* mi_host_ptr(N_Vector v)
{
  Mi355Content* c = CONTENT(v);
  if (c->host == NULL && c->length > 0)
  {
    c->host     = (sunrealtype*)calloc((size_t)c->length, sizeof(sunrealtype));
    c->own_host = SUNTRUE;
  }
  return c->host;
}
static sunrealtype* mi_dev_ptr(N_Vector v) { return DDATA(v); }
static void mi_set_ptr(sunrealtype* d, N_Vector v)
{
  Mi355Content* c = CONTENT(v);
  if (c->own_host && c->host) { free(c->host); }
  c->host     = d;
  c->own_host = SUNFALSE;
}

static inline bool aligned16(const void* p) { return (((uintptr_t)p) & 15u) == 0; }

static void mi_linearsum(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z)
{
  long n = (long)NVLEN(z);
  if (n == 0) { return; }
  Mi355Exec* ex = EXEC(z);
  int g         = stream_grid(ex, n >> 1, unroll_for(2));
  sunamd_prof_begin(SUNAMD_K_LINSUM, ex->stream);
  if (a == b)
  {
    k_linearsum<LS_SCALESUM><<<g, NV_BLOCK, 0, ex->stream>>>(a, DDATA(x), b, DDATA(y), DDATA(z), n);
  }
  else if (a == -b)
  {
    k_linearsum<LS_SCALEDIF><<<g, NV_BLOCK, 0, ex->stream>>>(a, DDATA(x), b, DDATA(y), DDATA(z), n);
  }
  else
  {
    k_linearsum<LS_GENERIC><<<g, NV_BLOCK, 0, ex->stream>>>(a, DDATA(x), b, DDATA(y), DDATA(z), n);
  }
  sunamd_prof_end(SUNAMD_K_LINSUM, ex->stream);
  LAUNCHCHK(z);
}

// launch accounting for kernels issued on a vector's stream from other files
extern "C" void sunamd_nv_count_launch(N_Vector v) { EXEC(v)->launches++; }

// the reduction scratch of a vector's execution context, for reductions
// issued on its stream from other files: *partials holds *capacity doubles on
// the device, *result one pinned, device-mapped slot (the same storage the
// vector's own reductions use, so calls on one context stay ordered by its
// stream and contexts on other streams never share it)
extern "C" int sunamd_nv_reduction_scratch(N_Vector v, double** partials, long* capacity,
                                           double** result)
{
  if (!SUNAMD_NVectorIsMi355(v) || !EXEC(v)->d_partials || !EXEC(v)->h_result) { return -1; }
  *partials = EXEC(v)->d_partials;
  *capacity = (long)NV_MAX_RED_BLOCKS * (NV_MAX_MULTI + 1);
  *result   = EXEC(v)->h_result;
  return 0;
}

extern "C" int SUNAMD_NVectorIsMi355(N_Vector v)
{
  return (v && v->ops && v->ops->nvlinearsum == mi_linearsum) ? 1 : 0;
}

static void mi_const(sunrealtype c, N_Vector z)
{
  long n = (long)NVLEN(z);
  if (n == 0) { return; }
  Mi355Exec* ex = EXEC(z);
  k_const<<<stream_grid(ex, n >> 1, 1), NV_BLOCK, 0, ex->stream>>>(c, DDATA(z), n);
  LAUNCHCHK(z);
}

template <int OP>
static void mi_ew(N_Vector x, N_Vector y, double c, N_Vector z)
{
  long n = (long)NVLEN(z);
  if (n == 0) { return; }
  Mi355Exec* ex = EXEC(z);
  const int nl = (OP == EW_PROD || OP == EW_DIV) ? 2 : 1;
  k_ew<OP><<<stream_grid(ex, n >> 1, unroll_for(nl)), NV_BLOCK, 0, ex->stream>>>(
    DDATA(x), y ? DDATA(y) : DDATA(x), c, DDATA(z), n);
  LAUNCHCHK(z);
}

static void mi_prod(N_Vector x, N_Vector y, N_Vector z) { mi_ew<EW_PROD>(x, y, 0.0, z); }
static void mi_div(N_Vector x, N_Vector y, N_Vector z) { mi_ew<EW_DIV>(x, y, 0.0, z); }
static void mi_scale(sunrealtype c, N_Vector x, N_Vector z)
{
  // c == 1 (copy) and c == -1 (negate) give the same bits as c*x
  hipStream_t s = EXEC(z)->stream;
  sunamd_prof_begin(SUNAMD_K_SCALE, s);
  mi_ew<EW_SCALE>(x, NULL, c, z);
  sunamd_prof_end(SUNAMD_K_SCALE, s);
}
static void mi_abs(N_Vector x, N_Vector z) { mi_ew<EW_ABS>(x, NULL, 0.0, z); }
static void mi_inv(N_Vector x, N_Vector z) { mi_ew<EW_INV>(x, NULL, 0.0, z); }
static void mi_addconst(N_Vector x, sunrealtype b, N_Vector z)
{
  mi_ew<EW_ADDCONST>(x, NULL, b, z);
}
static void mi_compare(sunrealtype c, N_Vector x, N_Vector z)
{
  mi_ew<EW_COMPARE>(x, NULL, c, z);
}

// run one reduction kernel, wait for it, return the host-visible result
template <int K>
static double mi_reduce(N_Vector x, N_Vector y, N_Vector id, double identity)
{
  long n        = (long)NVLEN(x);
  Mi355Exec* ex = EXEC(x);
  if (n == 0) { return identity; }
  int g = red_grid(ex, n >> 1, unroll_for(red_loads<K>()));
  const int fam = (K == RK_WSQR) ? SUNAMD_K_WRMS : -1;
  // min / max are exact in any order; sums follow the vector's reduction order
  const bool seq = (RedTraits<K>::op == RSUM) && CONTENT(x)->red_order == SUNAMD_REDUCE_SEQUENTIAL;
  if (fam >= 0) { sunamd_prof_begin(fam, ex->stream); }
  if (seq)
  {
    k_reduce_seq<K><<<1, 64, 0, ex->stream>>>(DDATA(x), y ? DDATA(y) : DDATA(x),
                                             id ? DDATA(id) : DDATA(x), n, ex->h_result);
  }
  else
  {
    k_reduce<K><<<g, NV_BLOCK, 0, ex->stream>>>(DDATA(x), y ? DDATA(y) : DDATA(x),
                                                id ? DDATA(id) : DDATA(x), n, identity,
                                                ex->d_partials);
    k_finish<RedTraits<K>::op, 1><<<1, NV_BLOCK, 0, ex->stream>>>(ex->d_partials, g,
                                                                 ex->h_result, identity);
  }
  if (fam >= 0) { sunamd_prof_end(fam, ex->stream); }
  LAUNCHCHK(x);
  HIPCHK_EX(x, hipStreamSynchronize(ex->stream));
  return (ex->err == hipSuccess) ? ex->h_result[0] : NAN;
}

static sunrealtype mi_dotprod(N_Vector x, N_Vector y) { return mi_reduce<RK_DOT>(x, y, NULL, 0.0); }
static sunrealtype mi_maxnorm(N_Vector x) { return mi_reduce<RK_MAXABS>(x, NULL, NULL, 0.0); }
static sunrealtype mi_wsqrsum(N_Vector x, N_Vector w) { return mi_reduce<RK_WSQR>(x, w, NULL, 0.0); }
static sunrealtype mi_wsqrsummask(N_Vector x, N_Vector w, N_Vector id)
{
  return mi_reduce<RK_WSQRMASK>(x, w, id, 0.0);
}
static inline double sqrt0(double v) { return (v <= 0.0) ? 0.0 : sqrt(v); }
static sunrealtype mi_wrmsnorm(N_Vector x, N_Vector w)
{
  return sqrt0(mi_wsqrsum(x, w) / (double)NVLEN(x));
}
static sunrealtype mi_wrmsnormmask(N_Vector x, N_Vector w, N_Vector id)
{
  return sqrt0(mi_wsqrsummask(x, w, id) / (double)NVLEN(x));
}
static sunrealtype mi_min(N_Vector x) { return mi_reduce<RK_MIN>(x, NULL, NULL, DBL_MAX); }
static sunrealtype mi_wl2norm(N_Vector x, N_Vector w) { return sqrt0(mi_wsqrsum(x, w)); }
static sunrealtype mi_l1norm(N_Vector x) { return mi_reduce<RK_L1>(x, NULL, NULL, 0.0); }

static sunbooleantype mi_invtest(N_Vector x, N_Vector z)
{
  long n        = (long)NVLEN(x);
  Mi355Exec* ex = EXEC(x);
  if (n == 0) { return SUNTRUE; }
  const int g = red_grid(ex, n >> 1, 4);
  k_invtest<<<g, NV_BLOCK, 0, ex->stream>>>(DDATA(x), DDATA(z), n, ex->d_partials);
  k_finish<RMAX, 1><<<1, NV_BLOCK, 0, ex->stream>>>(ex->d_partials, g, ex->h_result, 0.0);
  LAUNCHCHK(x);
  HIPCHK(x->sunctx, hipStreamSynchronize(ex->stream));
  return (ex->h_result[0] > 0.0) ? SUNFALSE : SUNTRUE;
}

static sunbooleantype mi_constrmask(N_Vector c, N_Vector x, N_Vector m)
{
  long n        = (long)NVLEN(x);
  Mi355Exec* ex = EXEC(x);
  if (n == 0) { return SUNTRUE; }
  const int g = red_grid(ex, n >> 1, 4);
  k_constrmask<<<g, NV_BLOCK, 0, ex->stream>>>(DDATA(c), DDATA(x), DDATA(m), n, ex->d_partials);
  k_finish<RMAX, 1><<<1, NV_BLOCK, 0, ex->stream>>>(ex->d_partials, g, ex->h_result, 0.0);
  LAUNCHCHK(x);
  HIPCHK(x->sunctx, hipStreamSynchronize(ex->stream));
  return (ex->h_result[0] > 0.0) ? SUNFALSE : SUNTRUE;
}

static sunrealtype mi_minquotient(N_Vector num, N_Vector den)
{
  long n        = (long)NVLEN(num);
  Mi355Exec* ex = EXEC(num);
  if (n == 0) { return SUN_BIG_REAL; }
  const int g = red_grid(ex, n >> 1, 4);
  k_minquotient<<<g, NV_BLOCK, 0, ex->stream>>>(DDATA(num), DDATA(den), n, ex->d_partials);
  k_finish<RMIN, 1><<<1, NV_BLOCK, 0, ex->stream>>>(ex->d_partials, g, ex->h_result, DBL_MAX);
  LAUNCHCHK(num);
  HIPCHK(num->sunctx, hipStreamSynchronize(ex->stream));
  return ex->h_result[0];
}

// ---- fused ops --------------------------------------------------------------
#define NV_CASES(MACRO) \
  MACRO(1) MACRO(2) MACRO(3) MACRO(4) MACRO(5) MACRO(6) MACRO(7) MACRO(8)

static SUNErrCode mi_dotprodmulti(int nvec, N_Vector x, N_Vector* Y, sunrealtype* d)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  long n        = (long)NVLEN(x);
  Mi355Exec* ex = EXEC(x);
  if (n == 0)
  {
    for (int k = 0; k < nvec; k++) { d[k] = 0.0; }
    return SUN_SUCCESS;
  }
  if (CONTENT(x)->red_order == SUNAMD_REDUCE_SEQUENTIAL)
  {
    for (int base = 0; base < nvec; base += NV_MAX_MULTI)
    {
      int nv = nvec - base < NV_MAX_MULTI ? nvec - base : NV_MAX_MULTI;
      PtrPack P;
      for (int k = 0; k < nv; k++) { P.p[k] = DDATA(Y[base + k]); }
      k_dotmulti_seq<<<1, 64, 0, ex->stream>>>(DDATA(x), P, nv, n, ex->h_result);
      LAUNCHCHK(x);
      HIPCHK_EX(x, hipStreamSynchronize(ex->stream));
      for (int k = 0; k < nv; k++) { d[base + k] = ex->h_result[k]; }
    }
    return exec_status(ex);
  }
  // chunks of up to 8 vectors keep the per-lane accumulators in registers;
  // results land in the pinned slots, read back every NV_MAX_MULTI products
  int group = 0;
  for (int base = 0; base < nvec; base += 8)
  {
    int nv = nvec - base < 8 ? nvec - base : 8;
    PtrPack P;
    for (int k = 0; k < nv; k++) { P.p[k] = DDATA(Y[base + k]); }
    double* slot = ex->h_result + (base - group);
    int g        = 0;
    sunamd_prof_begin(SUNAMD_K_DOTMULTI, ex->stream);
    switch (nv)
    {
#define DM_CASE(NVV)                                                                       \
  case NVV:                                                                                \
    g = red_grid(ex, n >> 1, unroll_for(NVV + 1));                                         \
    k_dotmulti<NVV><<<g, NV_BLOCK, 0, ex->stream>>>(DDATA(x), P, n, ex->d_partials);       \
    k_finish<RSUM, NVV><<<1, NV_BLOCK, 0, ex->stream>>>(ex->d_partials, g, slot, 0.0);     \
    break;
      NV_CASES(DM_CASE)
#undef DM_CASE
    }
    sunamd_prof_end(SUNAMD_K_DOTMULTI, ex->stream);
    LAUNCHCHK(x);
    int filled = base + nv - group;
    if (filled == NV_MAX_MULTI || base + nv == nvec)
    {
      HIPCHK_EX(x, hipStreamSynchronize(ex->stream));
      for (int k = 0; k < filled; k++) { d[group + k] = ex->h_result[k]; }
      group += filled;
    }
  }
  return exec_status(ex);
}

static SUNErrCode mi_lincomb(int nvec, sunrealtype* c, N_Vector* X, N_Vector z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  long n        = (long)NVLEN(z);
  Mi355Exec* ex = EXEC(z);
  if (n == 0) { return SUN_SUCCESS; }
  // first chunk writes z, later chunks accumulate into z (same rounding
  // order as the sequential loop)
  int done = 0;
  while (done < nvec)
  {
    CoefPack C;
    PtrPack P;
    int nv;
    if (done == 0)
    {
      nv = nvec < 8 ? nvec : 8;
      for (int k = 0; k < nv; k++)
      {
        C.c[k] = c[k];
        P.p[k] = DDATA(X[k]);
      }
    }
    else
    {
      nv     = (nvec - done) + 1 < 8 ? (nvec - done) + 1 : 8;
      C.c[0] = 1.0;
      P.p[0] = DDATA(z);
      for (int k = 1; k < nv; k++)
      {
        C.c[k] = c[done + k - 1];
        P.p[k] = DDATA(X[done + k - 1]);
      }
    }
    switch (nv)
    {
#define LC_CASE(NVV)                                                                    \
  case NVV:                                                                             \
    k_lincomb<NVV><<<stream_grid(ex, n >> 1, unroll_for(NVV)), NV_BLOCK, 0, ex->stream>>>( \
      C, P, DDATA(z), n);                                                                 \
    break;
      NV_CASES(LC_CASE)
#undef LC_CASE
    }
    LAUNCHCHK(z);
    done += (done == 0) ? nv : nv - 1;
  }
  return exec_status(ex);
}

static SUNErrCode mi_scaleaddmulti(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y,
                                   N_Vector* Z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  long n        = (long)NVLEN(x);
  Mi355Exec* ex = EXEC(x);
  if (n == 0) { return SUN_SUCCESS; }
  for (int base = 0; base < nvec; base += 8)
  {
    int nv = nvec - base < 8 ? nvec - base : 8;
    CoefPack C;
    PtrPack P;
    WPtrPack W;
    for (int k = 0; k < nv; k++)
    {
      C.c[k] = a[base + k];
      P.p[k] = DDATA(Y[base + k]);
      W.p[k] = DDATA(Z[base + k]);
    }
    switch (nv)
    {
#define SAM_CASE(NVV)                                                                        \
  case NVV:                                                                                  \
    k_scaleaddmulti<NVV><<<stream_grid(ex, n >> 1, unroll_for(NVV + 1)), NV_BLOCK, 0,           \
                           ex->stream>>>(C, DDATA(x), P, W, n);                                \
    break;
      NV_CASES(SAM_CASE)
#undef SAM_CASE
    }
    LAUNCHCHK(x);
  }
  return exec_status(ex);
}

static SUNErrCode mi_scalevectorarray(int nvec, sunrealtype* c, N_Vector* X, N_Vector* Z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  long n        = (long)NVLEN(Z[0]);
  Mi355Exec* ex = EXEC(Z[0]);
  if (n == 0) { return SUN_SUCCESS; }
  for (int base = 0; base < nvec; base += 8)
  {
    int nv = nvec - base < 8 ? nvec - base : 8;
    CoefPack C;
    PtrPack P;
    WPtrPack W;
    for (int k = 0; k < nv; k++)
    {
      C.c[k] = c[base + k];
      P.p[k] = DDATA(X[base + k]);
      W.p[k] = DDATA(Z[base + k]);
    }
    switch (nv)
    {
#define SVA_CASE(NVV)                                                                  \
  case NVV:                                                                            \
    k_scalearray<NVV><<<stream_grid(ex, n >> 1, unroll_for(NVV)), NV_BLOCK, 0,            \
                        ex->stream>>>(C, P, W, n);                                       \
    break;
      NV_CASES(SVA_CASE)
#undef SVA_CASE
    }
    LAUNCHCHK(Z[0]);
  }
  return exec_status(ex);
}

static SUNErrCode mi_linearsumvectorarray(int nvec, sunrealtype a, N_Vector* X, sunrealtype b,
                                          N_Vector* Y, N_Vector* Z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  for (int i = 0; i < nvec; i++) { mi_linearsum(a, X[i], b, Y[i], Z[i]); }
  return exec_status(EXEC(Z[0]));
}

static SUNErrCode mi_constvectorarray(int nvec, sunrealtype c, N_Vector* Z)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  for (int i = 0; i < nvec; i++) { mi_const(c, Z[i]); }
  return exec_status(EXEC(Z[0]));
}

static SUNErrCode mi_wrmsnormvectorarray(int nvec, N_Vector* X, N_Vector* W, sunrealtype* nrm)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  for (int i = 0; i < nvec; i++) { nrm[i] = mi_wrmsnorm(X[i], W[i]); }
  return exec_status(EXEC(X[0]));
}

static SUNErrCode mi_wrmsnormmaskvectorarray(int nvec, N_Vector* X, N_Vector* W, N_Vector id,
                                             sunrealtype* nrm)
{
  if (nvec < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  for (int i = 0; i < nvec; i++) { nrm[i] = mi_wrmsnormmask(X[i], W[i], id); }
  return exec_status(EXEC(X[0]));
}

static SUNErrCode mi_scaleaddmultivectorarray(int nvec, int nsum, sunrealtype* a, N_Vector* X,
                                              N_Vector** Y, N_Vector** Z)
{
  if (nvec < 1 || nsum < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  N_Vector YY[NV_MAX_MULTI], ZZ[NV_MAX_MULTI];
  if (nsum > NV_MAX_MULTI) { return SUN_ERR_ARG_OUTOFRANGE; }
  for (int i = 0; i < nvec; i++)
  {
    for (int j = 0; j < nsum; j++)
    {
      YY[j] = Y[j][i];
      ZZ[j] = Z[j][i];
    }
    SUNErrCode e = mi_scaleaddmulti(nsum, a, X[i], YY, ZZ);
    if (e) { return e; }
  }
  return SUN_SUCCESS;
}

static SUNErrCode mi_lincombvectorarray(int nvec, int nsum, sunrealtype* c, N_Vector** X,
                                        N_Vector* Z)
{
  if (nvec < 1 || nsum < 1) { return SUN_ERR_ARG_OUTOFRANGE; }
  N_Vector Yv[64];
  if (nsum > 64) { return SUN_ERR_ARG_OUTOFRANGE; }
  for (int i = 0; i < nvec; i++)
  {
    for (int j = 0; j < nsum; j++) { Yv[j] = X[j][i]; }
    SUNErrCode e = mi_lincomb(nsum, c, Yv, Z[i]);
    if (e) { return e; }
  }
  return SUN_SUCCESS;
}

// ---- XBraid-style buffers (host buffers) ------------------------------------
static SUNErrCode mi_bufsize(N_Vector x, sunindextype* size)
{
  *size = NVLEN(x) * (sunindextype)sizeof(sunrealtype);
  return SUN_SUCCESS;
}
static SUNErrCode mi_bufpack(N_Vector x, void* buf)
{
  Mi355Exec* ex = EXEC(x);
  hipError_t e  = hipMemcpyAsync(buf, DDATA(x), (size_t)NVLEN(x) * sizeof(double),
                                 hipMemcpyDeviceToHost, ex->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(ex->stream); }
  return (e == hipSuccess) ? SUN_SUCCESS : SUN_ERR_EXT_FAIL;
}
static SUNErrCode mi_bufunpack(N_Vector x, void* buf)
{
  Mi355Exec* ex = EXEC(x);
  hipError_t e  = hipMemcpyAsync(DDATA(x), buf, (size_t)NVLEN(x) * sizeof(double),
                                 hipMemcpyHostToDevice, ex->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(ex->stream); }
  return (e == hipSuccess) ? SUN_SUCCESS : SUN_ERR_EXT_FAIL;
}

static void mi_printfile(N_Vector x, FILE* f)
{
  sunrealtype* h = mi_host_ptr(x);
  N_VCopyFromDevice_Mi355(x);
  for (sunindextype i = 0; i < NVLEN(x); i++) { fprintf(f, "%.16e\n", h[i]); }
  fprintf(f, "\n");
}
static void mi_print(N_Vector x) { mi_printfile(x, stdout); }

static void set_ops(N_Vector v)
{
  N_Vector_Ops o               = v->ops;
  o->nvgetvectorid             = mi_getid;
  o->nvclone                   = mi_clone;
  o->nvcloneempty              = mi_clone_empty;
  o->nvdestroy                 = mi_destroy;
  o->nvspace                   = mi_space;
  o->nvgetarraypointer         = mi_host_ptr;
  o->nvgetdevicearraypointer   = mi_dev_ptr;
  o->nvsetarraypointer         = mi_set_ptr;
  o->nvgetlength               = mi_getlength;
  o->nvgetlocallength          = mi_getlength;
  o->nvlinearsum               = mi_linearsum;
  o->nvconst                   = mi_const;
  o->nvprod                    = mi_prod;
  o->nvdiv                     = mi_div;
  o->nvscale                   = mi_scale;
  o->nvabs                     = mi_abs;
  o->nvinv                     = mi_inv;
  o->nvaddconst                = mi_addconst;
  o->nvdotprod                 = mi_dotprod;
  o->nvmaxnorm                 = mi_maxnorm;
  o->nvwrmsnormmask            = mi_wrmsnormmask;
  o->nvwrmsnorm                = mi_wrmsnorm;
  o->nvmin                     = mi_min;
  o->nvwl2norm                 = mi_wl2norm;
  o->nvl1norm                  = mi_l1norm;
  o->nvcompare                 = mi_compare;
  o->nvinvtest                 = mi_invtest;
  o->nvconstrmask              = mi_constrmask;
  o->nvminquotient             = mi_minquotient;
  o->nvdotprodlocal            = mi_dotprod;
  o->nvmaxnormlocal            = mi_maxnorm;
  o->nvminlocal                = mi_min;
  o->nvl1normlocal             = mi_l1norm;
  o->nvinvtestlocal            = mi_invtest;
  o->nvconstrmasklocal         = mi_constrmask;
  o->nvminquotientlocal        = mi_minquotient;
  o->nvwsqrsumlocal            = mi_wsqrsum;
  o->nvwsqrsummasklocal        = mi_wsqrsummask;
  o->nvdotprodmultilocal       = mi_dotprodmulti;
  o->nvbufsize                 = mi_bufsize;
  o->nvbufpack                 = mi_bufpack;
  o->nvbufunpack               = mi_bufunpack;
  o->nvprint                   = mi_print;
  o->nvprintfile               = mi_printfile;
}

static void set_fused(N_Vector v, sunbooleantype tf)
{
  N_Vector_Ops o                    = v->ops;
  o->nvlinearcombination            = tf ? mi_lincomb : NULL;
  o->nvscaleaddmulti                = tf ? mi_scaleaddmulti : NULL;
  o->nvdotprodmulti                 = tf ? mi_dotprodmulti : NULL;
  o->nvlinearsumvectorarray         = tf ? mi_linearsumvectorarray : NULL;
  o->nvscalevectorarray             = tf ? mi_scalevectorarray : NULL;
  o->nvconstvectorarray             = tf ? mi_constvectorarray : NULL;
  o->nvwrmsnormvectorarray          = tf ? mi_wrmsnormvectorarray : NULL;
  o->nvwrmsnormmaskvectorarray      = tf ? mi_wrmsnormmaskvectorarray : NULL;
  o->nvscaleaddmultivectorarray     = tf ? mi_scaleaddmultivectorarray : NULL;
  o->nvlinearcombinationvectorarray = tf ? mi_lincombvectorarray : NULL;
}

static N_Vector new_empty(sunindextype length, SUNContext sunctx, Mi355Exec* ex,
                          bool fresh = false)
{
  N_Vector v = N_VNewEmpty(sunctx);
  if (!v) { return NULL; }
  set_ops(v);
  set_fused(v, SUNTRUE);
  Mi355Content* c = (Mi355Content*)calloc(1, sizeof(Mi355Content));
  if (!c)
  {
    N_VFreeEmpty(v);
    return NULL;
  }
  c->length = length;
  if (ex)
  {
    if (!fresh) { ex->refcount++; }
    c->exec = ex;
  }
  else
  {
    c->exec = exec_default(sunctx);
    if (!c->exec)
    {
      free(c);
      N_VFreeEmpty(v);
      return NULL;
    }
  }
  v->content = c;
  return v;
}

static int alloc_dev(N_Vector v)
{
  Mi355Content* c = CONTENT(v);
  if (c->length <= 0) { return 1; }
  // round up to a whole 128-bit pair so vector loads never straddle the end
  size_t bytes = (size_t)((c->length + 1) & ~(sunindextype)1) * sizeof(double);
  hipError_t e = hipMalloc((void**)&c->dev, bytes);
  if (e != hipSuccess)
  {
    hip_fail(v->sunctx, e, "hipMalloc(vector)", __LINE__);
    c->dev = NULL;
    return 0;
  }
  c->own_dev = SUNTRUE;
  return 1;
}

extern "C" N_Vector N_VNewEmpty_Mi355(sunindextype length, SUNContext sunctx)
{
  if (length < 0 || !sunctx) { return NULL; }
  return new_empty(length, sunctx, NULL);
}

extern "C" N_Vector N_VNew_Mi355(sunindextype length, SUNContext sunctx)
{
  N_Vector v = N_VNewEmpty_Mi355(length, sunctx);
  if (!v) { return NULL; }
  if (!alloc_dev(v))
  {
    mi_destroy(v);
    return NULL;
  }
  return v;
}

extern "C" N_Vector N_VNewWithStream_Mi355(sunindextype length, void* stream,
                                           SUNContext sunctx)
{
  if (length < 0 || !sunctx) { return NULL; }
  Mi355Exec* ex = exec_create(sunctx, (hipStream_t)stream, false);
  if (!ex) { return NULL; }
  N_Vector v = new_empty(length, sunctx, ex, true);
  if (!v)
  {
    exec_release(ex);
    return NULL;
  }
  if (!alloc_dev(v))
  {
    mi_destroy(v);
    return NULL;
  }
  return v;
}

extern "C" N_Vector N_VMake_Mi355(sunindextype length, sunrealtype* h_vdata,
                                  sunrealtype* d_vdata, SUNContext sunctx)
{
  if (!d_vdata || !aligned16(d_vdata)) { return NULL; }
  N_Vector v = N_VNewEmpty_Mi355(length, sunctx);
  if (!v) { return NULL; }
  CONTENT(v)->dev  = d_vdata;
  CONTENT(v)->host = h_vdata;
  return v;
}

static N_Vector mi_clone_empty(N_Vector w)
{
  N_Vector v = new_empty(NVLEN(w), w->sunctx, EXEC(w));
  if (!v) { return NULL; }
  N_VCopyOps(w, v);
  CONTENT(v)->red_order = CONTENT(w)->red_order;
  return v;
}

static N_Vector mi_clone(N_Vector w)
{
  N_Vector v = mi_clone_empty(w);
  if (!v) { return NULL; }
  if (!alloc_dev(v))
  {
    mi_destroy(v);
    return NULL;
  }
  return v;
}

static void mi_destroy(N_Vector v)
{
  if (!v) { return; }
  Mi355Content* c = CONTENT(v);
  if (c)
  {
    if (c->own_dev && c->dev)
    {
      hipStreamSynchronize(c->exec->stream);
      hipFree(c->dev);
    }
    if (c->own_host && c->host) { free(c->host); }
    exec_release(c->exec);
    free(c);
    v->content = NULL;
  }
  N_VFreeEmpty(v);
}

extern "C" SUNErrCode N_VCopyToDevice_Mi355(N_Vector v)
{
  Mi355Content* c = CONTENT(v);
  if (c->length == 0) { return SUN_SUCCESS; }
  if (!c->host) { return SUN_ERR_ARG_CORRUPT; }
  hipError_t e = hipMemcpyAsync(c->dev, c->host, (size_t)c->length * sizeof(double),
                                hipMemcpyHostToDevice, c->exec->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(c->exec->stream); }
  if (e != hipSuccess)
  {
    hip_fail(v->sunctx, e, "N_VCopyToDevice", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  return SUN_SUCCESS;
}

extern "C" SUNErrCode N_VCopyFromDevice_Mi355(N_Vector v)
{
  Mi355Content* c = CONTENT(v);
  if (c->length == 0) { return SUN_SUCCESS; }
  sunrealtype* h = mi_host_ptr(v);
  if (!h) { return SUN_ERR_MALLOC_FAIL; }
  hipError_t e = hipMemcpyAsync(h, c->dev, (size_t)c->length * sizeof(double),
                                hipMemcpyDeviceToHost, c->exec->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(c->exec->stream); }
  if (e != hipSuccess)
  {
    hip_fail(v->sunctx, e, "N_VCopyFromDevice", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  return SUN_SUCCESS;
}

extern "C" SUNErrCode N_VCopyFromHostArray_Mi355(N_Vector v, const sunrealtype* src)
{
  Mi355Content* c = CONTENT(v);
  if (c->length == 0) { return SUN_SUCCESS; }
  hipError_t e = hipMemcpyAsync(c->dev, src, (size_t)c->length * sizeof(double),
                                hipMemcpyHostToDevice, c->exec->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(c->exec->stream); }
  if (e != hipSuccess)
  {
    hip_fail(v->sunctx, e, "N_VCopyFromHostArray", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  return SUN_SUCCESS;
}

extern "C" SUNErrCode N_VCopyToHostArray_Mi355(N_Vector v, sunrealtype* dst)
{
  Mi355Content* c = CONTENT(v);
  if (c->length == 0) { return SUN_SUCCESS; }
  hipError_t e = hipMemcpyAsync(dst, c->dev, (size_t)c->length * sizeof(double),
                                hipMemcpyDeviceToHost, c->exec->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(c->exec->stream); }
  if (e != hipSuccess)
  {
    hip_fail(v->sunctx, e, "N_VCopyToHostArray", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  return SUN_SUCCESS;
}

extern "C" void* N_VGetStream_Mi355(N_Vector v) { return (void*)EXEC(v)->stream; }

extern "C" SUNErrCode N_VSynchronize_Mi355(N_Vector v)
{
  hipError_t e = hipStreamSynchronize(EXEC(v)->stream);
  if (e != hipSuccess)
  {
    hip_fail(v->sunctx, e, "hipStreamSynchronize", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  return SUN_SUCCESS;
}

extern "C" SUNErrCode N_VEnableFusedOps_Mi355(N_Vector v, sunbooleantype tf)
{
  if (!v || !v->ops) { return SUN_ERR_ARG_CORRUPT; }
  set_fused(v, tf);
  return SUN_SUCCESS;
}

extern "C" long int N_VGetNumLaunches_Mi355(N_Vector v) { return EXEC(v)->launches; }

// ----------------------------------------------------------------------------
// Device classical Gram-Schmidt step (SUNClassicalGS, sundials_iterative.c:116-170)
// ----------------------------------------------------------------------------
#define GS_MAX 128

// v_k <- v_k - sum_j d_j V_j for this chunk's vectors, in order, coefficients
// read from device memory; the last chunk also sums ||v_k||^2 (tree order)
template <int NV>
__global__ void __launch_bounds__(NV_BLOCK) k_gs_update(double* vk, PtrPack V,
                                                        const double* __restrict__ d, long n,
                                                        int last, double* partials)
{
  double nrm[1] = {0.0};
  double c[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) { c[j] = -d[j]; }
  const long np = n >> 1;
  chunk_pairs<unroll_for(NV + 1)>(
    np,
    [&](long i) {
      Pairs<NV + 1> p;
      p.v[0] = ld2(vk, i);
#pragma unroll
      for (int j = 0; j < NV; j++) { p.v[j + 1] = ld2(V.p[j], i); }
      return p;
    },
    [&](long i, const Pairs<NV + 1>& p) {
      dbl2 z = p.v[0];
#pragma unroll
      for (int j = 0; j < NV; j++) { z = z + c[j] * p.v[j + 1]; }
      st2(vk, i, z);
      nrm[0] += z.x * z.x;
      nrm[0] += z.y * z.y;
    });
  TAIL(n)
  {
    double z = vk[n - 1];
#pragma unroll
    for (int j = 0; j < NV; j++) { z = z + c[j] * V.p[j][n - 1]; }
    vk[n - 1] = z;
    nrm[0] += z * z;
  }
  if (last) { grid_finish<RSUM, 1>(nrm, partials); }
}

static bool gs_buffers(Mi355Exec* ex)
{
  if (ex->d_gs && ex->h_gs) { return true; }
  if (!ex->d_gs && hipMalloc((void**)&ex->d_gs, sizeof(double) * (GS_MAX + 2)) != hipSuccess)
  {
    return false;
  }
  if (!ex->h_gs && hipHostMalloc((void**)&ex->h_gs, sizeof(double) * (GS_MAX + 2),
                                 hipHostMallocDefault) != hipSuccess)
  {
    return false;
  }
  return true;
}

// Same arguments, coefficients and reorthogonalisation rule as SUNClassicalGS;
// the projections are computed into device memory, the update reads them
// there and sums the new norm in the same pass, so one synchronisation
// returns the Hessenberg column and both norms.  Vectors in sequential
// reduction order, or more than GS_MAX basis vectors, take SUNClassicalGS.
extern "C" SUNErrCode SUNAMD_ClassicalGS_Mi355(N_Vector* v, sunrealtype** h, int k, int p,
                                               sunrealtype* new_vk_norm, sunrealtype* stemp,
                                               N_Vector* vtemp)
{
  const int k_minus_1 = k - 1;
  const int i0        = (k - p > 0) ? k - p : 0;
  const int nv        = k - i0; // vectors projected out
  N_Vector vk         = v[k];
  Mi355Exec* ex       = EXEC(vk);
  const long n        = (long)NVLEN(vk);
  if (CONTENT(vk)->red_order == SUNAMD_REDUCE_SEQUENTIAL || nv < 1 || nv + 1 > GS_MAX ||
      n == 0 || !gs_buffers(ex))
  {
    return SUNClassicalGS(v, h, k, p, new_vk_norm, stemp, vtemp);
  }
  // projections d_j = v_k . v_(i0+j), j < nv, and v_k . v_k into d_gs[0..nv]
  for (int base = 0; base <= nv; base += 8)
  {
    int cnt = nv + 1 - base < 8 ? nv + 1 - base : 8;
    PtrPack P;
    for (int j = 0; j < cnt; j++) { P.p[j] = DDATA(v[i0 + base + j]); }
    int g = 0;
    switch (cnt)
    {
#define GD_CASE(NVV)                                                                        \
  case NVV:                                                                                 \
    g = red_grid(ex, n >> 1, unroll_for(NVV + 1));                                          \
    k_dotmulti<NVV><<<g, NV_BLOCK, 0, ex->stream>>>(DDATA(vk), P, n, ex->d_partials);       \
    k_finish<RSUM, NVV><<<1, NV_BLOCK, 0, ex->stream>>>(ex->d_partials, g, ex->d_gs + base,  \
                                                        0.0);                               \
    break;
      NV_CASES(GD_CASE)
#undef GD_CASE
    }
    LAUNCHCHK(vk);
  }
  // v_k -= sum_j d_j v_j (the reference's linear combination pairs the
  // coefficient of v_(i0+j) with v[j]; identical when i0 == 0), new norm^2
  // into d_gs[nv + 1]
  for (int base = 0; base < nv; base += 8)
  {
    int cnt  = nv - base < 8 ? nv - base : 8;
    int last = (base + cnt == nv) ? 1 : 0;
    PtrPack P;
    for (int j = 0; j < cnt; j++) { P.p[j] = DDATA(v[base + j]); }
    int g = 0;
    switch (cnt)
    {
#define GU_CASE(NVV)                                                                         \
  case NVV:                                                                                  \
    g = red_grid(ex, n >> 1, unroll_for(NVV + 1));                                           \
    k_gs_update<NVV><<<g, NV_BLOCK, 0, ex->stream>>>(DDATA(vk), P, ex->d_gs + base, n, last, \
                                                     ex->d_partials);                        \
    if (last)                                                                                \
    {                                                                                        \
      k_finish<RSUM, 1><<<1, NV_BLOCK, 0, ex->stream>>>(ex->d_partials, g, ex->d_gs + nv + 1, \
                                                        0.0);                                \
    }                                                                                        \
    break;
      NV_CASES(GU_CASE)
#undef GU_CASE
    }
    LAUNCHCHK(vk);
  }
  HIPCHK_EX(vk, hipMemcpyAsync(ex->h_gs, ex->d_gs, sizeof(double) * (nv + 2),
                               hipMemcpyDeviceToHost, ex->stream));
  HIPCHK_EX(vk, hipStreamSynchronize(ex->stream));
  if (ex->err != hipSuccess) { return SUN_ERR_EXT_FAIL; }
  const double* d      = ex->h_gs;
  const double vk_norm = (d[nv] <= 0.0) ? 0.0 : sqrt(d[nv]);
  for (int i = 0; i < nv; i++) { h[i][k_minus_1] = d[i]; }
  *new_vk_norm = (d[nv + 1] <= 0.0) ? 0.0 : sqrt(d[nv + 1]);

  // reorthogonalise with the reference's second pass when the projection
  // removed almost everything
  if ((1000.0 * (*new_vk_norm)) < vk_norm)
  {
    SUNErrCode err = mi_dotprodmulti(nv, vk, v + i0, stemp + 1);
    if (err != SUN_SUCCESS) { return err; }
    stemp[0] = 1.0;
    vtemp[0] = vk;
    for (int i = i0; i < k; i++)
    {
      h[i][k_minus_1] += stemp[i - i0 + 1];
      stemp[i - i0 + 1] = -stemp[i - i0 + 1];
      vtemp[i - i0 + 1] = v[i - i0];
    }
    err = mi_lincomb(k + 1, stemp, vtemp, vk);
    if (err != SUN_SUCCESS) { return err; }
    double nn    = mi_dotprod(vk, vk);
    *new_vk_norm = (nn <= 0.0) ? 0.0 : sqrt(nn);
  }
  return SUN_SUCCESS;
}

extern "C" int N_VGetHipError_Mi355(N_Vector v)
{
  if (!SUNAMD_NVectorIsMi355(v)) { return -1; }
  Mi355Exec* ex = EXEC(v);
  const int e   = (int)ex->err;
  ex->err       = hipSuccess;
  return e;
}

extern "C" SUNErrCode N_VSetReductionOrder_Mi355(N_Vector v, int order)
{
  if (!v || !v->content) { return SUN_ERR_ARG_CORRUPT; }
  if (order != SUNAMD_REDUCE_TREE && order != SUNAMD_REDUCE_SEQUENTIAL)
  {
    return SUN_ERR_ARG_OUTOFRANGE;
  }
  CONTENT(v)->red_order = order;
  return SUN_SUCCESS;
}
