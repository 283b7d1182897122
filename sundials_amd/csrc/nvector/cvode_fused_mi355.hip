// -----------------------------------------------------------------------------
// CVODE's fused error-weight and Newton-residual kernels on the MI355X vector:
// the counterparts of cvEwtSetSS_fused / cvEwtSetSV_fused / cvNlsResid_fused
// (src/cvode/cvode_fused_gpu.cpp:62-255), with the same arguments.  Each
// element is formed with the separate roundings of the unfused op sequence
// it replaces (Abs, Scale, AddConst, Inv -- or two LinearSums with the serial
// vector's three rounding forms), so switching the fused kernels on does not
// change a single bit of the integration; it saves launches and passes:
// error weights 64 -> 24 bytes per element, residual 48 -> 32.
// The same switch also fuses the Nordsieck predictor and its undo (cvPredict /
// cvRestore, cvode.c:2761-2768, 2791-2803; not fused in the reference): the
// q(q+1)/2 LinearSums zn[j-1] +- zn[j] run per element in registers, in the
// same order, so each element rounds as in the op-by-op sequence; at q = 5,
// 360 -> 88 bytes per element and 15 launches -> 1.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <cstdio>

#include "sunamd/nvector_mi355.h"
#include "sunamd/sundials_core.h"

extern "C" void sunamd_nv_count_launch(N_Vector v);

namespace
{

__device__ __forceinline__ double lsum3(double a, double x, double b, double y)
{
  if (a == b) { return a * (x + y); }
  if (a == -b) { return a * (x - y); }
  return (a * x) + (b * y);
}

__global__ void k_ewt_ss(double rtol, double atol, const double* __restrict__ y,
                         double* tempv, double* w, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
  {
    double t = fabs(y[i]);
    t        = rtol * t;
    t        = t + atol;
    tempv[i] = t;
    w[i]     = 1.0 / t;
  }
}

__global__ void k_ewt_sv(double rtol, const double* __restrict__ atol,
                         const double* __restrict__ y, double* tempv, double* w, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
  {
    double t = lsum3(rtol, fabs(y[i]), 1.0, atol[i]);
    tempv[i] = t;
    w[i]     = 1.0 / t;
  }
}

__global__ void k_nls_resid(double rl1, double ngamma, const double* __restrict__ zn1,
                            const double* __restrict__ ycor, const double* __restrict__ ftemp,
                            double* res, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
  {
    double r = lsum3(rl1, zn1[i], 1.0, ycor[i]);
    res[i]   = lsum3(ngamma, ftemp[i], 1.0, r);
  }
}

// Pascal-triangle update of the Nordsieck array, element pairs [0, n/2) with
// 128-bit accesses, the odd tail element by thread 0: zn[j-1] = zn[j-1] + s
// zn[j] for k = 1..Q, j = Q..k (s = +1 predictor, -1 restore)
struct ZnPtrs
{
  double* p[6];
};

typedef double dbl2_t __attribute__((ext_vector_type(2)));

template <int Q, int SIGN>
__device__ __forceinline__ void pascal_pair(dbl2_t (&z)[Q + 1])
{
#pragma unroll
  for (int k = 1; k <= Q; k++)
  {
#pragma unroll
    for (int j = Q; j >= k; j--) { z[j - 1] = (SIGN > 0) ? z[j - 1] + z[j] : z[j - 1] - z[j]; }
  }
}

// Element pairs in chunks of PASCAL_U * 256 consecutive pairs per workgroup,
// chunks strided over the grid; in a full chunk a lane requests all of its
// pairs' history entries before the first sum (PASCAL_U * (Q + 1) 16-byte
// loads in flight), the last, partial chunk goes pair by pair.
#define PASCAL_U 4
template <int Q, int SIGN>
__global__ void __launch_bounds__(256) k_pascal(ZnPtrs Z, long n)
{
  const long np    = n >> 1;
  const long chunk = (long)PASCAL_U * 256;
  for (long c = (long)blockIdx.x * chunk; c < np; c += (long)gridDim.x * chunk)
  {
    const long i0 = c + threadIdx.x;
    if (c + chunk <= np)
    {
      dbl2_t z[PASCAL_U][Q + 1];
#pragma unroll
      for (int u = 0; u < PASCAL_U; u++)
      {
#pragma unroll
        for (int j = 0; j <= Q; j++)
        {
          z[u][j] = __builtin_nontemporal_load(reinterpret_cast<const dbl2_t*>(Z.p[j]) + i0 + u * 256);
        }
      }
#pragma unroll
      for (int u = 0; u < PASCAL_U; u++)
      {
        pascal_pair<Q, SIGN>(z[u]);
#pragma unroll
        for (int j = 0; j < Q; j++)
        {
          __builtin_nontemporal_store(z[u][j], reinterpret_cast<dbl2_t*>(Z.p[j]) + i0 + u * 256);
        }
      }
    }
    else
    {
      for (long i = i0; i < np; i += 256)
      {
        dbl2_t z[Q + 1];
#pragma unroll
        for (int j = 0; j <= Q; j++)
        {
          z[j] = __builtin_nontemporal_load(reinterpret_cast<const dbl2_t*>(Z.p[j]) + i);
        }
        pascal_pair<Q, SIGN>(z);
#pragma unroll
        for (int j = 0; j < Q; j++)
        {
          __builtin_nontemporal_store(z[j], reinterpret_cast<dbl2_t*>(Z.p[j]) + i);
        }
      }
    }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0)
  {
    double z[Q + 1];
#pragma unroll
    for (int j = 0; j <= Q; j++) { z[j] = Z.p[j][n - 1]; }
#pragma unroll
    for (int k = 1; k <= Q; k++)
    {
#pragma unroll
      for (int j = Q; j >= k; j--) { z[j - 1] = (SIGN > 0) ? z[j - 1] + z[j] : z[j - 1] - z[j]; }
    }
#pragma unroll
    for (int j = 0; j < Q; j++) { Z.p[j][n - 1] = z[j]; }
  }
}

// chunked launch: one workgroup per chunk, at most 32 workgroups per CU
unsigned pascal_grid(long npairs)
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
  const long chunk = (long)PASCAL_U * 256;
  long g           = (npairs + chunk - 1) / chunk;
  const long cap   = 32L * cus;
  if (g > cap) { g = cap; }
  return (unsigned)(g > 0 ? g : 1);
}

unsigned grid_for(long n)
{
  long g = (n + 255) / 256;
  return (unsigned)(g < 4096 ? (g > 0 ? g : 1) : 4096);
}

int done(const char* what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
  {
    fprintf(stderr, "[sundials_amd cvode_fused] %s: %s\n", what, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

} // namespace

extern "C" int cvEwtSetSS_fused_Mi355(sunbooleantype atolMin0, sunrealtype reltol,
                                      sunrealtype Sabstol, N_Vector ycur, N_Vector tempv,
                                      N_Vector weight)
{
  (void)atolMin0; // the caller applies the N_VMin test to tempv
  if (!SUNAMD_NVectorIsMi355(weight)) { return -1; }
  const long n = (long)N_VGetLength(weight);
  if (n == 0) { return 0; }
  hipStream_t s = (hipStream_t)N_VGetStream_Mi355(weight);
  k_ewt_ss<<<grid_for(n), 256, 0, s>>>(reltol, Sabstol, N_VGetDeviceArrayPointer(ycur),
                                       N_VGetDeviceArrayPointer(tempv),
                                       N_VGetDeviceArrayPointer(weight), n);
  sunamd_nv_count_launch(weight);
  return done("ewt SS");
}

extern "C" int cvEwtSetSV_fused_Mi355(sunbooleantype atolMin0, sunrealtype reltol,
                                      N_Vector Vabstol, N_Vector ycur, N_Vector tempv,
                                      N_Vector weight)
{
  (void)atolMin0;
  if (!SUNAMD_NVectorIsMi355(weight)) { return -1; }
  const long n = (long)N_VGetLength(weight);
  if (n == 0) { return 0; }
  hipStream_t s = (hipStream_t)N_VGetStream_Mi355(weight);
  k_ewt_sv<<<grid_for(n), 256, 0, s>>>(reltol, N_VGetDeviceArrayPointer(Vabstol),
                                       N_VGetDeviceArrayPointer(ycur),
                                       N_VGetDeviceArrayPointer(tempv),
                                       N_VGetDeviceArrayPointer(weight), n);
  sunamd_nv_count_launch(weight);
  return done("ewt SV");
}

extern "C" int cvNlsResid_fused_Mi355(sunrealtype rl1, sunrealtype ngamma, N_Vector zn1,
                                      N_Vector ycor, N_Vector ftemp, N_Vector res)
{
  if (!SUNAMD_NVectorIsMi355(res)) { return -1; }
  const long n = (long)N_VGetLength(res);
  if (n == 0) { return 0; }
  hipStream_t s = (hipStream_t)N_VGetStream_Mi355(res);
  k_nls_resid<<<grid_for(n), 256, 0, s>>>(rl1, ngamma, N_VGetDeviceArrayPointer(zn1),
                                          N_VGetDeviceArrayPointer(ycor),
                                          N_VGetDeviceArrayPointer(ftemp),
                                          N_VGetDeviceArrayPointer(res), n);
  sunamd_nv_count_launch(res);
  return done("residual");
}

// zn[0..q] on one MI355X stream; q in 1..5
static int pascal(int q, N_Vector* zn, int sign)
{
  if (q < 1 || q > 5 || !SUNAMD_NVectorIsMi355(zn[0])) { return -1; }
  const long n = (long)N_VGetLength(zn[0]);
  if (n == 0) { return 0; }
  ZnPtrs Z = {};
  for (int j = 0; j <= q; j++)
  {
    if (!SUNAMD_NVectorIsMi355(zn[j])) { return -1; }
    Z.p[j] = N_VGetDeviceArrayPointer(zn[j]);
  }
  hipStream_t s     = (hipStream_t)N_VGetStream_Mi355(zn[0]);
  const unsigned g  = pascal_grid(n >> 1);
#define PASCAL_CASE(QQ)                                                              \
  case QQ:                                                                           \
    if (sign > 0) { k_pascal<QQ, 1><<<g, 256, 0, s>>>(Z, n); }                       \
    else { k_pascal<QQ, -1><<<g, 256, 0, s>>>(Z, n); }                               \
    break;
  switch (q)
  {
    PASCAL_CASE(1) PASCAL_CASE(2) PASCAL_CASE(3) PASCAL_CASE(4) PASCAL_CASE(5)
  }
#undef PASCAL_CASE
  sunamd_nv_count_launch(zn[0]);
  return done(sign > 0 ? "predict" : "restore");
}

extern "C" int cvPredict_fused_Mi355(int q, N_Vector* zn) { return pascal(q, zn, 1); }

extern "C" int cvRestore_fused_Mi355(int q, N_Vector* zn) { return pascal(q, zn, -1); }
