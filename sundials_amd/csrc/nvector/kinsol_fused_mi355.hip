// -----------------------------------------------------------------------------
// Fused KINSOL scaled norms for the MI355X vector.
//
// KINSOL evaluates its two scaled max norms as chains of vector ops
// (kinsol.c:2038-2062):
//   KINScFNorm(v, s) = MaxNorm(Prod(s, v))                      2 passes
//   KINScSNorm(v, u) = MaxNorm(Div(v, Inv(uscale) + Abs(u)))    5 passes
// Each pass reads and writes whole vectors.  The kernels here compute the
// same per-element values with the same IEEE operations in the same order and
// take their maximum in one pass; a maximum does not depend on the order it is
// taken in, so the results are bit-identical to the op chains (NaN entries are
// skipped, as the serial MaxNorm's '>' comparison skips them).  The op chains
// also leave their intermediates in KINSOL's work vectors; no caller reads
// them afterwards.
//
// Reduction: each workgroup folds a contiguous range into one partial; one
// workgroup folds the partials and writes the result into the pinned result
// slot of the vector's execution context, on the vector's stream, which the
// caller synchronises.  The partials are the context's reduction scratch too,
// so solves on different streams never share scratch (the same rule as every
// other reduction of the vector).
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "sunamd/nvector_mi355.h"
#include "sunamd/sundials_core.h"

namespace
{

constexpr int TPB      = 256;
constexpr int MAX_GRID = 2048;

} // namespace

// reduction scratch of a vector's execution context (nvector_mi355.hip)
extern "C" int sunamd_nv_reduction_scratch(N_Vector v, double** partials, long* capacity,
                                           double** result);

namespace
{

__device__ __forceinline__ double wave_max(double m)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
  {
    const double o = __shfl_xor(m, off, 64);
    m              = (o > m) ? o : m;
  }
  return m;
}

// block-wide maximum of m (every thread returns it)
__device__ __forceinline__ double block_max(double m)
{
  __shared__ double wmax[TPB / 64];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) { wmax[threadIdx.x >> 6] = m; }
  __syncthreads();
  m = wmax[0];
#pragma unroll
  for (int w = 1; w < TPB / 64; w++) { m = (wmax[w] > m) ? wmax[w] : m; }
  return m;
}

// KIND 0: |v_i / (1/s_i + |u_i|)|  (KINScSNorm, s = uscale)
// KIND 1: |s_i * v_i|              (KINScFNorm, s = fscale)
template <int KIND>
__global__ void __launch_bounds__(TPB) k_kin_scnorm(const double* __restrict__ v,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ s, long n,
                                                    long chunk, double* __restrict__ part)
{
  const long lo = (long)blockIdx.x * chunk;
  const long hi = (lo + chunk < n) ? lo + chunk : n;
  double m      = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += TPB)
  {
    double t;
    if (KIND == 0)
    {
      const double d = (1.0 / s[i]) + fabs(u[i]);
      t              = fabs(v[i] / d);
    }
    else { t = fabs(s[i] * v[i]); }
    m = (t > m) ? t : m;
  }
  m = block_max(m);
  if (threadIdx.x == 0) { part[blockIdx.x] = m; }
}

__global__ void __launch_bounds__(TPB) k_kin_scnorm_finish(const double* __restrict__ part,
                                                           int g, double* out)
{
  double m = 0.0;
  for (int i = threadIdx.x; i < g; i += TPB) { m = (part[i] > m) ? part[i] : m; }
  m = block_max(m);
  if (threadIdx.x == 0) { *out = m; }
}

template <int KIND>
double scnorm(N_Vector v, N_Vector u, N_Vector s)
{
  if (!SUNAMD_NVectorIsMi355(v) || !SUNAMD_NVectorIsMi355(s) ||
      (KIND == 0 && !SUNAMD_NVectorIsMi355(u)))
  {
    return -1.0;
  }
  const long n = (long)N_VGetLength(v);
  if (n == 0) { return 0.0; }
  double *partials = nullptr, *result = nullptr;
  long cap         = 0;
  if (sunamd_nv_reduction_scratch(v, &partials, &cap, &result) != 0) { return -1.0; }
  long g = (n + 4L * TPB - 1) / (4L * TPB); // at least 4 elements per thread
  if (g > MAX_GRID) { g = MAX_GRID; }
  if (g > cap) { g = cap; }
  const long chunk = (n + g - 1) / g;
  g                = (n + chunk - 1) / chunk;
  hipStream_t st   = (hipStream_t)N_VGetStream_Mi355(v);
  k_kin_scnorm<KIND><<<(unsigned)g, TPB, 0, st>>>(
    N_VGetDeviceArrayPointer(v), KIND == 0 ? N_VGetDeviceArrayPointer(u) : nullptr,
    N_VGetDeviceArrayPointer(s), n, chunk, partials);
  k_kin_scnorm_finish<<<1, TPB, 0, st>>>(partials, (int)g, result);
  if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess)
  {
    std::fprintf(stderr, "[sunamd] fused KINSOL norm: HIP error\n");
    return -1.0;
  }
  return *(volatile double*)result;
}

} // namespace

// KINScSNorm(v, u) with uscale: max_i |v_i / (1/uscale_i + |u_i|)|.
// Returns a negative value when a vector is not an MI355X vector or on a HIP
// error (the caller then runs the op chain).
extern "C" sunrealtype kinScSNorm_fused_Mi355(N_Vector v, N_Vector u, N_Vector uscale)
{
  return scnorm<0>(v, u, uscale);
}

// KINScFNorm(v, fscale): max_i |fscale_i * v_i|; negative as above.
extern "C" sunrealtype kinScFNorm_fused_Mi355(N_Vector v, N_Vector fscale)
{
  return scnorm<1>(v, nullptr, fscale);
}
