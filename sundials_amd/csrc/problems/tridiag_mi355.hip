// -----------------------------------------------------------------------------
// Device callbacks of the SPGMR test problem (tridiag(-1, 5, -1) with row
// and column scaling, Jacobi preconditioner), following
// test/unit_tests/sunlinsol/spgmr/serial/test_sunlinsol_spgmr_serial.c:433-503
// element by element (same products, same order of additions, v = 0 beyond
// the ends).  Launched on the output vector's stream.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <cstdio>

#include "sunamd/nvector_mi355.h"
#include "sunamd/sundials_core.h"
#include "sunamd/tridiag_mi355.h"

namespace
{

__global__ void k_tridiag(long n, const double* __restrict__ v, const double* __restrict__ s1,
                          const double* __restrict__ s2, double* __restrict__ z)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) { return; }
  double r;
  if (i == 0) { r = (5.0 * v[0] * s2[0] - v[1] * s2[1]) / s1[0]; }
  else if (i == n - 1)
  {
    r = (-v[n - 2] * s2[n - 2] + 5.0 * v[n - 1] * s2[n - 1]) / s1[n - 1];
  }
  else { r = (-v[i - 1] * s2[i - 1] + 5.0 * v[i] * s2[i] - v[i + 1] * s2[i + 1]) / s1[i]; }
  z[i] = r;
}

__global__ void k_jacobi(long n, const double* __restrict__ r, const double* __restrict__ d,
                         double* __restrict__ z)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { z[i] = r[i] / d[i]; }
}

int launch_check(const char* what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
  {
    fprintf(stderr, "[sundials_amd tridiag] %s: %s\n", what, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

} // namespace

extern "C" int SUNAMD_TridiagATimes_Mi355(void* data, N_Vector v, N_Vector z)
{
  const SUNAMD_TridiagData* P = (const SUNAMD_TridiagData*)data;
  const long n                = (long)P->N;
  if (n < 2) { return -1; }
  hipStream_t s = (hipStream_t)N_VGetStream_Mi355(z);
  k_tridiag<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(
    n, N_VGetDeviceArrayPointer(v), N_VGetDeviceArrayPointer(P->s1),
    N_VGetDeviceArrayPointer(P->s2), N_VGetDeviceArrayPointer(z));
  return launch_check("ATimes");
}

extern "C" int SUNAMD_TridiagPSetup_Mi355(void* data)
{
  (void)data;
  return 0;
}

extern "C" int SUNAMD_TridiagPSolve_Mi355(void* data, N_Vector r, N_Vector z, sunrealtype tol,
                                          int lr)
{
  (void)tol;
  (void)lr;
  const SUNAMD_TridiagData* P = (const SUNAMD_TridiagData*)data;
  const long n                = (long)P->N;
  hipStream_t s               = (hipStream_t)N_VGetStream_Mi355(z);
  k_jacobi<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, N_VGetDeviceArrayPointer(r),
                                                       N_VGetDeviceArrayPointer(P->d),
                                                       N_VGetDeviceArrayPointer(z));
  return launch_check("PSolve");
}
