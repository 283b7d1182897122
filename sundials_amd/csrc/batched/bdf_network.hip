// -----------------------------------------------------------------------------
// The batched integrator's in-kernel network evaluation, exposed on host
// arrays so that it can be checked against the training-side forward pass
// and a CPU reference (the integrator itself runs it inside its kernels).
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <cstdio>

#include "bdf_device.h"
#include "sunamd/learned_predictor.h"

// The integrator's in-kernel network evaluation (w_forward16) on rows of
// host inputs: one wavefront per 16 rows, the rows' inputs rounded to bf16
// into its LDS network area exactly as the integrator writes them.
__global__ void __launch_bounds__(64) k_mlp_wave(const unsigned short* __restrict__ W,
                                                 const float* __restrict__ bias,
                                                 const float* __restrict__ x, int rows,
                                                 float* __restrict__ y)
{
  __shared__ __align__(16) unsigned short area[sunamd_bdf::MLP_WAVE_BYTES / 2];
  const int lane = threadIdx.x;
  const int r0   = blockIdx.x * 16;
  for (int t = lane; t < sunamd_bdf::MLP_WAVE_BYTES / 2; t += 64) { area[t] = 0; }
  __syncthreads();
  for (int t = lane; t < 16 * 144; t += 64)
  {
    const int r = t / 144, k = t % 144;
    if (r0 + r < rows) { area[r * sunamd_bdf::MLP_XS + k] = sunamd_bdf::bf16_bits(x[(long)(r0 + r) * 144 + k]); }
  }
  __syncthreads();
  const float* Y = sunamd_bdf::w_forward16(W, bias, area);
  for (int t = lane; t < 16 * 64; t += 64)
  {
    const int r = t / 64;
    if (r0 + r < rows) { y[(long)(r0 + r) * 64 + t % 64] = Y[t]; }
  }
}

extern "C" int SUNAMD_MLPForwardWave(const void* weights, const float* x, int rows, float* y)
{
  if (!weights || !x || !y || rows < 1) { return -1; }
  float *dx = nullptr, *dy = nullptr;
  hipError_t e = hipMalloc((void**)&dx, sizeof(float) * 144 * (size_t)rows);
  if (e == hipSuccess) { e = hipMalloc((void**)&dy, sizeof(float) * 64 * (size_t)rows); }
  if (e == hipSuccess) { e = hipMemcpy(dx, x, sizeof(float) * 144 * (size_t)rows, hipMemcpyHostToDevice); }
  if (e == hipSuccess)
  {
    const unsigned short* W = (const unsigned short*)weights;
    k_mlp_wave<<<(rows + 15) / 16, 64>>>(W, (const float*)(W + SUNAMD_MLP_WEIGHTS_BF16), dx, rows,
                                        dy);
    e = hipGetLastError();
  }
  if (e == hipSuccess) { e = hipMemcpy(y, dy, sizeof(float) * 64 * (size_t)rows, hipMemcpyDeviceToHost); }
  if (dx) { (void)hipFree(dx); }
  if (dy) { (void)hipFree(dy); }
  if (e != hipSuccess)
  {
    fprintf(stderr, "[sundials_amd batched_bdf] network evaluation: %s\n", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

