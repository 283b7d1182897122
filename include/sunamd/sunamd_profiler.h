/* -----------------------------------------------------------------------------
 * sunamd_profiler.h -- HIP-event timing of selected kernel families, and the
 * N_Vector microbenchmark.
 *
 * The reference times with host wall clocks around whole functions
 * (SUNProfiler, src/sundials/sundials_profiler.c:171-612) and needs
 * synchronous launches to attribute GPU time.  Here, when enabled, the
 * library records a HIP event pair around each launch of a timed kernel
 * family on the stream it is launched on; durations are read back when
 * queried, so timing does not serialise the stream.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_PROFILER_H
#define SUNAMD_PROFILER_H

#include "sundials_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* timed kernel families */
#define SUNAMD_K_GETRS     0 /* batched block LU solve */
#define SUNAMD_K_GETRF     1 /* batched block LU factorisation */
#define SUNAMD_K_LINSUM    2 /* N_VLinearSum */
#define SUNAMD_K_WRMS      3 /* N_VWrmsNorm */
#define SUNAMD_K_DOTMULTI  4 /* N_VDotProdMulti */
#define SUNAMD_K_RHS       5 /* mass-action right-hand side */
#define SUNAMD_K_BDF_BATCH 6 /* batched per-system BDF */
#define SUNAMD_K_ADR_FE    7 /* advection-diffusion explicit RHS (IMEX problem) */
#define SUNAMD_K_SCALE     8 /* N_VScale (also copies and negations) */
#define SUNAMD_K_COUNT     9

void SUNAMD_ProfilerEnable(int on);
void SUNAMD_ProfilerReset(void);
/* total milliseconds and launch count recorded for a family (synchronises
   on the recorded events) */
int SUNAMD_ProfilerGet(int family, double* total_ms, long int* count);

/* N_Vector microbenchmark on n FP64 values: x, y ~ U(-1,1), w ~ U(1e-3,1),
   generated on the device.  Average milliseconds per call over `reps`
   timed calls (after one warm-up), measured with HIP events on the vector
   stream: out_ms[0] LinearSum z = a x + b y, [1] WrmsNorm(x, w),
   [2] DotProdMulti(nvec_multi, x, Y), [3] LinearSum y = x + y (in place),
   [4] DotProd(x, y); then the copy z = x that bounds them: [5] in the vector
   kernels' chunked form (256 threads, 8 pairs per lane, 32 workgroups per
   CU), [6] the same with 1024 threads and 8 workgroups per CU (the fastest
   copy of a sweep of shapes), [7] hipMemcpyAsync device to device.  out_ms
   holds 8 values; n must be even. */
int SUNAMD_BenchVectorOps(sunindextype n, int nvec_multi, int reps, double* out_ms);

/* block LU factorisation benchmark: nblocks blocks of m x m (m = 32 or 64,
   entries U(-1,1)/m plus the identity, generated on the device); average ms
   per factorisation of the whole batch over reps launches after a warm-up.
   variant 0: FP64 register-row kernel, 1: FP32 register-row kernel, 2: FP32
   blocked kernel with MFMA trailing updates. */
int SUNAMD_BenchBlockLU(long int nblocks, int m, int variant, int reps, double* ms_out);

#ifdef __cplusplus
}
#endif

#endif
