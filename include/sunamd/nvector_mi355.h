/* -----------------------------------------------------------------------------
 * nvector_mi355.h -- N_Vector whose data lives in MI355X (gfx950) HBM.
 *
 * Fills the reference N_Vector ops table (include/sundials/sundials_nvector.h:
 * 101-195) so CVODE / ARKODE / KINSOL use it unchanged.  It replaces the
 * reference HIP vector's constructors (src/nvector/hip/nvector_hip.hip.cpp:
 * N_VNew_Hip :220, N_VMake_Hip :370, N_VCopyToDevice_Hip / N_VCopyFromDevice_Hip
 * :500-560) with the same roles:
 *   - every standard op, the fused LinearCombination / ScaleAddMulti /
 *     DotProdMulti and the vector-array ops are device kernels;
 *   - reductions finish on the device and hand one value back through a
 *     pinned host slot (one stream synchronisation per reduction, no
 *     per-call allocation or pointer-array copies);
 *   - N_VGetArrayPointer returns a host mirror (allocated on first use),
 *     N_VGetDeviceArrayPointer the HBM array; the mirror is synchronised only
 *     by N_VCopyToDevice_Mi355 / N_VCopyFromDevice_Mi355.
 * N_VGetVectorID returns SUNDIALS_NVEC_CUSTOM.
 * Elementwise ops round exactly like the reference serial vector (same
 * formula per element, no fused multiply-add); reductions sum in a fixed
 * tree order, so they are deterministic but not bit-equal to a sequential sum.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_NVECTOR_MI355_H
#define SUNAMD_NVECTOR_MI355_H

#include "sundials_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* new vector of the given length (device data allocated, host mirror lazy).
   All vectors made this way on one device share one stream, so work issued
   from one host thread is ordered; clones share their parent's stream. */
N_Vector N_VNew_Mi355(sunindextype length, SUNContext sunctx);
/* new vector whose ops (and its clones') run on the caller's hipStream_t */
N_Vector N_VNewWithStream_Mi355(sunindextype length, void* stream, SUNContext sunctx);
/* vector with no data */
N_Vector N_VNewEmpty_Mi355(sunindextype length, SUNContext sunctx);
/* wrap caller-owned device (and optionally host) arrays */
N_Vector N_VMake_Mi355(sunindextype length, sunrealtype* h_vdata,
                       sunrealtype* d_vdata, SUNContext sunctx);

/* host <-> device copies of the whole vector (synchronous) */
SUNErrCode N_VCopyToDevice_Mi355(N_Vector v);
SUNErrCode N_VCopyFromDevice_Mi355(N_Vector v);
/* copy raw host data in / out without a host mirror (synchronous) */
SUNErrCode N_VCopyFromHostArray_Mi355(N_Vector v, const sunrealtype* src);
SUNErrCode N_VCopyToHostArray_Mi355(N_Vector v, sunrealtype* dst);

/* the HIP stream (hipStream_t) all ops of this vector and its clones run on */
void* N_VGetStream_Mi355(N_Vector v);
/* wait for all work queued on the vector's stream */
SUNErrCode N_VSynchronize_Mi355(N_Vector v);
/* enable (default) or disable the fused and vector-array ops */
SUNErrCode N_VEnableFusedOps_Mi355(N_Vector v, sunbooleantype tf);

/* number of HIP kernel launches issued by this vector family (for profiling) */
long int N_VGetNumLaunches_Mi355(N_Vector v);

/* Summation order of the sum-type reductions (DotProd, WrmsNorm[Mask],
   WL2Norm, L1Norm, DotProdMulti); min / max are exact in any order.
   TREE (default): fixed device-wide tree whose shape depends only on the
     vector length -- deterministic, full bandwidth.
   SEQUENTIAL: left-to-right, the serial vector's order
     (nvector_serial.c:605-761) -- bit-identical to it, latency-bound
     (one lane adds every term); for exact-parity runs.
   Set on a vector before cloning it; clones inherit the order. */
#define SUNAMD_REDUCE_TREE       0
#define SUNAMD_REDUCE_SEQUENTIAL 1
SUNErrCode N_VSetReductionOrder_Mi355(N_Vector v, int order);
/* HIP failures: ops that return SUNErrCode return SUN_ERR_EXT_FAIL, and the
   first failure on a vector's execution context (its stream) is also kept
   there, because void ops (N_VLinearSum, ...) cannot return it: after one,
   reductions return NaN until it is read.  Returns that hipError_t code (0:
   none) and clears it; -1 for a vector that is not an MI355X vector.  (The
   reference reports through SUNContext's last error; this library never
   writes into the SUNContext, whose layout belongs to the linked core.) */
int N_VGetHipError_Mi355(N_Vector v);

/* Classical Gram-Schmidt step on MI355X vectors with SUNClassicalGS's
   arguments, coefficients and reorthogonalisation rule
   (include/sundials/sundials_iterative.h:231-233): the k projections and
   ||v_k||^2 are summed into device memory, the update v_k -= sum h_i v_i
   reads them there and sums the new norm in the same pass, and one
   synchronisation returns the Hessenberg column -- 8(2k+3) bytes per element
   and one host round trip instead of 8(2k+5) and two.  Vectors in
   sequential reduction order take SUNClassicalGS itself.  Install with
   SUNLinSol_SPGMRSetClassicalGSFn. */
SUNErrCode SUNAMD_ClassicalGS_Mi355(N_Vector* v, sunrealtype** h, int k, int p,
                                    sunrealtype* new_vk_norm, sunrealtype* stemp,
                                    N_Vector* vtemp);

/* 1 if v is an MI355X vector of this library, else 0 (no counterpart) */
int SUNAMD_NVectorIsMi355(N_Vector v);

/* CVODE's fused error-weight and Newton-residual kernels for this vector,
   with the arguments of cvEwtSetSS_fused / cvEwtSetSV_fused /
   cvNlsResid_fused (src/cvode/cvode_fused_gpu.cpp:81, 131, 232); results equal
   the unfused op sequences bit for bit.  They return -1 for a vector that is
   not an MI355X vector.  CVodeSetUseIntegratorFusedKernels switches the
   repo's CVODE onto them. */
int cvEwtSetSS_fused_Mi355(sunbooleantype atolMin0, sunrealtype reltol, sunrealtype Sabstol,
                           N_Vector ycur, N_Vector tempv, N_Vector weight);
int cvEwtSetSV_fused_Mi355(sunbooleantype atolMin0, sunrealtype reltol, N_Vector Vabstol,
                           N_Vector ycur, N_Vector tempv, N_Vector weight);
int cvNlsResid_fused_Mi355(sunrealtype rl1, sunrealtype ngamma, N_Vector zn1, N_Vector ycor,
                           N_Vector ftemp, N_Vector res);

/* No counterpart in the reference (which fuses neither): the Nordsieck
   predictor and its undo of cvPredict / cvRestore (src/cvode/cvode.c:2748,
   2791) as one kernel over zn[0..q], q in 1..5 -- zn[j-1] = zn[j-1] +- zn[j]
   for k = 1..q, j = q..k, element by element in that order, so the result
   equals the q(q+1)/2 LinearSums bit for bit.  Used by the repo's CVODE under
   CVodeSetUseIntegratorFusedKernels; -1 for vectors that are not MI355X
   vectors. */
int cvPredict_fused_Mi355(int q, N_Vector* zn);
int cvRestore_fused_Mi355(int q, N_Vector* zn);

/* No counterpart in the reference (KINSOL has no fused kernels): KINSOL's
   scaled norms (src/kinsol/kinsol.c:2038-2062) in one pass each,
   KINScSNorm = max_i |v_i / (1/uscale_i + |u_i|)| and
   KINScFNorm = max_i |fscale_i v_i|, bit-identical to the op chains.  Used
   by the repo's KINSOL under KINSetUseFusedKernels; negative for vectors that
   are not MI355X vectors or on a HIP error.  Scratch is per device and shared
   by all calls: calls that overlap on one device race. */
sunrealtype kinScSNorm_fused_Mi355(N_Vector v, N_Vector u, N_Vector uscale);
sunrealtype kinScFNorm_fused_Mi355(N_Vector v, N_Vector fscale);

#ifdef __cplusplus
}
#endif

#endif
