/* -----------------------------------------------------------------------------
 * Block-diagonal SUNMatrix in MI355X HBM and its batched LU linear solver.
 *
 * Matrix: nblocks dense m x m blocks, block-contiguous, column-major inside a
 * block (block k, column j, row i at k*m*m + j*m + i) -- the layout of the
 * reference's batched precedent SUNMatrix_MagmaDenseBlock
 * (include/sunmatrix/sunmatrix_magmadense.h:41-69,
 *  src/sunmatrix/magmadense/sunmatrix_magmadense.cpp:84-110).
 * Ops fill the reference SUNMatrix table (include/sundials/sundials_matrix.h:
 * 89-105) with the dense matrix's per-element arithmetic
 * (src/sunmatrix/dense/sunmatrix_dense.c:229-334).  The vector of a block
 * system is the concatenation of the per-block vectors.
 *
 * Solver (SUNLINEARSOLVER_DIRECT): replaces SUNLinSol_MagmaDense
 * (src/sunlinsol/magmadense/sunlinsol_magmadense.cpp:314-440, which calls
 * magma_dgetrf/getrs_batched) with hand-written HIP batched LU.  In FP64 mode
 * the factors, pivots and solutions are bit-for-bit those of the reference
 * dense GETRF / GETRS (src/sundials/sundials_dense.c:85-186) applied block by
 * block.  In FP32 mode the blocks are factored in single precision and the
 * solve runs in single precision on FP64 right-hand sides (the Newton
 * iteration around it stays FP64).  A zero pivot in any block makes setup
 * return SUNLS_LUFACT_FAIL, with SUNLinSolLastFlag = (global column index of
 * the first zero pivot) + 1, like the dense solver.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_SUNMATRIX_MI355BLOCK_H
#define SUNAMD_SUNMATRIX_MI355BLOCK_H

#include "sundials_core.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SUNAMD_LU_FP64 0
#define SUNAMD_LU_FP32 1

/* stream: the hipStream_t the owning vectors run on (N_VGetStream_Mi355) */
SUNMatrix SUNMatrix_Mi355BlockDense(sunindextype nblocks, sunindextype m, void* stream,
                                    SUNContext sunctx);
sunindextype SUNMatrix_Mi355BlockDense_NumBlocks(SUNMatrix A);
sunindextype SUNMatrix_Mi355BlockDense_BlockSize(SUNMatrix A);
sunrealtype* SUNMatrix_Mi355BlockDense_DeviceData(SUNMatrix A);
/* host mirror (allocated on first use) and synchronous copies */
sunrealtype* SUNMatrix_Mi355BlockDense_HostData(SUNMatrix A);
SUNErrCode SUNMatrix_Mi355BlockDense_CopyToDevice(SUNMatrix A);
SUNErrCode SUNMatrix_Mi355BlockDense_CopyFromDevice(SUNMatrix A);

SUNLinearSolver SUNLinSol_Mi355BlockLU(N_Vector y, SUNMatrix A, int precision,
                                       SUNContext sunctx);
/* FP32 factors of 32 x 32 and 64 x 64 blocks by the blocked kernel whose
   trailing updates run on MFMA (v_mfma_f32_16x16x4f32; panels of 16 columns,
   the block staged in LDS) instead of the register-row kernel; no
   counterpart in the reference (MAGMA's batched GETRF plays this part for
   sunlinsol_magmadense.cpp:339).  Same pivots and factor layout, so the FP32
   solve is shared.  SUN_ERR_ARG_INCOMPATIBLE for FP64 solvers or other m. */
SUNErrCode SUNLinSol_Mi355BlockLUSetMFMA(SUNLinearSolver S, int onoff);
/* device pivot array (int32, nblocks*m, pivot row of step k of block b at b*m+k) */
int* SUNLinSol_Mi355BlockLU_DevicePivots(SUNLinearSolver S);
/* copy the pivots to a host array of nblocks*m ints (synchronous) */
SUNErrCode SUNLinSol_Mi355BlockLU_GetPivots(SUNLinearSolver S, int* host_out);
/* The Newton-matrix setup in one pass over the blocks, in place of the three
   calls of cvode_ls.c:1311-1395 (cvLsLinSys + cvLsSetup):
     save_to_J == 0:  SUNMatCopy(J, A); SUNMatScaleAddI(c, A); SUNLinSolSetup(S, A)
     save_to_J != 0:  SUNMatCopy(A, J); SUNMatScaleAddI(c, A); SUNLinSolSetup(S, A)
   with the same per-element roundings, so the factors, pivots and what A
   holds afterwards equal the three calls' bit for bit.  Returns what
   SUNLinSolSetup returns (0 or SUNLS_LUFACT_FAIL), or SUN_ERR_ARG_INCOMPATIBLE
   without doing anything when the three calls must run instead: S not a
   block LU of A's shape, J not a block matrix of that shape or J == A, a
   matrix on another stream, m outside 1-10, 12, 16, 24, 32, 48, 64, or the
   MFMA factorisation switched on.  No counterpart in the reference, whose
   matrix and solver ops are separate calls. */
int SUNLinSol_Mi355BlockLUSetupScaled(SUNLinearSolver S, SUNMatrix A, SUNMatrix J, sunrealtype c,
                                      int save_to_J);
/* The direct solve in one pass, in place of the calls of cvode_ls.c:1744-1784
   (cvLsSolve): N_VConst(0, x); SUNLinSolSolve(S, A, x, b); N_VScale(1, x, b);
   and, when the solution is scaled, N_VScale(scale, b, b).  b is solved in
   place and each entry is written as scale * x (N_VScale's product; scale = 1
   leaves x unchanged), so b ends bit-identical to the calls' result (x itself
   is not written).  Returns SUN_SUCCESS, or SUN_ERR_ARG_INCOMPATIBLE without
   doing anything when the calls must run instead: S not a block LU of A's
   shape, b not an MI355X vector on the solver's stream, or m outside 1-10,
   12, 16, 24, 32, 48, 64.  No counterpart in the reference. */
int SUNLinSol_Mi355BlockLUSolveScaled(SUNLinearSolver S, SUNMatrix A, N_Vector b,
                                      sunrealtype scale);

#ifdef __cplusplus
}
#endif

#endif
