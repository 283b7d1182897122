/* -----------------------------------------------------------------------------
 * CPU ORACLE -- test infrastructure only.
 *
 * Plain-C restatement of the reference's serial N_Vector and dense LU, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker for the HIP path.  Nothing in the product library links or calls
 * this code.
 *
 *   N_VNew_Serial & ops          src/nvector/serial/nvector_serial.c:80-1592
 *   SUNDlsMat_denseGETRF/GETRS   src/sundials/sundials_dense.c:85-186
 *   SUNDenseMatrix & ops         src/sunmatrix/dense/sunmatrix_dense.c:229-334
 *   SUNLinSol_Dense              src/sunlinsol/dense/sunlinsol_dense.c:128-183
 *   Block-dense variants: per-block application of the same GETRF/GETRS;
 *   equal bit-for-bit to the full-dense routines on a block-diagonal matrix
 *   (strict '>' pivot search never picks an off-block zero and zero updates
 *   are skipped, sundials_dense.c:99-145) -- pinned by tests/test_oracle.py.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_ORACLE_H
#define SUNAMD_ORACLE_H

#include "sunamd/sundials_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* serial vector */
N_Vector N_VNew_Serial(sunindextype length, SUNContext sunctx);
N_Vector N_VNewEmpty_Serial(sunindextype length, SUNContext sunctx);
N_Vector N_VMake_Serial(sunindextype length, sunrealtype* v_data, SUNContext sunctx);
SUNErrCode N_VEnableFusedOps_Serial(N_Vector v, sunbooleantype tf);

/* dense LU kernels (column pointers) */
sunindextype SUNDlsMat_denseGETRF(sunrealtype** a, sunindextype m, sunindextype n,
                                  sunindextype* p);
void SUNDlsMat_denseGETRS(sunrealtype** a, sunindextype n, sunindextype* p,
                          sunrealtype* b);
/* same kernels on one column-major block (lda = m) */
sunindextype oracle_block_getrf(sunrealtype* a, sunindextype m, sunindextype* p);
void oracle_block_getrs(const sunrealtype* a, sunindextype m, const sunindextype* p,
                        sunrealtype* b);

/* dense matrix + dense direct solver */
SUNMatrix SUNDenseMatrix(sunindextype M, sunindextype N, SUNContext sunctx);
sunrealtype* SUNDenseMatrix_Data(SUNMatrix A);
sunrealtype** SUNDenseMatrix_Cols(SUNMatrix A);
sunindextype SUNDenseMatrix_Rows(SUNMatrix A);
sunindextype SUNDenseMatrix_Columns(SUNMatrix A);
SUNLinearSolver SUNLinSol_Dense(N_Vector y, SUNMatrix A, SUNContext sunctx);

/* block-diagonal dense matrix (nblocks blocks of m x m, block-contiguous,
   column-major per block) + per-block dense direct solver */
SUNMatrix SUNBlockDenseMatrix_Oracle(sunindextype nblocks, sunindextype m,
                                     SUNContext sunctx);
sunrealtype* SUNBlockDenseMatrix_Oracle_Data(SUNMatrix A);
SUNLinearSolver SUNLinSol_BlockDense_Oracle(N_Vector y, SUNMatrix A, SUNContext sunctx);

/* host N_Vector microbenchmark (bench_port.c): ms per call of [0] LinearSum
   z = 0.3x + 0.7y, [1] WrmsNorm, [2] DotProdMulti(nvec), [3] y = x + y,
   [4] DotProd; nthreads 1 = serial vector, > 1 = OpenMP loops */
int oracle_bench_vector_ops(long n, int nvec, int reps, int nthreads, double* out_ms);

#ifdef __cplusplus
}
#endif

#endif
