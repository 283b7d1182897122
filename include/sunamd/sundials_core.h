/* -----------------------------------------------------------------------------
 * sundials_core.h -- the generic-object layer (context, dispatch wrappers,
 * empty constructors) that integrators call.  Same names, arguments and
 * fallback semantics as the reference's generic layer, so integrator code
 * reads the same; this library supplies them where no SUNDIALS core library
 * is installed (e.g. on a GPU host without the reference build).
 *
 *   SUNContext_Create/Free/GetLastError   src/sundials/sundials_context.c
 *   N_VNewEmpty/N_VFreeEmpty/N_VCopyOps   src/sundials/sundials_nvector.c:46-278
 *   N_V* dispatch + fused/array fallbacks src/sundials/sundials_nvector.c:286-831
 *   N_VCloneVectorArray etc.              src/sundials/sundials_nvector.c:1042-1090
 *   SUNMat* wrappers                      src/sundials/sundials_matrix.c
 *   SUNLinSol* wrappers                   src/sundials/sundials_linearsolver.c
 *   SUNNonlinSol* wrappers                src/sundials/sundials_nonlinearsolver.c
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_SUNDIALS_CORE_H
#define SUNAMD_SUNDIALS_CORE_H

#include "sundials_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* context */
SUNErrCode SUNContext_Create(SUNComm comm, SUNContext* sunctx_out);
SUNErrCode SUNContext_Free(SUNContext* ctx);
SUNErrCode SUNContext_GetLastError(SUNContext sunctx);
SUNErrCode SUNContext_PeekLastError(SUNContext sunctx);
/* records an error on the context (used by void/real-valued ops) */
void SUNAMD_SetLastError(SUNContext sunctx, SUNErrCode err);

/* N_Vector generic layer */
N_Vector N_VNewEmpty(SUNContext sunctx);
void N_VFreeEmpty(N_Vector v);
SUNErrCode N_VCopyOps(N_Vector w, N_Vector v);
N_Vector_ID N_VGetVectorID(N_Vector w);
N_Vector N_VClone(N_Vector w);
N_Vector N_VCloneEmpty(N_Vector w);
void N_VDestroy(N_Vector v);
void N_VSpace(N_Vector v, sunindextype* lrw, sunindextype* liw);
sunrealtype* N_VGetArrayPointer(N_Vector v);
sunrealtype* N_VGetDeviceArrayPointer(N_Vector v);
void N_VSetArrayPointer(sunrealtype* v_data, N_Vector v);
SUNComm N_VGetCommunicator(N_Vector v);
sunindextype N_VGetLength(N_Vector v);
sunindextype N_VGetLocalLength(N_Vector v);

void N_VLinearSum(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z);
void N_VConst(sunrealtype c, N_Vector z);
void N_VProd(N_Vector x, N_Vector y, N_Vector z);
void N_VDiv(N_Vector x, N_Vector y, N_Vector z);
void N_VScale(sunrealtype c, N_Vector x, N_Vector z);
void N_VAbs(N_Vector x, N_Vector z);
void N_VInv(N_Vector x, N_Vector z);
void N_VAddConst(N_Vector x, sunrealtype b, N_Vector z);
sunrealtype N_VDotProd(N_Vector x, N_Vector y);
sunrealtype N_VMaxNorm(N_Vector x);
sunrealtype N_VWrmsNorm(N_Vector x, N_Vector w);
sunrealtype N_VWrmsNormMask(N_Vector x, N_Vector w, N_Vector id);
sunrealtype N_VMin(N_Vector x);
sunrealtype N_VWL2Norm(N_Vector x, N_Vector w);
sunrealtype N_VL1Norm(N_Vector x);
void N_VCompare(sunrealtype c, N_Vector x, N_Vector z);
sunbooleantype N_VInvTest(N_Vector x, N_Vector z);
sunbooleantype N_VConstrMask(N_Vector c, N_Vector x, N_Vector m);
sunrealtype N_VMinQuotient(N_Vector num, N_Vector denom);

SUNErrCode N_VLinearCombination(int nvec, sunrealtype* c, N_Vector* X, N_Vector z);
SUNErrCode N_VScaleAddMulti(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y,
                            N_Vector* Z);
SUNErrCode N_VDotProdMulti(int nvec, N_Vector x, N_Vector* Y, sunrealtype* dotprods);

SUNErrCode N_VLinearSumVectorArray(int nvec, sunrealtype a, N_Vector* X,
                                   sunrealtype b, N_Vector* Y, N_Vector* Z);
SUNErrCode N_VScaleVectorArray(int nvec, sunrealtype* c, N_Vector* X, N_Vector* Z);
SUNErrCode N_VConstVectorArray(int nvec, sunrealtype c, N_Vector* Z);
SUNErrCode N_VWrmsNormVectorArray(int nvec, N_Vector* X, N_Vector* W,
                                  sunrealtype* nrm);
SUNErrCode N_VWrmsNormMaskVectorArray(int nvec, N_Vector* X, N_Vector* W,
                                      N_Vector id, sunrealtype* nrm);
SUNErrCode N_VScaleAddMultiVectorArray(int nvec, int nsum, sunrealtype* a,
                                       N_Vector* X, N_Vector** Y, N_Vector** Z);
SUNErrCode N_VLinearCombinationVectorArray(int nvec, int nsum, sunrealtype* c,
                                           N_Vector** X, N_Vector* Z);

sunrealtype N_VDotProdLocal(N_Vector x, N_Vector y);
sunrealtype N_VMaxNormLocal(N_Vector x);
sunrealtype N_VMinLocal(N_Vector x);
sunrealtype N_VL1NormLocal(N_Vector x);
sunrealtype N_VWSqrSumLocal(N_Vector x, N_Vector w);
sunrealtype N_VWSqrSumMaskLocal(N_Vector x, N_Vector w, N_Vector id);
sunbooleantype N_VInvTestLocal(N_Vector x, N_Vector z);
sunbooleantype N_VConstrMaskLocal(N_Vector c, N_Vector x, N_Vector m);
sunrealtype N_VMinQuotientLocal(N_Vector num, N_Vector denom);
SUNErrCode N_VDotProdMultiLocal(int nvec, N_Vector x, N_Vector* Y,
                                sunrealtype* dotprods);
SUNErrCode N_VDotProdMultiAllReduce(int nvec_total, N_Vector x, sunrealtype* sum);
SUNErrCode N_VBufSize(N_Vector x, sunindextype* size);
SUNErrCode N_VBufPack(N_Vector x, void* buf);
SUNErrCode N_VBufUnpack(N_Vector x, void* buf);

N_Vector* N_VNewVectorArray(int count, SUNContext sunctx);
N_Vector* N_VCloneEmptyVectorArray(int count, N_Vector w);
N_Vector* N_VCloneVectorArray(int count, N_Vector w);
void N_VDestroyVectorArray(N_Vector* vs, int count);
void N_VPrint(N_Vector v);
void N_VPrintFile(N_Vector v, FILE* outfile);

/* SUNMatrix generic layer */
SUNMatrix SUNMatNewEmpty(SUNContext sunctx);
void SUNMatFreeEmpty(SUNMatrix A);
SUNErrCode SUNMatCopyOps(SUNMatrix A, SUNMatrix B);
SUNMatrix_ID SUNMatGetID(SUNMatrix A);
SUNMatrix SUNMatClone(SUNMatrix A);
void SUNMatDestroy(SUNMatrix A);
SUNErrCode SUNMatZero(SUNMatrix A);
SUNErrCode SUNMatCopy(SUNMatrix A, SUNMatrix B);
SUNErrCode SUNMatScaleAdd(sunrealtype c, SUNMatrix A, SUNMatrix B);
SUNErrCode SUNMatScaleAddI(sunrealtype c, SUNMatrix A);
SUNErrCode SUNMatMatvecSetup(SUNMatrix A);
SUNErrCode SUNMatMatvec(SUNMatrix A, N_Vector x, N_Vector y);
SUNErrCode SUNMatSpace(SUNMatrix A, long int* lenrw, long int* leniw);

/* SUNLinearSolver generic layer */
SUNLinearSolver SUNLinSolNewEmpty(SUNContext sunctx);
void SUNLinSolFreeEmpty(SUNLinearSolver S);
SUNLinearSolver_Type SUNLinSolGetType(SUNLinearSolver S);
SUNLinearSolver_ID SUNLinSolGetID(SUNLinearSolver S);
SUNErrCode SUNLinSolSetATimes(SUNLinearSolver S, void* A_data, SUNATimesFn ATimes);
SUNErrCode SUNLinSolSetPreconditioner(SUNLinearSolver S, void* P_data,
                                      SUNPSetupFn Pset, SUNPSolveFn Psol);
SUNErrCode SUNLinSolSetScalingVectors(SUNLinearSolver S, N_Vector s1, N_Vector s2);
SUNErrCode SUNLinSolSetZeroGuess(SUNLinearSolver S, sunbooleantype onoff);
SUNErrCode SUNLinSolInitialize(SUNLinearSolver S);
int SUNLinSolSetup(SUNLinearSolver S, SUNMatrix A);
int SUNLinSolSolve(SUNLinearSolver S, SUNMatrix A, N_Vector x, N_Vector b,
                   sunrealtype tol);
int SUNLinSolNumIters(SUNLinearSolver S);
sunrealtype SUNLinSolResNorm(SUNLinearSolver S);
N_Vector SUNLinSolResid(SUNLinearSolver S);
sunindextype SUNLinSolLastFlag(SUNLinearSolver S);
SUNErrCode SUNLinSolSpace(SUNLinearSolver S, long int* lenrwLS, long int* leniwLS);
SUNErrCode SUNLinSolFree(SUNLinearSolver S);

/* SUNNonlinearSolver generic layer */
SUNNonlinearSolver SUNNonlinSolNewEmpty(SUNContext sunctx);
void SUNNonlinSolFreeEmpty(SUNNonlinearSolver NLS);
SUNNonlinearSolver_Type SUNNonlinSolGetType(SUNNonlinearSolver NLS);
SUNErrCode SUNNonlinSolInitialize(SUNNonlinearSolver NLS);
int SUNNonlinSolSetup(SUNNonlinearSolver NLS, N_Vector y, void* mem);
int SUNNonlinSolSolve(SUNNonlinearSolver NLS, N_Vector y0, N_Vector y, N_Vector w,
                      sunrealtype tol, sunbooleantype callLSetup, void* mem);
SUNErrCode SUNNonlinSolFree(SUNNonlinearSolver NLS);
SUNErrCode SUNNonlinSolSetSysFn(SUNNonlinearSolver NLS, SUNNonlinSolSysFn SysFn);
SUNErrCode SUNNonlinSolSetLSetupFn(SUNNonlinearSolver NLS, SUNNonlinSolLSetupFn SetupFn);
SUNErrCode SUNNonlinSolSetLSolveFn(SUNNonlinearSolver NLS, SUNNonlinSolLSolveFn SolveFn);
SUNErrCode SUNNonlinSolSetConvTestFn(SUNNonlinearSolver NLS,
                                     SUNNonlinSolConvTestFn CTestFn, void* ctest_data);
SUNErrCode SUNNonlinSolSetMaxIters(SUNNonlinearSolver NLS, int maxiters);
SUNErrCode SUNNonlinSolGetNumIters(SUNNonlinearSolver NLS, long int* niters);
SUNErrCode SUNNonlinSolGetCurIter(SUNNonlinearSolver NLS, int* iter);
SUNErrCode SUNNonlinSolGetNumConvFails(SUNNonlinearSolver NLS, long int* nconvfails);

/* Newton nonlinear solver (src/sunnonlinsol/newton/sunnonlinsol_newton.c) */
SUNNonlinearSolver SUNNonlinSol_Newton(N_Vector y, SUNContext sunctx);
/* No counterpart in the reference: the solve loop hands LSolve the residual
   itself instead of its negation (for an LSolve that solves J x = -b); used by
   the integrator's fused mode with the block LU's one-pass solve.
   SUN_ERR_ARG_INCOMPATIBLE for another solver type. */
SUNErrCode SUNNonlinSol_NewtonSetNegatedLSolve(SUNNonlinearSolver NLS, sunbooleantype on);

/* small math helpers with the reference's semantics (sundials_math.c:37-52,
   sundials_math.h:83,107,235) */
sunrealtype SUNRpowerI(sunrealtype base, int exponent);
sunbooleantype SUNRCompare(sunrealtype a, sunrealtype b);
sunbooleantype SUNRCompareTol(sunrealtype a, sunrealtype b, sunrealtype tol);

#ifdef __cplusplus
}
#endif

#endif
