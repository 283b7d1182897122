/* -----------------------------------------------------------------------------
 * kinsol_sunamd.h -- KINSOL: Newton iteration with or without a line search
 * (KIN_NONE, KIN_LINESEARCH) over a matrix-based direct linear solver, and the
 * fixed-point iteration with Anderson acceleration (KIN_FP).
 *
 * Public names, argument meaning and return codes follow the reference
 * KINSOL API (include/kinsol/kinsol.h:38-79, 100-200; kinsol_ls.h:39-80) so a
 * driver written against KINSOL runs unchanged.  Provided: Newton with
 * inequality constraints, the alpha/beta line search, residual monitoring and
 * modified-Newton setup control, a user Jacobian or the dense difference-
 * quotient Jacobian; KIN_FP with optional damping (constant or user
 * function), Anderson acceleration of depth m with delay, a user depth
 * function, and the four QR updates of the acceleration space
 * (KIN_ORTH_MGS, _ICWY, _CGS2, _DCGS2).  Picard, iterative (matrix-free)
 * linear solvers and preconditioners are not part of this library:
 * KINSol returns KIN_ILL_INPUT for KIN_PICARD and KINSetLinearSolver
 * KINLS_ILL_INPUT for a solver that is not SUNLINEARSOLVER_DIRECT.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_KINSOL_H
#define SUNAMD_KINSOL_H

#include "sundials_core.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KIN_SUCCESS          0
#define KIN_INITIAL_GUESS_OK 1
#define KIN_STEP_LT_STPTOL   2
#define KIN_WARNING 99
#define KIN_MEM_NULL            -1
#define KIN_ILL_INPUT           -2
#define KIN_NO_MALLOC           -3
#define KIN_MEM_FAIL            -4
#define KIN_LINESEARCH_NONCONV  -5
#define KIN_MAXITER_REACHED     -6
#define KIN_MXNEWT_5X_EXCEEDED  -7
#define KIN_LINESEARCH_BCFAIL   -8
#define KIN_LINSOLV_NO_RECOVERY -9
#define KIN_LINIT_FAIL          -10
#define KIN_LSETUP_FAIL         -11
#define KIN_LSOLVE_FAIL         -12
#define KIN_SYSFUNC_FAIL        -13
#define KIN_FIRST_SYSFUNC_ERR   -14
#define KIN_REPTD_SYSFUNC_ERR   -15
#define KIN_VECTOROP_ERR        -16
#define KIN_DAMPING_FN_ERR      -18
#define KIN_DEPTH_FN_ERR        -19

#define KIN_ORTH_MGS   0
#define KIN_ORTH_ICWY  1
#define KIN_ORTH_CGS2  2
#define KIN_ORTH_DCGS2 3

#define KIN_NONE       0
#define KIN_LINESEARCH 1
#define KIN_PICARD     2
#define KIN_FP         3

/* kinsol_ls.h:39-48 */
#define KINLS_SUCCESS     0
#define KINLS_MEM_NULL    -1
#define KINLS_LMEM_NULL   -2
#define KINLS_ILL_INPUT   -3
#define KINLS_MEM_FAIL    -4
#define KINLS_PMEM_NULL   -5
#define KINLS_JACFUNC_ERR -6
#define KINLS_SUNMAT_FAIL -7
#define KINLS_SUNLS_FAIL  -8

/* kinsol.h:84-98 */
typedef int (*KINSysFn)(N_Vector uu, N_Vector fval, void* user_data);
typedef int (*KINDampingFn)(long int iter, N_Vector u_val, N_Vector g_val,
                            sunrealtype* qt_fn, long int depth, void* user_data,
                            sunrealtype* damping_factor);
typedef int (*KINDepthFn)(long int iter, N_Vector u_val, N_Vector g_val,
                          N_Vector f_val, N_Vector* df, sunrealtype* R_mat,
                          long int depth, void* user_data, long int* new_depth,
                          sunbooleantype* remove_indices);
/* kinsol_ls.h:54-56 */
typedef int (*KINLsJacFn)(N_Vector u, N_Vector fu, SUNMatrix J, void* user_data,
                          N_Vector tmp1, N_Vector tmp2);

void* KINCreate(SUNContext sunctx);
int KINInit(void* kinmem, KINSysFn func, N_Vector tmpl);
int KINSetUserData(void* kinmem, void* user_data);
int KINSetMAA(void* kinmem, long int maa);
int KINSetDelayAA(void* kinmem, long int delay);
int KINSetOrthAA(void* kinmem, int orthaa);
int KINSetDampingAA(void* kinmem, sunrealtype beta);
int KINSetDamping(void* kinmem, sunrealtype beta);
int KINSetDampingFn(void* kinmem, KINDampingFn damping_fn);
int KINSetDepthFn(void* kinmem, KINDepthFn depth_fn);
int KINSetFuncNormTol(void* kinmem, sunrealtype fnormtol);
int KINSetNumMaxIters(void* kinmem, long int mxiter);
int KINSetReturnNewest(void* kinmem, sunbooleantype ret_newest);
/* Newton options (kinsol.h:120-160); constraints are copied, 0 entries are
   free, +-1 mean >= 0 / <= 0 and +-2 mean > 0 / < 0 */
int KINSetConstraints(void* kinmem, N_Vector constraints);
int KINSetScaledStepTol(void* kinmem, sunrealtype scsteptol);
int KINSetMaxSetupCalls(void* kinmem, long int msbset);
int KINSetMaxSubSetupCalls(void* kinmem, long int msbsetsub);
int KINSetMaxNewtonStep(void* kinmem, sunrealtype mxnewtstep);
int KINSetNoInitSetup(void* kinmem, sunbooleantype noInitSetup);
int KINSetNoResMon(void* kinmem, sunbooleantype noResMon);
int KINSetMaxBetaFails(void* kinmem, long int mxnbcf);
int KINSetRelErrFunc(void* kinmem, sunrealtype relfunc);
int KINSetResMonParams(void* kinmem, sunrealtype omegamin, sunrealtype omegamax);
int KINSetResMonConstVal(void* kinmem, sunrealtype omegaconst);
/* linear solver interface (kinsol_ls.h:62-80); LS must be a direct solver
   and A its matrix.  Without KINSetJacFn the Jacobian is the dense
   difference quotient, which needs a SUNMATRIX_DENSE A and a host-addressable
   vector (N_VGetArrayPointer / N_VSetArrayPointer). */
int KINSetLinearSolver(void* kinmem, SUNLinearSolver LS, SUNMatrix A);
int KINSetJacFn(void* kinmem, KINLsJacFn jac);
/* No counterpart in the reference: evaluate KINSOL's scaled norms
   (KINScFNorm, KINScSNorm) as one fused device pass each when the vectors are
   MI355X vectors (libsundials_amd.so loaded); results are bit-identical to the
   op chains.  Returns KIN_ILL_INPUT (and leaves the op chains on) otherwise. */
int KINSetUseFusedKernels(void* kinmem, sunbooleantype onoff);
/* strategy KIN_NONE, KIN_LINESEARCH or KIN_FP; KIN_FP does not use u_scale,
   and its f_scale weights the max norm of G(u) - u */
int KINSol(void* kinmem, N_Vector uu, int strategy, N_Vector u_scale, N_Vector f_scale);
int KINGetNumNonlinSolvIters(void* kinmem, long int* nniters);
int KINGetNumFuncEvals(void* kinmem, long int* nfevals);
int KINGetFuncNorm(void* kinmem, sunrealtype* fnorm);
int KINGetNumBetaCondFails(void* kinmem, long int* nbcfails);
int KINGetNumBacktrackOps(void* kinmem, long int* nbacktr);
int KINGetStepLength(void* kinmem, sunrealtype* steplength);
int KINGetNumJacEvals(void* kinmem, long int* njevals);
int KINGetNumLinFuncEvals(void* kinmem, long int* nfevals);
int KINGetLastLinFlag(void* kinmem, long int* flag);
void KINFree(void** kinmem);

#ifdef __cplusplus
}
#endif

#endif
