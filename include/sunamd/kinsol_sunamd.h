/* -----------------------------------------------------------------------------
 * kinsol_sunamd.h -- KINSOL's fixed-point iteration with Anderson acceleration
 * (the KIN_FP strategy): the caller of the vector hot path for nonlinear
 * systems written as u = G(u).
 *
 * Public names, argument meaning and return codes follow the reference
 * KINSOL API (include/kinsol/kinsol.h:38-79, 100-200) so a driver written
 * against KINSOL's fixed-point interface runs unchanged.  Provided: KIN_FP
 * with optional damping (constant or user function), Anderson acceleration
 * of depth m with delay, a user depth function, and the four QR updates of
 * the acceleration space (KIN_ORTH_MGS, _ICWY, _CGS2, _DCGS2).  Newton,
 * line search, Picard and the linear-solver interface are not part of this
 * library: KINSol returns KIN_ILL_INPUT for those strategies.
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
#define KIN_MAXITER_REACHED     -6
#define KIN_SYSFUNC_FAIL        -13
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

/* kinsol.h:84-98 */
typedef int (*KINSysFn)(N_Vector uu, N_Vector fval, void* user_data);
typedef int (*KINDampingFn)(long int iter, N_Vector u_val, N_Vector g_val,
                            sunrealtype* qt_fn, long int depth, void* user_data,
                            sunrealtype* damping_factor);
typedef int (*KINDepthFn)(long int iter, N_Vector u_val, N_Vector g_val,
                          N_Vector f_val, N_Vector* df, sunrealtype* R_mat,
                          long int depth, void* user_data, long int* new_depth,
                          sunbooleantype* remove_indices);

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
/* strategy KIN_FP only; u_scale is not used by KIN_FP, f_scale weights the
   max norm of G(u) - u */
int KINSol(void* kinmem, N_Vector uu, int strategy, N_Vector u_scale, N_Vector f_scale);
int KINGetNumNonlinSolvIters(void* kinmem, long int* nniters);
int KINGetNumFuncEvals(void* kinmem, long int* nfevals);
int KINGetFuncNorm(void* kinmem, sunrealtype* fnorm);
void KINFree(void** kinmem);

#ifdef __cplusplus
}
#endif

#endif
