/* -----------------------------------------------------------------------------
 * arkode_sunamd.h -- additive Runge-Kutta (ARKStep IMEX) integrator: the
 * caller of the vector / SPGMR / preconditioner hot path for split problems
 * y' = fe(t,y) + fi(t,y) (fe explicit, fi implicit).
 *
 * Public names, argument meaning and return codes follow the reference
 * ARKODE API (include/arkode/arkode.h:188-295, include/arkode/arkode_arkstep.h,
 * include/arkode/arkode_ls.h) so a driver written against ARKStep runs
 * unchanged.  Provided: the IMEX pairs ARK4(3)7L[2]SA (order 4, the
 * reference's default) and ARK3(2)4L[2]SA (order 3), adaptive or fixed steps
 * with the I, PI or PID controller, the trivial stage predictor, linearly
 * implicit and autonomous problem options, Newton iterations through any
 * SUNNonlinearSolver's callbacks and a matrix-free iterative linear solver
 * (SPGMR) with a user preconditioner and a user or difference-quotient
 * Jacobian-vector product, ARK_NORMAL / ARK_ONE_STEP, stop time, Hermite or
 * no interpolation bookkeeping (no dense output).  Mass matrices, relaxation,
 * rootfinding, constraints, MRI forcing, adjoints and other tables are not
 * part of this library.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_ARKODE_H
#define SUNAMD_ARKODE_H

#include "sundials_core.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ARK_NORMAL   1
#define ARK_ONE_STEP 2

#define ARK_INTERP_NONE    -1
#define ARK_INTERP_HERMITE 0

#define ARK_SUCCESS      0
#define ARK_TSTOP_RETURN 1
#define ARK_ROOT_RETURN  2
#define ARK_WARNING 99
#define ARK_TOO_MUCH_WORK    -1
#define ARK_TOO_MUCH_ACC     -2
#define ARK_ERR_FAILURE      -3
#define ARK_CONV_FAILURE     -4
#define ARK_LINIT_FAIL       -5
#define ARK_LSETUP_FAIL      -6
#define ARK_LSOLVE_FAIL      -7
#define ARK_RHSFUNC_FAIL     -8
#define ARK_FIRST_RHSFUNC_ERR -9
#define ARK_REPTD_RHSFUNC_ERR -10
#define ARK_UNREC_RHSFUNC_ERR -11
#define ARK_NLS_INIT_FAIL    -14
#define ARK_NLS_SETUP_FAIL   -15
#define ARK_MEM_FAIL         -20
#define ARK_MEM_NULL         -21
#define ARK_ILL_INPUT        -22
#define ARK_NO_MALLOC        -23
#define ARK_TOO_CLOSE        -27
#define ARK_VECTOROP_ERR     -28
#define ARK_NLS_OP_ERR       -32
#define ARK_USER_PREDICT_FAIL -39
#define ARK_CONTROLLER_ERR   -47

#define ARKLS_SUCCESS   0
#define ARKLS_MEM_NULL  -1
#define ARKLS_LMEM_NULL -2
#define ARKLS_ILL_INPUT -3
#define ARKLS_MEM_FAIL  -4

typedef int (*ARKRhsFn)(sunrealtype t, N_Vector y, N_Vector ydot, void* user_data);
/* arkode_ls.h:73-96 */
typedef int (*ARKLsPrecSetupFn)(sunrealtype t, N_Vector y, N_Vector fy,
                                sunbooleantype jok, sunbooleantype* jcurPtr,
                                sunrealtype gamma, void* user_data);
typedef int (*ARKLsPrecSolveFn)(sunrealtype t, N_Vector y, N_Vector fy, N_Vector r,
                                N_Vector z, sunrealtype gamma, sunrealtype delta,
                                int lr, void* user_data);
typedef int (*ARKLsJacTimesSetupFn)(sunrealtype t, N_Vector y, N_Vector fy,
                                    void* user_data);
typedef int (*ARKLsJacTimesVecFn)(N_Vector v, N_Vector Jv, sunrealtype t, N_Vector y,
                                  N_Vector fy, void* user_data, N_Vector tmp);

/* arkode.h:188: a user stage predictor, called for every implicit stage
   after the built-in (trivial) prediction with the stage time and the
   prediction zpred, which it may overwrite; < 0: unrecoverable failure
   (ARK_USER_PREDICT_FAIL), > 0: the step attempt is retried as after a
   convergence failure */
typedef int (*ARKStagePredictFn)(sunrealtype t, N_Vector zpred, void* user_data);

/* arkode_arkstep.h: ARKStepCreate(fe, fi, t0, y0, sunctx) */
void* ARKStepCreate(ARKRhsFn fe, ARKRhsFn fi, sunrealtype t0, N_Vector y0,
                    SUNContext sunctx);
void ARKodeFree(void** arkode_mem);
int ARKodeSStolerances(void* arkode_mem, sunrealtype reltol, sunrealtype abstol);
int ARKodeSetUserData(void* arkode_mem, void* user_data);
int ARKodeSetOrder(void* arkode_mem, int maxord);
int ARKodeSetStopTime(void* arkode_mem, sunrealtype tstop);
int ARKodeSetMaxNumSteps(void* arkode_mem, long int mxsteps);
/* arkode.h:295 (ARKodeSetStagePredictFn); NULL removes it */
int ARKodeSetStagePredictFn(void* arkode_mem, ARKStagePredictFn PredictStage);
int ARKodeSetFixedStep(void* arkode_mem, sunrealtype hfixed);
int ARKodeSetInitStep(void* arkode_mem, sunrealtype hin);
int ARKodeSetInterpolantType(void* arkode_mem, int itype);
/* "I", "PI" or "PID" (arkode_io.c:930-1010) */
int ARKodeSetAdaptControllerByName(void* arkode_mem, const char* cname);
int ARKodeSetErrorBias(void* arkode_mem, sunrealtype bias);
int ARKodeSetMaxNonlinIters(void* arkode_mem, int maxcor);
int ARKodeSetNonlinConvCoef(void* arkode_mem, sunrealtype nlscoef);
int ARKodeSetLSetupFrequency(void* arkode_mem, int msbp);
int ARKodeSetNonlinearSolver(void* arkode_mem, SUNNonlinearSolver NLS);
/* linearly implicit fi (one Newton iteration per stage; timedepend: its
   Jacobian depends on t) / undo; an autonomous problem lets the stages'
   first residuals reuse fi(tn, yn) (arkode_arkstep_io.c:853-900) */
int ARKodeSetLinear(void* arkode_mem, int timedepend);
int ARKodeSetNonlinear(void* arkode_mem);
int ARKodeSetAutonomous(void* arkode_mem, sunbooleantype autonomous);
/* matrix-free iterative linear solver (A must be NULL) */
int ARKodeSetLinearSolver(void* arkode_mem, SUNLinearSolver LS, SUNMatrix A);
int ARKodeSetPreconditioner(void* arkode_mem, ARKLsPrecSetupFn psetup,
                            ARKLsPrecSolveFn psolve);
int ARKodeSetJacTimes(void* arkode_mem, ARKLsJacTimesSetupFn jtsetup,
                      ARKLsJacTimesVecFn jtimes);
int ARKodeSetEpsLin(void* arkode_mem, sunrealtype eplifac);
int ARKodeSetJacEvalFrequency(void* arkode_mem, long int msbj);

int ARKodeEvolve(void* arkode_mem, sunrealtype tout, N_Vector yout, sunrealtype* tret,
                 int itask);

int ARKodeGetNumSteps(void* arkode_mem, long int* nsteps);
int ARKodeGetNumStepAttempts(void* arkode_mem, long int* nstep_attempts);
int ARKodeGetNumErrTestFails(void* arkode_mem, long int* netfails);
int ARKodeGetNumConvFails(void* arkode_mem, long int* ncfails);
int ARKStepGetNumRhsEvals(void* arkode_mem, long int* nfe_evals, long int* nfi_evals);
int ARKodeGetNumNonlinSolvIters(void* arkode_mem, long int* nniters);
int ARKodeGetNumNonlinSolvConvFails(void* arkode_mem, long int* nnfails);
int ARKodeGetNumLinSolvSetups(void* arkode_mem, long int* nlinsetups);
int ARKodeGetNumLinIters(void* arkode_mem, long int* nliters);
int ARKodeGetNumLinConvFails(void* arkode_mem, long int* nlcfails);
int ARKodeGetNumPrecEvals(void* arkode_mem, long int* npevals);
int ARKodeGetNumPrecSolves(void* arkode_mem, long int* npsolves);
int ARKodeGetNumJtimesEvals(void* arkode_mem, long int* njvevals);
/* fi evaluations of the difference-quotient Jv (arkode_ls.c:1299) */
int ARKodeGetNumLinRhsEvals(void* arkode_mem, long int* nfevalsLS);
int ARKodeGetLastStep(void* arkode_mem, sunrealtype* hlast);
int ARKodeGetActualInitStep(void* arkode_mem, sunrealtype* hinused);
int ARKodeGetCurrentStep(void* arkode_mem, sunrealtype* hcur);
int ARKodeGetCurrentTime(void* arkode_mem, sunrealtype* tcur);

/* the pair of a method order (ARKodeButcherTable_LoadERK / LoadDIRK of the
   ARKSTEP_DEFAULT_ARK_*TABLE_q pair): order 3 ARK3(2)4L[2]SA
   (arkode_butcher_erk.def:218 / arkode_butcher_dirk.def:248), order 4 or
   <= 0 ARK4(3)7L[2]SA (arkode_butcher_erk.def:371 / arkode_butcher_dirk.def:474).
   implicit_part 0 the explicit table, 1 the ESDIRK; s stages, A (s x s
   row-major, s <= 7), b, c, d, order q, embedding order p */
int SUNAMD_ARKTable(int order, int implicit_part, int* s, double* A, double* b, double* c,
                    double* d, int* q, int* p);
/* the order-3 pair with A 4x4 row-major */
int SUNAMD_ARK324Table(int implicit_part, double* A, double* b, double* c, double* d, int* q,
                       int* p);

#ifdef __cplusplus
}
#endif

#endif
