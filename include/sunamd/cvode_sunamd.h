/* -----------------------------------------------------------------------------
 * cvode_sunamd.h -- BDF integrator with modified Newton and a matrix-based
 * linear solver interface: the caller of the vector / block-matrix hot path.
 *
 * Public names, argument meaning and return codes follow the reference
 * CVODE API (include/cvode/cvode.h:38-100, include/cvode/cvode_ls.h:54-80) so
 * a driver written against CVODE runs unchanged.  Only the BDF method with
 * scalar or vector absolute tolerances, a user Jacobian and a direct
 * (matrix-based) or matrix-free linear solver is provided; rootfinding,
 * projection, constraints, stability-limit detection and the Adams method
 * are not part of this library.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_CVODE_H
#define SUNAMD_CVODE_H

#include "sundials_core.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CV_ADAMS 1
#define CV_BDF   2

#define CV_NORMAL   1
#define CV_ONE_STEP 2

#define CV_SUCCESS      0
#define CV_TSTOP_RETURN 1
#define CV_ROOT_RETURN  2
#define CV_WARNING 99
#define CV_TOO_MUCH_WORK -1
#define CV_TOO_MUCH_ACC  -2
#define CV_ERR_FAILURE   -3
#define CV_CONV_FAILURE  -4
#define CV_LINIT_FAIL        -5
#define CV_LSETUP_FAIL       -6
#define CV_LSOLVE_FAIL       -7
#define CV_RHSFUNC_FAIL      -8
#define CV_FIRST_RHSFUNC_ERR -9
#define CV_REPTD_RHSFUNC_ERR -10
#define CV_UNREC_RHSFUNC_ERR -11
#define CV_RTFUNC_FAIL       -12
#define CV_NLS_INIT_FAIL     -13
#define CV_NLS_SETUP_FAIL    -14
#define CV_CONSTR_FAIL       -15
#define CV_NLS_FAIL          -16
#define CV_MEM_FAIL     -20
#define CV_MEM_NULL     -21
#define CV_ILL_INPUT    -22
#define CV_NO_MALLOC    -23
#define CV_BAD_K        -24
#define CV_BAD_T        -25
#define CV_BAD_DKY      -26
#define CV_TOO_CLOSE    -27
#define CV_VECTOROP_ERR -28

#define CVLS_SUCCESS         0
#define CVLS_MEM_NULL        -1
#define CVLS_LMEM_NULL       -2
#define CVLS_ILL_INPUT       -3
#define CVLS_MEM_FAIL        -4
#define CVLS_PMEM_NULL       -5
#define CVLS_JACFUNC_UNRECVR -6
#define CVLS_JACFUNC_RECVR   -7
#define CVLS_SUNMAT_FAIL     -8
#define CVLS_SUNLS_FAIL      -9

typedef int (*CVRhsFn)(sunrealtype t, N_Vector y, N_Vector ydot, void* user_data);
typedef int (*CVEwtFn)(N_Vector y, N_Vector ewt, void* user_data);
typedef int (*CVLsJacFn)(sunrealtype t, N_Vector y, N_Vector fy, SUNMatrix Jac,
                         void* user_data, N_Vector tmp1, N_Vector tmp2,
                         N_Vector tmp3);
/* optional initial Newton correction (the reference always starts from zero,
   cvode.c:3095); returning nonzero falls back to the zero correction */
typedef int (*CVInitialCorrectionFn)(sunrealtype t, sunrealtype h, sunrealtype gamma,
                                     N_Vector ypred, N_Vector acor,
                                     void* user_data);

void* CVodeCreate(int lmm, SUNContext sunctx);
int CVodeInit(void* cvode_mem, CVRhsFn f, sunrealtype t0, N_Vector y0);
int CVodeReInit(void* cvode_mem, sunrealtype t0, N_Vector y0);
int CVodeSStolerances(void* cvode_mem, sunrealtype reltol, sunrealtype abstol);
int CVodeSVtolerances(void* cvode_mem, sunrealtype reltol, N_Vector abstol);
int CVodeWFtolerances(void* cvode_mem, CVEwtFn efun);
int CVodeSetUserData(void* cvode_mem, void* user_data);
int CVodeSetMaxNumSteps(void* cvode_mem, long int mxsteps);
int CVodeSetMaxOrd(void* cvode_mem, int maxord);
int CVodeSetInitStep(void* cvode_mem, sunrealtype hin);
int CVodeSetMinStep(void* cvode_mem, sunrealtype hmin);
int CVodeSetMaxStep(void* cvode_mem, sunrealtype hmax);
int CVodeSetLSetupFrequency(void* cvode_mem, long int msbp);
int CVodeSetStopTime(void* cvode_mem, sunrealtype tstop);
int CVodeSetNonlinearSolver(void* cvode_mem, SUNNonlinearSolver NLS);
int CVodeSetLinearSolver(void* cvode_mem, SUNLinearSolver LS, SUNMatrix A);
int CVodeSetJacFn(void* cvode_mem, CVLsJacFn jac);
int CVodeSetInitialCorrectionFn(void* cvode_mem, CVInitialCorrectionFn fn);
/* cvode_io.c:1006: fused error-weight and residual kernels (MI355X vectors) */
int CVodeSetUseIntegratorFusedKernels(void* cvode_mem, sunbooleantype onoff);

/* the x^y of the step-size ratios (SUNRpowerR, cvode.c:3394 ...): correctly
   rounded evaluation shared with the device integrator (no reference
   counterpart) */
double SUNAMD_Pow(double x, double y);
/* diagnostics of its two paths (no reference counterpart): the fast path's
   approximation (hi + lo) 2^scale with its error bound bound 2^scale, and
   whether it decided the rounding (return 1, *res the result); the
   double-double path alone */
int SUNAMD_PowFastParts(double x, double y, double* res, double* hi, double* lo, double* bound,
                        int* scale);
double SUNAMD_PowExact(double x, double y);

int CVode(void* cvode_mem, sunrealtype tout, N_Vector yout, sunrealtype* tret,
          int itask);
int CVodeGetDky(void* cvode_mem, sunrealtype t, int k, N_Vector dky);

int CVodeGetNumSteps(void* cvode_mem, long int* nsteps);
int CVodeGetNumRhsEvals(void* cvode_mem, long int* nfevals);
int CVodeGetNumLinSolvSetups(void* cvode_mem, long int* nlinsetups);
int CVodeGetNumErrTestFails(void* cvode_mem, long int* netfails);
int CVodeGetNumNonlinSolvIters(void* cvode_mem, long int* nniters);
int CVodeGetNumNonlinSolvConvFails(void* cvode_mem, long int* nnfails);
int CVodeGetNumStepSolveFails(void* cvode_mem, long int* ncnf);
int CVodeGetNumJacEvals(void* cvode_mem, long int* njevals);
int CVodeGetLastOrder(void* cvode_mem, int* qlast);
int CVodeGetCurrentOrder(void* cvode_mem, int* qcur);
int CVodeGetActualInitStep(void* cvode_mem, sunrealtype* hinused);
int CVodeGetLastStep(void* cvode_mem, sunrealtype* hlast);
int CVodeGetCurrentStep(void* cvode_mem, sunrealtype* hcur);
int CVodeGetCurrentTime(void* cvode_mem, sunrealtype* tcur);
int CVodeGetNumInitialCorrectionsUsed(void* cvode_mem, long int* nused);

/* all counters at once: nst, nfe, nsetups, netf, nni, nnf, ncfn, nje, qu, qcur,
   hinused, hu, hcur, tcur (order of the reference stats table) */
int CVodeGetAllStats(void* cvode_mem, long int* counts /* [8] */,
                     int* orders /* [2] */, sunrealtype* steps /* [4] */);

/* every statistic, as CVodePrintAllStats (cvode_io.c:1626-1714) writes it:
   SUN_OUTPUTFORMAT_TABLE one "name = value" line each (names padded to 29
   columns, reals "%.15g"), SUN_OUTPUTFORMAT_CSV "name,value" pairs joined by
   commas (reals "% .15e"), with no line end, as the reference writes them;
   rootfinding is not part of this library, so its "Root fn evals" entry is 0 */
int CVodePrintAllStats(void* cvode_mem, FILE* outfile, SUNOutputFormat fmt);

void CVodeFree(void** cvode_mem);

#ifdef __cplusplus
}
#endif

#endif
