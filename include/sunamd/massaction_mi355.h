/* -----------------------------------------------------------------------------
 * massaction_mi355.h -- the mass-action workload on the MI355X: device copy of
 * the problem, right-hand side and block Jacobian as HIP kernels wrapped as
 * CVODE callbacks (CVRhsFn / CVLsJacFn, include/cvode/cvode.h:93,
 * include/cvode/cvode_ls.h:54), and a driver that integrates all cells as one
 * coupled system through the N_Vector / block-LU drop-ins -- the structure of
 * the reference's batched example (examples/cvode/magma/
 * cv_bruss_batched_magma.cpp: one vector holding every cell, a block-diagonal
 * Jacobian, BDF + Newton + batched direct solve).
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_MASSACTION_MI355_H
#define SUNAMD_MASSACTION_MI355_H

#include "massaction.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SUNAMD_MassActionDev_ SUNAMD_MassActionDev;

/* copy the problem to HBM; kernels run on `stream` (a hipStream_t).  NULL for
 * m outside 1..64, a species index outside 0..m-1, or a network whose
 * per-cell Jacobian work area (m*m + m + 2R doubles, plus 8R + m + 2 ints of
 * tables) exceeds 64 KiB of LDS (R up to ~640 at m = 64) */
SUNAMD_MassActionDev* SUNAMD_MassActionToDevice(const SUNAMD_MassAction* p, void* stream);
void SUNAMD_MassActionDevFree(SUNAMD_MassActionDev* d);

/* CVODE callbacks: y, ydot are N_VNew_Mi355 vectors, J a SUNMatrix_Mi355BlockDense */
int SUNAMD_MassActionRhsDev(sunrealtype t, N_Vector y, N_Vector ydot, void* user_data);
int SUNAMD_MassActionJacDev(sunrealtype t, N_Vector y, N_Vector fy, SUNMatrix J,
                            void* user_data, N_Vector tmp1, N_Vector tmp2, N_Vector tmp3);

/* raw device-array evaluation (ncells*m vectors, ncells*m*m blocks) */
int SUNAMD_MassActionRhsDevArrays(SUNAMD_MassActionDev* d, const double* dy, double* df);
int SUNAMD_MassActionJacDevArrays(SUNAMD_MassActionDev* d, const double* dy, double* dJ);

typedef struct
{
  long int nst, nfe, nsetups, netf, nni, nnf, ncfn, nje;
  int qlast, qcur;
  double hinused, hlast, hcur, tcur;
  double wall_s;     /* wall time of the CVode call (device idle at both ends) */
  long int launches; /* N_Vector kernel launches during the call */
  int flag;          /* CVode return flag */
} SUNAMD_SolveStats;

/* integrate every cell of p from t = 0 to tout as one coupled system:
   N_VNew_Mi355 + SUNMatrix_Mi355BlockDense + SUNLinSol_Mi355BlockLU
   (lu_precision SUNAMD_LU_FP64 / SUNAMD_LU_FP32), scalar tolerances,
   reductions in SUNAMD_REDUCE_TREE or SUNAMD_REDUCE_SEQUENTIAL order.
   yout (ncells*m, host) receives y(tout). */
int SUNAMD_MassActionSolveCoupled_Mi355(const SUNAMD_MassAction* p, double rtol,
                                        double atol, double tout, int lu_precision,
                                        int reduction_order, double* yout,
                                        SUNAMD_SolveStats* stats);

/* the same in three parts, so repeated integrations reuse the resident
   problem, vectors, matrices and integrator: create (allocate + upload),
   run (reset y to y0 on the device, re-initialise, integrate to tout),
   free.  run may be called any number of times. */
typedef struct SUNAMD_CoupledRun_ SUNAMD_CoupledRun;
SUNAMD_CoupledRun* SUNAMD_CoupledCreate_Mi355(const SUNAMD_MassAction* p, double rtol,
                                              double atol, int lu_precision,
                                              int reduction_order);
int SUNAMD_CoupledRun_Mi355(SUNAMD_CoupledRun* r, double tout, SUNAMD_SolveStats* stats);
int SUNAMD_CoupledGetState_Mi355(SUNAMD_CoupledRun* r, double* yout);
void SUNAMD_CoupledFree_Mi355(SUNAMD_CoupledRun* r);
/* switch the integrator onto the fused error-weight / residual kernels
   (CVodeSetUseIntegratorFusedKernels); results are unchanged bit for bit */
int SUNAMD_CoupledSetFusedKernels_Mi355(SUNAMD_CoupledRun* r, int onoff);
/* Device stepping: SUNAMD_CoupledRun_Mi355 then integrates with one
   cooperative launch in which every decision of the integrator -- Newton
   convergence, Jacobian reuse, the error test, step size and order -- is
   taken on the device, with no host synchronisation between steps (the
   step / Newton / error-test logic of cvode.c cvStep, cvode_nls.c and
   sunnonlinsol_newton.c restated per workgroup; each WRMS norm a grid-wide
   reduction in a fixed order).  Per cell it streams the state through one
   fused pass per Newton iteration and one per accepted step.  With
   SUNAMD_REDUCE_SEQUENTIAL the norms are summed in element order and the
   run equals the host-stepped one bit for bit (counters, step sizes, final
   state); with SUNAMD_REDUCE_TREE the order is fixed by the launch shape.
   The saved Jacobian is kept as the state it was taken at; the block
   matrix's storage holds the factors and the solver's pivot array the row
   permutations.  Supported for m in 1-8, 10, 12, 16, 24, 32, 48, 64 (other m:
   SUNAMD_CoupledRun_Mi355 returns -100).  No counterpart in the reference,
   whose integrators decide on the host after every norm. */
int SUNAMD_CoupledSetDeviceStepping_Mi355(SUNAMD_CoupledRun* r, int onoff);
/* launch shape of the last device-stepped run: workgroups of the
   cooperative grid, threads per workgroup, dynamic LDS bytes per workgroup
   (-1 before any run) */
int SUNAMD_CoupledDeviceLastShape(int* workgroups, int* threads, int* lds_bytes);

#ifdef __cplusplus
}
#endif

#endif
