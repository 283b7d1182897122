/* -----------------------------------------------------------------------------
 * kin_newton.h -- the Newton test problem that drives KINSOL's Newton path
 * (a user problem, not a reference interface; modelled on
 * examples/kinsol/serial/kinFerTron_dns.c, whose 6-unknown constrained system
 * is replicated over independent cells, n = 6 ncells unknowns, one 6 x 6
 * Jacobian block per cell).
 *
 * Variant 0: the example's functions (sin, exp of the math library; CPU
 * only); variant 1: the same with deterministic sin / exp built from IEEE
 * operations, so device and CPU evaluations agree bit for bit.  Constraints
 * l >= 0, L <= 0 on the bound unknowns, fnormtol / scsteptol as given, unit
 * scaling, and the Jacobian either analytic per block (jac = 1: a
 * block-diagonal matrix and batched LU) or the dense difference quotient
 * (jac = 0: CPU only, a dense matrix).
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_KIN_NEWTON_H
#define SUNAMD_KIN_NEWTON_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  long int ncells;
  int variant;       /* 0 or 1 (the MI355X path requires 1) */
  int guess;         /* 1: lower bounds, 2: middle of the box */
  int perturb;       /* per-cell perturbation of the initial guess */
  int strategy;      /* KIN_NONE or KIN_LINESEARCH */
  long int msbset;   /* KINSetMaxSetupCalls (1: exact Newton, 0: default 10) */
  double fnormtol;   /* KINSetFuncNormTol */
  double scsteptol;  /* KINSetScaledStepTol */
  long int maxiter;  /* KINSetNumMaxIters (0: default 200) */
  int jac;           /* 1: analytic block Jacobian; 0: dense DQ (CPU only) */
  int sequential;    /* MI355X vector only: reductions in serial order */
  int fused;         /* MI355X vector only: KINSetUseFusedKernels (scaled norms) */
} SUNAMD_KinNewtonOptions;

/* stats (doubles): return flag, nni, nfe, nje, nfeDQ, nbcf, nbacktrack,
   wall seconds of KINSol */
#define SUNAMD_KINN_STATS 8

/* MI355X path (libsundials_amd.so): KINSol(KIN_NONE / KIN_LINESEARCH) on an
   MI355X vector, the cell functions and Jacobian blocks evaluated by device
   kernels, the Newton systems solved by SUNLinSol_Mi355BlockLU on a
   SUNMatrix_Mi355BlockDense.  Fills stats and, if u is not NULL, the final
   iterate (6 ncells, host).  Returns the KINSol flag, or -100 on a setup
   failure (bad options, allocation, HIP error). */
int SUNAMD_KinNewtonSolve_Mi355(const SUNAMD_KinNewtonOptions* o, double* stats, double* u);

#ifdef __cplusplus
}
#endif

#endif
