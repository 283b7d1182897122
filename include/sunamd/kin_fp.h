/* -----------------------------------------------------------------------------
 * kin_fp.h -- the fixed-point problem that drives KINSOL's Anderson-
 * accelerated fixed-point path (a user problem, not a reference interface;
 * modelled on examples/kinsol/serial/kinAnalytic_fp.c, whose 3-equation map
 * is replicated over independent cells, n = 3 ncells unknowns).
 *
 * Variant 0: the example's map (cos, sin, exp); variant 1: the same with
 * cos, sin and exp replaced by their first Taylor terms (only +, -, *, /,
 * sqrt: device and CPU evaluations agree bit for bit).  The damping function
 * (when enabled) is the example's: 0.5 at depth 0, else
 * 0.9 - 0.5 sqrt(1 - ||Q^T f||^2 / ||f||^2); the depth function keeps depth
 * 1 for the first two iterations.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_KIN_FP_H
#define SUNAMD_KIN_FP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  long int ncells;
  int variant;        /* 0 or 1 */
  int perturb;        /* per-cell perturbation of the initial guess */
  double tol;         /* KINSetFuncNormTol */
  long int maxiter;   /* KINSetNumMaxIters */
  long int m_aa;      /* KINSetMAA */
  long int delay_aa;  /* KINSetDelayAA */
  int orth_aa;        /* KINSetOrthAA */
  double damping_fp;  /* KINSetDamping (m_aa == 0) */
  double damping_aa;  /* KINSetDampingAA (m_aa > 0) */
  int damping_fn;     /* attach the damping function */
  int depth_fn;       /* attach the depth function */
  int sequential;     /* MI355X vector only: reductions in serial order */
  int fused;          /* MI355X vector only: KINSetUseFusedKernels (scaled norm) */
} SUNAMD_KinOptions;

/* stats (doubles): return flag, nni, nfe, fnorm, wall seconds */
#define SUNAMD_KIN_STATS 5

/* MI355X path (libsundials_amd.so): KINSol(KIN_FP) on an MI355X vector with
   the map evaluated by a device kernel; fills stats and, if u is not NULL,
   the final iterate (3 ncells, host).  Returns the KINSol flag. */
int SUNAMD_KinSolve_Mi355(const SUNAMD_KinOptions* o, double* stats, double* u);

#ifdef __cplusplus
}
#endif

#endif
