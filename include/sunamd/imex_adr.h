/* -----------------------------------------------------------------------------
 * imex_adr.h -- the advection-diffusion-reaction problem that drives the
 * ARKStep IMEX + SPGMR + block-Jacobi path (a user problem, not a reference
 * interface; modelled on benchmarks/advection_reaction_3D/kokkos/rhs3D.hpp,
 * whose per-cell 3x3 block-Jacobi preconditioner is rhs3D.hpp:337-426).
 *
 *   periodic grid of n^3 cells on [0,1)^3, spacing dx = 1/n, 3 species
 *   (u, v, w) stored cell by cell: y[3*cell + s], cell = i + n*(j + n*k)
 *   fe (explicit): first-order upwind advection with velocity (cx, cy, cz)
 *       (all >= 0: backward differences) + second-order diffusion d,
 *       per species
 *   fi (implicit, cell-local Brusselator with a stiff third species):
 *       u' = A - (w + 1) u + u^2 v
 *       v' = w u - u^2 v
 *       w' = kappa (B - w) - w u,   kappa = 10^U(0, 5) per cell
 *   preconditioner: P = I - gamma J_fi per cell (3x3, partial pivoting),
 *       J_fi taken at the state of the last refresh (jok = false)
 *   Jacobian-vector product: J_fi(y) v per cell
 *   initial state: u = A + 0.1 g, v = B/A + 0.1 g, w = B + 0.1 g with g a
 *       sum of 4 periodic Gaussian bumps (width 0.1) at seeded centres
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_IMEX_ADR_H
#define SUNAMD_IMEX_ADR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  int n;                  /* cells per dimension */
  int64_t ncells;         /* n^3 */
  double cx, cy, cz, d;   /* advection velocity, diffusion */
  double A, B;            /* Brusselator feed rates */
  double* kappa;          /* ncells reaction rates */
  double* y0;             /* 3 * ncells initial state */
  uint64_t seed;
} SUNAMD_Adr;

/* the problem of the IMEX configuration: c = (1e-3, 1e-3, 1e-3), d = 1e-2,
   A = 0.6, B = 2 */
SUNAMD_Adr* SUNAMD_AdrCreate(int n, uint64_t seed);
void SUNAMD_AdrFree(SUNAMD_Adr* p);

/* statistics of one integration (doubles): nst, nst_attempts, netf, ncfn,
   nfe, nfi, nni, nnf, nsetups, nli, npe, nps, njtimes, ncfl, wall seconds,
   last h */
#define SUNAMD_ADR_STATS 16

/* options of an ARKStep solve of the problem */
typedef struct
{
  double rtol, atol, tf;
  int maxl;           /* SPGMR Krylov dimension (restarts 0) */
  int gstype;         /* SUN_MODIFIED_GS (1) or SUN_CLASSICAL_GS (2) */
  int controller;     /* 0 = I (the ARKStep default), 1 = PI, 2 = PID */
  int sequential;     /* MI355X vector only: reductions in serial order */
  int prec_fp32;      /* MI355X only: preconditioner LU in single precision */
  long int mxsteps;
} SUNAMD_AdrOptions;

/* MI355X path (libsundials_amd.so): ARKStep IMEX (ARK3(2)4L[2]SA) + SPGMR
   (left block-Jacobi preconditioning, no restarts) on MI355X vectors, the
   callbacks as device kernels; integrates [0, tf] (stop time tf), fills
   stats and, if yout is not NULL, the final state (3 ncells, host).
   Returns 0 on success, the ARKodeEvolve flag otherwise. */
int SUNAMD_AdrSolve_Mi355(const SUNAMD_Adr* p, const SUNAMD_AdrOptions* o, double* stats,
                          double* yout);

#ifdef __cplusplus
}
#endif

#endif
