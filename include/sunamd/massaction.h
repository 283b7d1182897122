/* -----------------------------------------------------------------------------
 * massaction.h -- seeded synthetic mass-action chemistry used as the
 * integration workload (independent cells, block-diagonal Jacobian).
 *
 * Network: m species, R reversible bimolecular reactions A + B <-> C + D with
 * the four indices drawn uniformly from [0, m).  Base rate constants kf0, kb0
 * are log-uniform on [1e-2, 1e6].  Each cell multiplies every one of its 2R
 * rate constants by its own lognormal(0, sigma) factor (default sigma 0.2).
 * y0 per cell: uniform on [0.1, 1], normalised to sum 1 (summed in species
 * order).  Random numbers come from a counter-based generator (SplitMix64
 * finaliser over (seed, stream, counter)), so any cell can be generated
 * independently; normals use Box-Muller.
 *
 * Per reaction the net rate is w = (kf*y_a)*y_b - (kb*y_c)*y_d, applied as
 * f_a -= w, f_b -= w, f_c += w, f_d += w in reaction order.  The Jacobian of a
 * cell is accumulated in reaction order, for each reaction over the target
 * rows (a, b, c, d) and, for each, over the columns (a, b, c, d) with partials
 * kf*y_b, kf*y_a, -(kb*y_d), -(kb*y_c).  The device kernels keep this order,
 * so CPU and GPU evaluations agree bit-for-bit.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_MASSACTION_H
#define SUNAMD_MASSACTION_H

#include "sundials_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  int m;            /* species per cell */
  int R;            /* reversible reactions */
  int64_t ncells;
  uint64_t seed;
  double sigma;     /* lognormal width of the per-cell multipliers */
  int* idx;         /* 4*R: a, b, c, d per reaction */
  double* kf0;      /* R base forward rates */
  double* kb0;      /* R base backward rates */
  double* kf;       /* ncells*R, cell-major */
  double* kb;       /* ncells*R, cell-major */
  double* y0;       /* ncells*m, cell-major */
} SUNAMD_MassAction;

/* uniform double in (0, 1) from the counter-based generator */
double SUNAMD_Uniform(uint64_t seed, uint64_t stream, uint64_t counter);

SUNAMD_MassAction* SUNAMD_MassActionCreate(int m, int R, int64_t ncells, uint64_t seed,
                                           double sigma);
/* same network and multipliers, but only cells [cell0, cell0+ncells) of the
   seed's infinite cell sequence (for sampling and sharding) */
SUNAMD_MassAction* SUNAMD_MassActionCreateRange(int m, int R, int64_t cell0,
                                                int64_t ncells, uint64_t seed,
                                                double sigma);
/* tile: cell i of the result is cell (i mod base->ncells) of base */
SUNAMD_MassAction* SUNAMD_MassActionTile(const SUNAMD_MassAction* base, int64_t ncells);
void SUNAMD_MassActionFree(SUNAMD_MassAction* p);

/* CPU evaluation on host arrays for cells [c0, c1): y, f are ncells*m;
   J is ncells*m*m column-major per cell */
void SUNAMD_MassActionRhsHost(const SUNAMD_MassAction* p, const double* y, double* f,
                              int64_t c0, int64_t c1);
void SUNAMD_MassActionJacHost(const SUNAMD_MassAction* p, const double* y, double* J,
                              int64_t c0, int64_t c1);

#ifdef __cplusplus
}
#endif

#endif
