/* -----------------------------------------------------------------------------
 * Generator of the IMEX advection-diffusion-reaction problem (imex_adr.h):
 * per-cell reaction rates kappa = 10^U(0,5) and the bump initial state, from
 * the counter-based SUNAMD_Uniform stream of the problem seed (stream 11:
 * rates, stream 12: bump centres).
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>

#include "sunamd/imex_adr.h"
#include "sunamd/massaction.h"

#define NBUMPS 4

/* squared periodic distance between two coordinates in [0, 1) */
static double pdist2(double a, double b)
{
  double t = fabs(a - b);
  if (t > 0.5) { t = 1.0 - t; }
  return t * t;
}

SUNAMD_Adr* SUNAMD_AdrCreate(int n, uint64_t seed)
{
  if (n < 2) { return NULL; }
  SUNAMD_Adr* p = (SUNAMD_Adr*)calloc(1, sizeof(SUNAMD_Adr));
  if (p == NULL) { return NULL; }
  p->n      = n;
  p->ncells = (int64_t)n * n * n;
  p->cx = p->cy = p->cz = 1.0e-3;
  p->d                  = 1.0e-2;
  p->A                  = 0.6;
  p->B                  = 2.0;
  p->seed               = seed;
  p->kappa              = (double*)malloc(sizeof(double) * (size_t)p->ncells);
  p->y0                 = (double*)malloc(sizeof(double) * 3 * (size_t)p->ncells);
  if (p->kappa == NULL || p->y0 == NULL)
  {
    SUNAMD_AdrFree(p);
    return NULL;
  }
  double cen[NBUMPS][3];
  for (int b = 0; b < NBUMPS; b++)
  {
    for (int a = 0; a < 3; a++) { cen[b][a] = SUNAMD_Uniform(seed, 12, 3 * (uint64_t)b + a); }
  }
  const double dx = 1.0 / n, s2 = 2.0 * 0.1 * 0.1;
  for (int k = 0; k < n; k++)
  {
    for (int j = 0; j < n; j++)
    {
      for (int i = 0; i < n; i++)
      {
        const int64_t c = i + (int64_t)n * (j + (int64_t)n * k);
        const double x = (i + 0.5) * dx, y = (j + 0.5) * dx, z = (k + 0.5) * dx;
        double g       = 0.0;
        for (int b = 0; b < NBUMPS; b++)
        {
          g += exp(-(pdist2(x, cen[b][0]) + pdist2(y, cen[b][1]) + pdist2(z, cen[b][2])) / s2);
        }
        p->kappa[c]     = pow(10.0, 5.0 * SUNAMD_Uniform(seed, 11, (uint64_t)c));
        p->y0[3 * c]     = p->A + 0.1 * g;
        p->y0[3 * c + 1] = p->B / p->A + 0.1 * g;
        p->y0[3 * c + 2] = p->B + 0.1 * g;
      }
    }
  }
  return p;
}

void SUNAMD_AdrFree(SUNAMD_Adr* p)
{
  if (p == NULL) { return; }
  free(p->kappa);
  free(p->y0);
  free(p);
}
