/* -----------------------------------------------------------------------------
 * Seeded synthetic mass-action chemistry: generator and CPU evaluation.
 * The generator's draw layout (stream, counter) is fixed so that a cell's
 * data depends only on (seed, m, R, cell index).
 *   stream 0: reaction indices, counter 4r+q
 *   stream 1: base rates, counter 2r (forward) / 2r+1 (backward)
 *   stream 2: per-cell multipliers, counters 2*(cell*2R + q) and +1 (Box-Muller)
 *   stream 3: initial state, counter cell*m + s
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sunamd/massaction.h"

static inline uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

double SUNAMD_Uniform(uint64_t seed, uint64_t stream, uint64_t counter)
{
  uint64_t z = mix64(seed * 0x9E3779B97F4A7C15ULL + mix64((stream + 1) * 0xD1B54A32D192ED03ULL +
                                                         counter));
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static double normal(uint64_t seed, uint64_t stream, uint64_t k)
{
  double u1 = SUNAMD_Uniform(seed, stream, 2 * k);
  double u2 = SUNAMD_Uniform(seed, stream, 2 * k + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static SUNAMD_MassAction* alloc_problem(int m, int R, int64_t ncells)
{
  SUNAMD_MassAction* p = (SUNAMD_MassAction*)calloc(1, sizeof(SUNAMD_MassAction));
  if (!p) { return NULL; }
  p->m      = m;
  p->R      = R;
  p->ncells = ncells;
  p->idx    = (int*)malloc(sizeof(int) * 4 * (size_t)R);
  p->kf0    = (double*)malloc(sizeof(double) * (size_t)R);
  p->kb0    = (double*)malloc(sizeof(double) * (size_t)R);
  p->kf     = (double*)malloc(sizeof(double) * (size_t)R * (size_t)ncells);
  p->kb     = (double*)malloc(sizeof(double) * (size_t)R * (size_t)ncells);
  p->y0     = (double*)malloc(sizeof(double) * (size_t)m * (size_t)ncells);
  if (!p->idx || !p->kf0 || !p->kb0 || !p->kf || !p->kb || !p->y0)
  {
    SUNAMD_MassActionFree(p);
    return NULL;
  }
  return p;
}

SUNAMD_MassAction* SUNAMD_MassActionCreateRange(int m, int R, int64_t cell0,
                                                int64_t ncells, uint64_t seed, double sigma)
{
  if (m <= 0 || R <= 0 || ncells <= 0 || cell0 < 0) { return NULL; }
  SUNAMD_MassAction* p = alloc_problem(m, R, ncells);
  if (!p) { return NULL; }
  p->seed  = seed;
  p->sigma = sigma;
  for (int r = 0; r < R; r++)
  {
    for (int q = 0; q < 4; q++)
    {
      int v = (int)(SUNAMD_Uniform(seed, 0, 4 * (uint64_t)r + q) * m);
      p->idx[4 * r + q] = (v >= m) ? m - 1 : v;
    }
    p->kf0[r] = pow(10.0, -2.0 + 8.0 * SUNAMD_Uniform(seed, 1, 2 * (uint64_t)r));
    p->kb0[r] = pow(10.0, -2.0 + 8.0 * SUNAMD_Uniform(seed, 1, 2 * (uint64_t)r + 1));
  }
  for (int64_t c = 0; c < ncells; c++)
  {
    uint64_t gc = (uint64_t)(cell0 + c);
    for (int r = 0; r < R; r++)
    {
      double mf = exp(sigma * normal(seed, 2, gc * 2 * (uint64_t)R + 2 * (uint64_t)r));
      double mb = exp(sigma * normal(seed, 2, gc * 2 * (uint64_t)R + 2 * (uint64_t)r + 1));
      p->kf[c * R + r] = p->kf0[r] * mf;
      p->kb[c * R + r] = p->kb0[r] * mb;
    }
    double sum = 0.0;
    for (int s = 0; s < m; s++)
    {
      double v          = 0.1 + 0.9 * SUNAMD_Uniform(seed, 3, gc * (uint64_t)m + s);
      p->y0[c * m + s]  = v;
      sum              += v;
    }
    for (int s = 0; s < m; s++) { p->y0[c * m + s] = p->y0[c * m + s] / sum; }
  }
  return p;
}

SUNAMD_MassAction* SUNAMD_MassActionCreate(int m, int R, int64_t ncells, uint64_t seed,
                                           double sigma)
{
  return SUNAMD_MassActionCreateRange(m, R, 0, ncells, seed, sigma);
}

SUNAMD_MassAction* SUNAMD_MassActionTile(const SUNAMD_MassAction* b, int64_t ncells)
{
  if (!b || ncells <= 0) { return NULL; }
  SUNAMD_MassAction* p = alloc_problem(b->m, b->R, ncells);
  if (!p) { return NULL; }
  p->seed  = b->seed;
  p->sigma = b->sigma;
  memcpy(p->idx, b->idx, sizeof(int) * 4 * (size_t)b->R);
  memcpy(p->kf0, b->kf0, sizeof(double) * (size_t)b->R);
  memcpy(p->kb0, b->kb0, sizeof(double) * (size_t)b->R);
  for (int64_t c = 0; c < ncells; c++)
  {
    int64_t s = c % b->ncells;
    memcpy(p->kf + c * b->R, b->kf + s * b->R, sizeof(double) * (size_t)b->R);
    memcpy(p->kb + c * b->R, b->kb + s * b->R, sizeof(double) * (size_t)b->R);
    memcpy(p->y0 + c * b->m, b->y0 + s * b->m, sizeof(double) * (size_t)b->m);
  }
  return p;
}

void SUNAMD_MassActionFree(SUNAMD_MassAction* p)
{
  if (!p) { return; }
  free(p->idx);
  free(p->kf0);
  free(p->kb0);
  free(p->kf);
  free(p->kb);
  free(p->y0);
  free(p);
}

void SUNAMD_MassActionRhsHost(const SUNAMD_MassAction* p, const double* y, double* f,
                              int64_t c0, int64_t c1)
{
  const int m = p->m, R = p->R;
  for (int64_t c = c0; c < c1; c++)
  {
    const double* yc = y + c * m;
    double* fc       = f + c * m;
    for (int s = 0; s < m; s++) { fc[s] = 0.0; }
    for (int r = 0; r < R; r++)
    {
      const int* ix = p->idx + 4 * r;
      double wf     = (p->kf[c * R + r] * yc[ix[0]]) * yc[ix[1]];
      double wb     = (p->kb[c * R + r] * yc[ix[2]]) * yc[ix[3]];
      double w      = wf - wb;
      fc[ix[0]]     = fc[ix[0]] - w;
      fc[ix[1]]     = fc[ix[1]] - w;
      fc[ix[2]]     = fc[ix[2]] + w;
      fc[ix[3]]     = fc[ix[3]] + w;
    }
  }
}

void SUNAMD_MassActionJacHost(const SUNAMD_MassAction* p, const double* y, double* J,
                              int64_t c0, int64_t c1)
{
  const int m = p->m, R = p->R;
  for (int64_t c = c0; c < c1; c++)
  {
    const double* yc = y + c * m;
    double* Jc       = J + c * (int64_t)m * m;
    for (int e = 0; e < m * m; e++) { Jc[e] = 0.0; }
    for (int r = 0; r < R; r++)
    {
      const int* ix = p->idx + 4 * r;
      double kf = p->kf[c * R + r], kb = p->kb[c * R + r];
      double dw[4] = {kf * yc[ix[1]], kf * yc[ix[0]], -(kb * yc[ix[3]]), -(kb * yc[ix[2]])};
      for (int t = 0; t < 4; t++)
      {
        int row = ix[t];
        for (int q = 0; q < 4; q++)
        {
          double* e = Jc + (int64_t)ix[q] * m + row;
          if (t < 2) { *e = *e - dw[q]; }
          else { *e = *e + dw[q]; }
        }
      }
    }
  }
}
