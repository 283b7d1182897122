/* -----------------------------------------------------------------------------
 * Uneven batches of independent mass-action systems (size buckets).
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sunamd/batched_bdf.h"

SUNAMD_Batch* SUNAMD_BatchCreate(int64_t nsys, uint64_t seed, int nvalues, const int* m_values,
                                 const double* weights, double rfactor, double sigma)
{
  if (nsys <= 0 || nvalues <= 0 || nvalues > SUNAMD_BATCH_MAXB) { return NULL; }
  SUNAMD_Batch* B = (SUNAMD_Batch*)calloc(1, sizeof(SUNAMD_Batch));
  if (!B) { return NULL; }
  B->nbuckets   = nvalues;
  B->nsys       = nsys;
  B->seed       = seed;
  B->sys_bucket = (int*)malloc(sizeof(int) * (size_t)nsys);
  B->sys_cell   = (int64_t*)malloc(sizeof(int64_t) * (size_t)nsys);
  B->sys_yoff   = (int64_t*)malloc(sizeof(int64_t) * (size_t)nsys);
  if (!B->sys_bucket || !B->sys_cell || !B->sys_yoff)
  {
    SUNAMD_BatchFree(B);
    return NULL;
  }
  double wsum = 0.0;
  for (int b = 0; b < nvalues; b++) { wsum += weights[b]; }
  int64_t count[SUNAMD_BATCH_MAXB] = {0};
  int64_t yoff                     = 0;
  for (int64_t s = 0; s < nsys; s++)
  {
    double u   = SUNAMD_Uniform(seed, 5, (uint64_t)s) * wsum;
    int b      = 0;
    double acc = weights[0];
    while (u >= acc && b < nvalues - 1)
    {
      b++;
      acc += weights[b];
    }
    B->sys_bucket[s] = b;
    B->sys_cell[s]   = count[b]++;
    B->sys_yoff[s]   = yoff;
    yoff += m_values[b];
  }
  B->ytotal = yoff;
  for (int b = 0; b < nvalues; b++)
  {
    int m = m_values[b];
    int R = (int)floor(rfactor * m + 0.5);
    /* an empty bucket still gets a one-cell network so indices stay valid */
    int64_t nc   = count[b] > 0 ? count[b] : 1;
    B->bucket[b] = SUNAMD_MassActionCreate(m, R, nc, seed + 7919ULL * (uint64_t)(b + 1), sigma);
    if (!B->bucket[b])
    {
      SUNAMD_BatchFree(B);
      return NULL;
    }
    if (count[b] == 0) { B->bucket[b]->ncells = 0; }
  }
  return B;
}

void SUNAMD_BatchFree(SUNAMD_Batch* B)
{
  if (!B) { return; }
  for (int b = 0; b < B->nbuckets; b++) { SUNAMD_MassActionFree(B->bucket[b]); }
  free(B->sys_bucket);
  free(B->sys_cell);
  free(B->sys_yoff);
  free(B);
}
