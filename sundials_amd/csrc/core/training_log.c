/* -----------------------------------------------------------------------------
 * Versioned binary file of (predictor, converged iterate) training pairs for
 * the learned Newton initial iterate (no counterpart in the reference).
 *
 * Layout (little-endian, every field naturally aligned):
 *   offset  0  char[8]   magic "SUNAMDTP"
 *           8  uint32    format version (1)
 *          12  uint32    header bytes (64)
 *          16  uint64    rows
 *          24  uint32    inputs per row (144: SUNAMD_MLP_IN_PAD)
 *          28  uint32    targets per row (64: SUNAMD_MLP_OUT)
 *          32  uint64    seed of the batch the pairs were logged from
 *          40  double    rtol of the logging runs
 *          48  double    atol of the logging runs
 *          56  uint32    log stride (every stride-th step of every system)
 *          60  uint32    reserved (0)
 *   then per row: uint32 m (species of the system, 1..targets), float32
 *   x[inputs], float32 t[targets]; the species mask is t's first m entries.
 * ---------------------------------------------------------------------------*/
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sunamd/learned_predictor.h"

static const char TLOG_MAGIC[8] = {'S', 'U', 'N', 'A', 'M', 'D', 'T', 'P'};
#define TLOG_VERSION      1u
#define TLOG_HEADER_BYTES 64u

typedef struct
{
  char magic[8];
  uint32_t version, header_bytes;
  uint64_t rows;
  uint32_t ninputs, ntargets;
  uint64_t seed;
  double rtol, atol;
  uint32_t stride, reserved;
} TLogHeader;

_Static_assert(sizeof(TLogHeader) == TLOG_HEADER_BYTES, "training-log header layout");

static int host_is_little_endian(void)
{
  const uint32_t one = 1;
  unsigned char b;
  memcpy(&b, &one, 1);
  return b == 1;
}

int SUNAMD_TrainingLogWrite(const char* path, const float* x, const float* t,
                            const float* mask, long rows, const SUNAMD_TrainingLogInfo* info)
{
  if (!path || rows < 0 || (rows > 0 && (!x || !t || !mask)) || !host_is_little_endian())
  {
    return -1;
  }
  TLogHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, TLOG_MAGIC, 8);
  h.version      = TLOG_VERSION;
  h.header_bytes = TLOG_HEADER_BYTES;
  h.rows         = (uint64_t)rows;
  h.ninputs      = SUNAMD_MLP_IN_PAD;
  h.ntargets     = SUNAMD_MLP_OUT;
  if (info)
  {
    h.seed   = info->seed;
    h.rtol   = info->rtol;
    h.atol   = info->atol;
    h.stride = (uint32_t)info->stride;
  }
  FILE* f = fopen(path, "wb");
  if (!f) { return -2; }
  int rc = (fwrite(&h, sizeof(h), 1, f) == 1) ? 0 : -3;
  for (long r = 0; r < rows && rc == 0; r++)
  {
    const float* mr = mask + (size_t)r * SUNAMD_MLP_OUT;
    uint32_t m      = 0;
    while (m < SUNAMD_MLP_OUT && mr[m] != 0.0f) { m++; }
    /* the mask must be a prefix of ones: the first m species of a system */
    for (uint32_t j = m; j < SUNAMD_MLP_OUT; j++)
    {
      if (mr[j] != 0.0f) { rc = -4; }
    }
    if (m == 0) { rc = -4; }
    if (rc) { break; }
    if (fwrite(&m, sizeof(m), 1, f) != 1 ||
        fwrite(x + (size_t)r * SUNAMD_MLP_IN_PAD, sizeof(float), SUNAMD_MLP_IN_PAD, f) !=
          SUNAMD_MLP_IN_PAD ||
        fwrite(t + (size_t)r * SUNAMD_MLP_OUT, sizeof(float), SUNAMD_MLP_OUT, f) !=
          SUNAMD_MLP_OUT)
    {
      rc = -3;
    }
  }
  if (fclose(f) != 0 && rc == 0) { rc = -3; }
  return rc;
}

static int read_header(FILE* f, TLogHeader* h)
{
  if (fread(h, sizeof(*h), 1, f) != 1) { return -3; }
  if (memcmp(h->magic, TLOG_MAGIC, 8) != 0) { return -5; }
  if (h->version != TLOG_VERSION || h->header_bytes != TLOG_HEADER_BYTES) { return -6; }
  if (h->ninputs != SUNAMD_MLP_IN_PAD || h->ntargets != SUNAMD_MLP_OUT) { return -6; }
  return 0;
}

int SUNAMD_TrainingLogInfoRead(const char* path, SUNAMD_TrainingLogInfo* info)
{
  if (!path || !info || !host_is_little_endian()) { return -1; }
  FILE* f = fopen(path, "rb");
  if (!f) { return -2; }
  TLogHeader h;
  int rc = read_header(f, &h);
  fclose(f);
  if (rc) { return rc; }
  info->version = (int)h.version;
  info->rows    = (long)h.rows;
  info->seed    = h.seed;
  info->rtol    = h.rtol;
  info->atol    = h.atol;
  info->stride  = (int)h.stride;
  return 0;
}

int SUNAMD_TrainingLogRead(const char* path, long row0, long nrows, float* x, float* t,
                           float* mask)
{
  if (!path || row0 < 0 || nrows < 0 || !host_is_little_endian()) { return -1; }
  if (nrows > 0 && (!x || !t || !mask)) { return -1; }
  FILE* f = fopen(path, "rb");
  if (!f) { return -2; }
  TLogHeader h;
  int rc = read_header(f, &h);
  if (rc == 0 && (uint64_t)row0 + (uint64_t)nrows > h.rows) { rc = -7; }
  const long row_bytes = (long)(sizeof(uint32_t) + sizeof(float) * (SUNAMD_MLP_IN_PAD + SUNAMD_MLP_OUT));
  if (rc == 0 && fseek(f, (long)TLOG_HEADER_BYTES + row0 * row_bytes, SEEK_SET) != 0) { rc = -3; }
  for (long r = 0; r < nrows && rc == 0; r++)
  {
    uint32_t m = 0;
    float* xr  = x + (size_t)r * SUNAMD_MLP_IN_PAD;
    float* tr  = t + (size_t)r * SUNAMD_MLP_OUT;
    float* mr  = mask + (size_t)r * SUNAMD_MLP_OUT;
    if (fread(&m, sizeof(m), 1, f) != 1 ||
        fread(xr, sizeof(float), SUNAMD_MLP_IN_PAD, f) != SUNAMD_MLP_IN_PAD ||
        fread(tr, sizeof(float), SUNAMD_MLP_OUT, f) != SUNAMD_MLP_OUT)
    {
      rc = -3;
      break;
    }
    if (m < 1 || m > SUNAMD_MLP_OUT)
    {
      rc = -8;
      break;
    }
    for (uint32_t j = 0; j < SUNAMD_MLP_OUT; j++) { mr[j] = (j < m) ? 1.0f : 0.0f; }
  }
  fclose(f);
  return rc;
}
