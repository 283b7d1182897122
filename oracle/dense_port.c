/* -----------------------------------------------------------------------------
 * CPU ORACLE (test infrastructure): dense LU, dense SUNMatrix, dense direct
 * linear solver, and their block-diagonal forms.
 *
 *   SUNDlsMat_denseGETRF  src/sundials/sundials_dense.c:85-151
 *     partial pivoting (first maximum |a_ik|, strict '>'), full-row swaps,
 *     multipliers = a_ik * (1/a_kk) stored below the diagonal, column-wise
 *     rank-1 update skipped where a_kj == 0; returns k+1 on a zero pivot.
 *   SUNDlsMat_denseGETRS  src/sundials/sundials_dense.c:153-186
 *   dense matrix ops      src/sunmatrix/dense/sunmatrix_dense.c:229-334
 *     (ScaleAddI: a_ij *= c, then += 1 on the diagonal)
 *   SUNLinSol_Dense       src/sunlinsol/dense/sunlinsol_dense.c:128-183
 *     (setup = GETRF -> SUNLS_LUFACT_FAIL on zero pivot; solve = copy b to x,
 *      GETRS in place)
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ZERO 0.0
#define ONE  1.0

sunindextype SUNDlsMat_denseGETRF(sunrealtype** a, sunindextype m, sunindextype n,
                                  sunindextype* p)
{
  for (sunindextype k = 0; k < n; k++)
  {
    sunrealtype* col_k = a[k];
    sunindextype l     = k;
    for (sunindextype i = k + 1; i < m; i++)
    {
      if (fabs(col_k[i]) > fabs(col_k[l])) { l = i; }
    }
    p[k] = l;
    if (col_k[l] == ZERO) { return k + 1; }
    if (l != k)
    {
      for (sunindextype i = 0; i < n; i++)
      {
        sunrealtype temp = a[i][l];
        a[i][l]          = a[i][k];
        a[i][k]          = temp;
      }
    }
    sunrealtype mult = ONE / col_k[k];
    for (sunindextype i = k + 1; i < m; i++) { col_k[i] *= mult; }
    for (sunindextype j = k + 1; j < n; j++)
    {
      sunrealtype* col_j = a[j];
      sunrealtype a_kj   = col_j[k];
      if (a_kj != ZERO)
      {
        for (sunindextype i = k + 1; i < m; i++) { col_j[i] -= a_kj * col_k[i]; }
      }
    }
  }
  return 0;
}

void SUNDlsMat_denseGETRS(sunrealtype** a, sunindextype n, sunindextype* p,
                          sunrealtype* b)
{
  for (sunindextype k = 0; k < n; k++)
  {
    sunindextype pk = p[k];
    if (pk != k)
    {
      sunrealtype tmp = b[k];
      b[k]            = b[pk];
      b[pk]           = tmp;
    }
  }
  for (sunindextype k = 0; k < n - 1; k++)
  {
    sunrealtype* col_k = a[k];
    for (sunindextype i = k + 1; i < n; i++) { b[i] -= col_k[i] * b[k]; }
  }
  for (sunindextype k = n - 1; k > 0; k--)
  {
    sunrealtype* col_k = a[k];
    b[k] /= col_k[k];
    for (sunindextype i = 0; i < k; i++) { b[i] -= col_k[i] * b[k]; }
  }
  b[0] /= a[0][0];
}

/* single column-major block with leading dimension m: same arithmetic */
sunindextype oracle_block_getrf(sunrealtype* a, sunindextype m, sunindextype* p)
{
  sunrealtype* cols[256];
  sunrealtype** c = cols;
  sunrealtype** heap = NULL;
  if (m > 256)
  {
    heap = (sunrealtype**)malloc((size_t)m * sizeof(sunrealtype*));
    c    = heap;
  }
  for (sunindextype j = 0; j < m; j++) { c[j] = a + j * m; }
  sunindextype r = SUNDlsMat_denseGETRF(c, m, m, p);
  free(heap);
  return r;
}

void oracle_block_getrs(const sunrealtype* a, sunindextype m, const sunindextype* p,
                        sunrealtype* b)
{
  sunrealtype* cols[256];
  sunrealtype** c = cols;
  sunrealtype** heap = NULL;
  if (m > 256)
  {
    heap = (sunrealtype**)malloc((size_t)m * sizeof(sunrealtype*));
    c    = heap;
  }
  for (sunindextype j = 0; j < m; j++) { c[j] = (sunrealtype*)a + j * m; }
  SUNDlsMat_denseGETRS(c, m, (sunindextype*)p, b);
  free(heap);
}

/* ------------------------------------------------------------------------- */
/* dense SUNMatrix                                                           */
/* ------------------------------------------------------------------------- */
typedef struct
{
  sunindextype M, N;
  sunrealtype* data;
  sunindextype ldata;
  sunrealtype** cols;
} DenseContent;

#define DC(A) ((DenseContent*)((A)->content))

static SUNMatrix_ID dense_getid(SUNMatrix A)
{
  (void)A;
  return SUNMATRIX_DENSE;
}

static void dense_destroy(SUNMatrix A)
{
  if (!A) { return; }
  if (A->content)
  {
    free(DC(A)->data);
    free(DC(A)->cols);
    free(A->content);
    A->content = NULL;
  }
  SUNMatFreeEmpty(A);
}

static SUNMatrix dense_clone(SUNMatrix A)
{
  return SUNDenseMatrix(DC(A)->M, DC(A)->N, A->sunctx);
}

static SUNErrCode dense_zero(SUNMatrix A)
{
  for (sunindextype i = 0; i < DC(A)->ldata; i++) { DC(A)->data[i] = ZERO; }
  return SUN_SUCCESS;
}

static SUNErrCode dense_copy(SUNMatrix A, SUNMatrix B)
{
  if (DC(A)->M != DC(B)->M || DC(A)->N != DC(B)->N) { return SUN_ERR_ARG_DIMSMISMATCH; }
  for (sunindextype j = 0; j < DC(A)->N; j++)
  {
    for (sunindextype i = 0; i < DC(A)->M; i++) { DC(B)->cols[j][i] = DC(A)->cols[j][i]; }
  }
  return SUN_SUCCESS;
}

static SUNErrCode dense_scaleaddi(sunrealtype c, SUNMatrix A)
{
  for (sunindextype j = 0; j < DC(A)->N; j++)
  {
    for (sunindextype i = 0; i < DC(A)->M; i++)
    {
      DC(A)->cols[j][i] *= c;
      if (i == j) { DC(A)->cols[j][i] += ONE; }
    }
  }
  return SUN_SUCCESS;
}

static SUNErrCode dense_scaleadd(sunrealtype c, SUNMatrix A, SUNMatrix B)
{
  if (DC(A)->M != DC(B)->M || DC(A)->N != DC(B)->N) { return SUN_ERR_ARG_DIMSMISMATCH; }
  for (sunindextype j = 0; j < DC(A)->N; j++)
  {
    for (sunindextype i = 0; i < DC(A)->M; i++)
    {
      DC(A)->cols[j][i] = c * DC(A)->cols[j][i] + DC(B)->cols[j][i];
    }
  }
  return SUN_SUCCESS;
}

static SUNErrCode dense_matvec(SUNMatrix A, N_Vector x, N_Vector y)
{
  sunrealtype* xd = N_VGetArrayPointer(x);
  sunrealtype* yd = N_VGetArrayPointer(y);
  if (!xd || !yd || xd == yd) { return SUN_ERR_MEM_FAIL; }
  for (sunindextype i = 0; i < DC(A)->M; i++) { yd[i] = ZERO; }
  for (sunindextype j = 0; j < DC(A)->N; j++)
  {
    sunrealtype* col_j = DC(A)->cols[j];
    for (sunindextype i = 0; i < DC(A)->M; i++) { yd[i] += col_j[i] * xd[j]; }
  }
  return SUN_SUCCESS;
}

SUNMatrix SUNDenseMatrix(sunindextype M, sunindextype N, SUNContext sunctx)
{
  if (M <= 0 || N <= 0) { return NULL; }
  SUNMatrix A = SUNMatNewEmpty(sunctx);
  if (!A) { return NULL; }
  A->ops->getid     = dense_getid;
  A->ops->clone     = dense_clone;
  A->ops->destroy   = dense_destroy;
  A->ops->zero      = dense_zero;
  A->ops->copy      = dense_copy;
  A->ops->scaleadd  = dense_scaleadd;
  A->ops->scaleaddi = dense_scaleaddi;
  A->ops->matvec    = dense_matvec;
  DenseContent* c   = (DenseContent*)calloc(1, sizeof(DenseContent));
  A->content        = c;
  c->M              = M;
  c->N              = N;
  c->ldata          = M * N;
  c->data           = (sunrealtype*)calloc((size_t)(M * N), sizeof(sunrealtype));
  c->cols           = (sunrealtype**)malloc((size_t)N * sizeof(sunrealtype*));
  if (!c->data || !c->cols)
  {
    dense_destroy(A);
    return NULL;
  }
  for (sunindextype j = 0; j < N; j++) { c->cols[j] = c->data + j * M; }
  return A;
}

sunrealtype* SUNDenseMatrix_Data(SUNMatrix A) { return DC(A)->data; }
sunrealtype** SUNDenseMatrix_Cols(SUNMatrix A) { return DC(A)->cols; }
sunindextype SUNDenseMatrix_Rows(SUNMatrix A) { return DC(A)->M; }
sunindextype SUNDenseMatrix_Columns(SUNMatrix A) { return DC(A)->N; }

/* ------------------------------------------------------------------------- */
/* dense direct linear solver                                                */
/* ------------------------------------------------------------------------- */
typedef struct
{
  sunindextype N;
  sunindextype* pivots;
  sunindextype last_flag;
  sunindextype nblocks; /* 0 = full dense; >0 = block-dense */
  sunindextype m;
} DenseLSContent;

#define LC(S) ((DenseLSContent*)((S)->content))

static SUNLinearSolver_Type ls_gettype(SUNLinearSolver S)
{
  (void)S;
  return SUNLINEARSOLVER_DIRECT;
}
static SUNLinearSolver_ID ls_getid(SUNLinearSolver S)
{
  (void)S;
  return SUNLINEARSOLVER_DENSE;
}
static SUNErrCode ls_initialize(SUNLinearSolver S)
{
  LC(S)->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}
static int ls_setup(SUNLinearSolver S, SUNMatrix A)
{
  if (SUNMatGetID(A) != SUNMATRIX_DENSE) { return SUN_ERR_ARG_WRONGTYPE; }
  LC(S)->last_flag = SUNDlsMat_denseGETRF(DC(A)->cols, DC(A)->M, DC(A)->N, LC(S)->pivots);
  if (LC(S)->last_flag > 0) { return SUNLS_LUFACT_FAIL; }
  return SUN_SUCCESS;
}
static int ls_solve(SUNLinearSolver S, SUNMatrix A, N_Vector x, N_Vector b,
                    sunrealtype tol)
{
  (void)tol;
  N_VScale(ONE, b, x);
  sunrealtype* xd = N_VGetArrayPointer(x);
  if (!xd) { return SUN_ERR_ARG_CORRUPT; }
  SUNDlsMat_denseGETRS(DC(A)->cols, DC(A)->M, LC(S)->pivots, xd);
  LC(S)->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}
static sunindextype ls_lastflag(SUNLinearSolver S) { return LC(S)->last_flag; }
static SUNErrCode ls_free(SUNLinearSolver S)
{
  if (!S) { return SUN_SUCCESS; }
  if (S->content)
  {
    free(LC(S)->pivots);
    free(S->content);
    S->content = NULL;
  }
  SUNLinSolFreeEmpty(S);
  return SUN_SUCCESS;
}

SUNLinearSolver SUNLinSol_Dense(N_Vector y, SUNMatrix A, SUNContext sunctx)
{
  if (SUNMatGetID(A) != SUNMATRIX_DENSE) { return NULL; }
  if (DC(A)->M != DC(A)->N || N_VGetLength(y) != DC(A)->M) { return NULL; }
  SUNLinearSolver S = SUNLinSolNewEmpty(sunctx);
  if (!S) { return NULL; }
  S->ops->gettype    = ls_gettype;
  S->ops->getid      = ls_getid;
  S->ops->initialize = ls_initialize;
  S->ops->setup      = ls_setup;
  S->ops->solve      = ls_solve;
  S->ops->lastflag   = ls_lastflag;
  S->ops->free       = ls_free;
  DenseLSContent* c  = (DenseLSContent*)calloc(1, sizeof(DenseLSContent));
  S->content         = c;
  c->N               = DC(A)->M;
  c->pivots          = (sunindextype*)calloc((size_t)c->N, sizeof(sunindextype));
  return S;
}

/* ------------------------------------------------------------------------- */
/* block-diagonal dense matrix + per-block solver                            */
/* ------------------------------------------------------------------------- */
typedef struct
{
  sunindextype nblocks, m;
  sunrealtype* data; /* nblocks * m * m, block b column j row i at b*m*m + j*m + i */
} BlockContent;

#define BC(A) ((BlockContent*)((A)->content))

static SUNMatrix_ID blk_getid(SUNMatrix A)
{
  (void)A;
  return SUNMATRIX_CUSTOM;
}
static void blk_destroy(SUNMatrix A)
{
  if (!A) { return; }
  if (A->content)
  {
    free(BC(A)->data);
    free(A->content);
    A->content = NULL;
  }
  SUNMatFreeEmpty(A);
}
static SUNMatrix blk_clone(SUNMatrix A)
{
  return SUNBlockDenseMatrix_Oracle(BC(A)->nblocks, BC(A)->m, A->sunctx);
}
static SUNErrCode blk_zero(SUNMatrix A)
{
  sunindextype n = BC(A)->nblocks * BC(A)->m * BC(A)->m;
  for (sunindextype i = 0; i < n; i++) { BC(A)->data[i] = ZERO; }
  return SUN_SUCCESS;
}
static SUNErrCode blk_copy(SUNMatrix A, SUNMatrix B)
{
  sunindextype n = BC(A)->nblocks * BC(A)->m * BC(A)->m;
  for (sunindextype i = 0; i < n; i++) { BC(B)->data[i] = BC(A)->data[i]; }
  return SUN_SUCCESS;
}
static SUNErrCode blk_scaleaddi(sunrealtype c, SUNMatrix A)
{
  sunindextype m = BC(A)->m;
  for (sunindextype b = 0; b < BC(A)->nblocks; b++)
  {
    sunrealtype* a = BC(A)->data + b * m * m;
    for (sunindextype j = 0; j < m; j++)
    {
      for (sunindextype i = 0; i < m; i++)
      {
        a[j * m + i] *= c;
        if (i == j) { a[j * m + i] += ONE; }
      }
    }
  }
  return SUN_SUCCESS;
}
static SUNErrCode blk_scaleadd(sunrealtype c, SUNMatrix A, SUNMatrix B)
{
  sunindextype n = BC(A)->nblocks * BC(A)->m * BC(A)->m;
  for (sunindextype i = 0; i < n; i++) { BC(A)->data[i] = c * BC(A)->data[i] + BC(B)->data[i]; }
  return SUN_SUCCESS;
}
static SUNErrCode blk_matvec(SUNMatrix A, N_Vector x, N_Vector y)
{
  sunrealtype* xd = N_VGetArrayPointer(x);
  sunrealtype* yd = N_VGetArrayPointer(y);
  sunindextype m  = BC(A)->m;
  for (sunindextype b = 0; b < BC(A)->nblocks; b++)
  {
    sunrealtype* a = BC(A)->data + b * m * m;
    for (sunindextype i = 0; i < m; i++) { yd[b * m + i] = ZERO; }
    for (sunindextype j = 0; j < m; j++)
    {
      for (sunindextype i = 0; i < m; i++) { yd[b * m + i] += a[j * m + i] * xd[b * m + j]; }
    }
  }
  return SUN_SUCCESS;
}

SUNMatrix SUNBlockDenseMatrix_Oracle(sunindextype nblocks, sunindextype m,
                                     SUNContext sunctx)
{
  if (nblocks <= 0 || m <= 0) { return NULL; }
  SUNMatrix A = SUNMatNewEmpty(sunctx);
  if (!A) { return NULL; }
  A->ops->getid     = blk_getid;
  A->ops->clone     = blk_clone;
  A->ops->destroy   = blk_destroy;
  A->ops->zero      = blk_zero;
  A->ops->copy      = blk_copy;
  A->ops->scaleadd  = blk_scaleadd;
  A->ops->scaleaddi = blk_scaleaddi;
  A->ops->matvec    = blk_matvec;
  BlockContent* c   = (BlockContent*)calloc(1, sizeof(BlockContent));
  A->content        = c;
  c->nblocks        = nblocks;
  c->m              = m;
  c->data           = (sunrealtype*)calloc((size_t)(nblocks * m * m), sizeof(sunrealtype));
  if (!c->data)
  {
    blk_destroy(A);
    return NULL;
  }
  return A;
}

sunrealtype* SUNBlockDenseMatrix_Oracle_Data(SUNMatrix A) { return BC(A)->data; }

static int blk_ls_setup(SUNLinearSolver S, SUNMatrix A)
{
  sunindextype m = BC(A)->m;
  for (sunindextype b = 0; b < BC(A)->nblocks; b++)
  {
    sunindextype r = oracle_block_getrf(BC(A)->data + b * m * m, m, LC(S)->pivots + b * m);
    if (r > 0)
    {
      /* first zero pivot, as a global column index + 1 */
      LC(S)->last_flag = b * m + r;
      return SUNLS_LUFACT_FAIL;
    }
  }
  LC(S)->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}

static int blk_ls_solve(SUNLinearSolver S, SUNMatrix A, N_Vector x, N_Vector b,
                        sunrealtype tol)
{
  (void)tol;
  N_VScale(ONE, b, x);
  sunrealtype* xd = N_VGetArrayPointer(x);
  sunindextype m  = BC(A)->m;
  for (sunindextype k = 0; k < BC(A)->nblocks; k++)
  {
    oracle_block_getrs(BC(A)->data + k * m * m, m, LC(S)->pivots + k * m, xd + k * m);
  }
  LC(S)->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}

SUNLinearSolver SUNLinSol_BlockDense_Oracle(N_Vector y, SUNMatrix A, SUNContext sunctx)
{
  if (N_VGetLength(y) != BC(A)->nblocks * BC(A)->m) { return NULL; }
  SUNLinearSolver S = SUNLinSolNewEmpty(sunctx);
  if (!S) { return NULL; }
  S->ops->gettype    = ls_gettype;
  S->ops->getid      = ls_getid;
  S->ops->initialize = ls_initialize;
  S->ops->setup      = blk_ls_setup;
  S->ops->solve      = blk_ls_solve;
  S->ops->lastflag   = ls_lastflag;
  S->ops->free       = ls_free;
  DenseLSContent* c  = (DenseLSContent*)calloc(1, sizeof(DenseLSContent));
  S->content         = c;
  c->N               = N_VGetLength(y);
  c->nblocks         = BC(A)->nblocks;
  c->m               = BC(A)->m;
  c->pivots          = (sunindextype*)calloc((size_t)c->N, sizeof(sunindextype));
  return S;
}
