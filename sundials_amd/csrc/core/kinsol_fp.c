/* -----------------------------------------------------------------------------
 * KINSOL fixed-point iteration with Anderson acceleration (strategy KIN_FP).
 * Restatement of the reference's control flow for that strategy:
 *
 *   KINCreate defaults                 kinsol.c:230-320
 *   KINInit, KINAllocVectors           kinsol.c:334-440, 847-930
 *   KINSol (KIN_FP branch)             kinsol.c:445-560
 *   KINFP                              kinsol.c:2445-2595
 *   AndersonAccQRDelete                kinsol.c:2601-2673
 *   AndersonAcc                        kinsol.c:2675-2949
 *   KINScFNorm                         kinsol.c:2038-2042
 *   KINInitAA                          kinsol_aa.c:20-130
 *   KINInitOrth                        kinsol_orth.c:20-110
 *   KINSetDamping / MAA / DelayAA / OrthAA / DampingAA / DampingFn /
 *   DepthFn / FuncNormTol / NumMaxIters / ReturnNewest
 *                                      kinsol_io.c:80-330
 * The QR updates of the acceleration space are SUNQRAdd_MGS / _ICWY / _CGS2 /
 * _DCGS2 (core/sundials_iterative.c).  Every vector operation goes through
 * the N_Vector ops in the reference's order, so on a device vector the
 * iteration's vector work runs on the device.
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kinsol_impl.h"
#include "sunamd_pow.h"

void* KINCreate(SUNContext sunctx)
{
  if (!sunctx) { return NULL; }
  KinMem* k = (KinMem*)calloc(1, sizeof(KinMem));
  if (!k) { return NULL; }
  k->ctx      = sunctx;
  k->mxiter   = MXITER_DEFAULT;
  k->fnormtol = sunamd_pow(UROUND, ONETHIRD);
  k->beta     = ONE;
  k->beta_aa  = ONE;
  k->orth_aa  = KIN_ORTH_MGS;
  /* Newton defaults (kinsol.c:262-290) */
  k->scsteptol    = sunamd_pow(UROUND, TWOTHIRDS);
  k->sqrt_relfunc = sqrt(UROUND);
  k->msbset       = MSBSET_DEFAULT;
  k->msbset_sub   = MSBSET_SUB_DEFAULT;
  k->mxnbcf       = MXNBCF_DEFAULT;
  k->omega_min    = OMEGA_MIN_DEFAULT;
  k->omega_max    = OMEGA_MAX_DEFAULT;
  return k;
}

int KINInit(void* kinmem, KINSysFn func, N_Vector tmpl)
{
  KinMem* k = (KinMem*)kinmem;
  if (!k) { return KIN_MEM_NULL; }
  if (!func || !tmpl) { return KIN_ILL_INPUT; }
  const struct _generic_N_Vector_Ops* o = tmpl->ops;
  if (!o->nvclone || !o->nvdestroy || !o->nvlinearsum || !o->nvprod || !o->nvscale ||
      !o->nvmaxnorm || !o->nvdotprod)
  {
    return KIN_ILL_INPUT;
  }
  k->unew   = N_VClone(tmpl);
  k->fval   = N_VClone(tmpl);
  k->vtemp1 = N_VClone(tmpl);
  k->vtemp2 = N_VClone(tmpl);
  k->pp     = N_VClone(tmpl);
  if (!k->unew || !k->fval || !k->vtemp1 || !k->vtemp2 || !k->pp) { return KIN_MEM_FAIL; }
  k->func        = func;
  k->malloc_done = 1;
  return KIN_SUCCESS;
}

#define KMEM()                          \
  KinMem* k = (KinMem*)kinmem;          \
  if (!k) { return KIN_MEM_NULL; }

int KINSetUserData(void* kinmem, void* user_data)
{
  KMEM();
  k->user_data = user_data;
  return KIN_SUCCESS;
}

int KINSetMAA(void* kinmem, long int maa)
{
  KMEM();
  if (maa < 0) { return KIN_ILL_INPUT; }
  k->m_aa = maa; /* capped below mxiter when the workspace is set up */
  return KIN_SUCCESS;
}

int KINSetDelayAA(void* kinmem, long int delay)
{
  KMEM();
  if (delay < 0) { return KIN_ILL_INPUT; }
  k->delay_aa = delay;
  return KIN_SUCCESS;
}

int KINSetOrthAA(void* kinmem, int orthaa)
{
  KMEM();
  if (orthaa < KIN_ORTH_MGS || orthaa > KIN_ORTH_DCGS2) { return KIN_ILL_INPUT; }
  k->orth_aa = orthaa;
  return KIN_SUCCESS;
}

/* beta < 1 enables damping, beta >= 1 disables it */
int KINSetDampingAA(void* kinmem, sunrealtype beta)
{
  KMEM();
  if (beta <= ZERO) { return KIN_ILL_INPUT; }
  k->damping_aa = beta < ONE;
  k->beta_aa    = beta < ONE ? beta : ONE;
  return KIN_SUCCESS;
}

int KINSetDamping(void* kinmem, sunrealtype beta)
{
  KMEM();
  if (beta <= ZERO) { return KIN_ILL_INPUT; }
  k->damping = beta < ONE;
  k->beta    = beta < ONE ? beta : ONE;
  return KIN_SUCCESS;
}

int KINSetDampingFn(void* kinmem, KINDampingFn damping_fn)
{
  KMEM();
  k->damping_fn = damping_fn;
  return KIN_SUCCESS;
}

int KINSetDepthFn(void* kinmem, KINDepthFn depth_fn)
{
  KMEM();
  k->depth_fn = depth_fn;
  return KIN_SUCCESS;
}

/* fnormtol == 0 restores the default uround^(1/3) */
int KINSetFuncNormTol(void* kinmem, sunrealtype fnormtol)
{
  KMEM();
  if (fnormtol < ZERO) { return KIN_ILL_INPUT; }
  k->fnormtol = (fnormtol == ZERO) ? sunamd_pow(UROUND, ONETHIRD) : fnormtol;
  return KIN_SUCCESS;
}

/* mxiter == 0 restores the default */
int KINSetNumMaxIters(void* kinmem, long int mxiter)
{
  KMEM();
  if (mxiter < 0) { return KIN_ILL_INPUT; }
  k->mxiter = (mxiter == 0) ? MXITER_DEFAULT : mxiter;
  return KIN_SUCCESS;
}

int KINSetReturnNewest(void* kinmem, sunbooleantype ret_newest)
{
  KMEM();
  k->ret_newest = ret_newest ? 1 : 0;
  return KIN_SUCCESS;
}

int KINGetNumNonlinSolvIters(void* kinmem, long int* nniters)
{
  KMEM();
  *nniters = k->nni;
  return KIN_SUCCESS;
}

int KINGetNumFuncEvals(void* kinmem, long int* nfevals)
{
  KMEM();
  *nfevals = k->nfe;
  return KIN_SUCCESS;
}

int KINGetFuncNorm(void* kinmem, sunrealtype* fnorm)
{
  KMEM();
  *fnorm = k->fnorm;
  return KIN_SUCCESS;
}

/* ---- Anderson-acceleration workspace ----------------------------------------- */
static void free_aa(KinMem* k)
{
  free(k->gamma_aa);
  free(k->R_aa);
  free(k->cv);
  free(k->Xv);
  k->gamma_aa = k->R_aa = k->cv = NULL;
  k->Xv                         = NULL;
  if (k->q_aa) { N_VDestroyVectorArray(k->q_aa, (int)k->m_aa_alloc); }
  if (k->df_aa) { N_VDestroyVectorArray(k->df_aa, (int)k->m_aa_alloc); }
  if (k->dg_aa) { N_VDestroyVectorArray(k->dg_aa, (int)k->m_aa_alloc); }
  k->q_aa = k->df_aa = k->dg_aa = NULL;
  if (k->fold_aa) { N_VDestroy(k->fold_aa); }
  if (k->gold_aa) { N_VDestroy(k->gold_aa); }
  k->fold_aa = k->gold_aa = NULL;
  k->m_aa_alloc           = 0;
}

static void free_orth(KinMem* k)
{
  if (k->vtemp3) { N_VDestroy(k->vtemp3); }
  k->vtemp3 = NULL;
  free(k->T_aa);
  k->T_aa       = NULL;
  k->orth_alloc = 0;
}

static int init_aa(KinMem* k)
{
  if (k->m_aa >= k->mxiter) { k->m_aa = k->mxiter - 1; }
  k->current_depth = 0;
  if (k->m_aa > k->m_aa_alloc)
  {
    free_aa(k);
    const long int m = k->m_aa;
    k->m_aa_alloc    = m;
    k->gamma_aa      = (sunrealtype*)malloc(sizeof(sunrealtype) * m);
    k->R_aa          = (sunrealtype*)malloc(sizeof(sunrealtype) * m * m);
    k->q_aa          = N_VCloneVectorArray((int)m, k->unew);
    k->df_aa         = N_VCloneVectorArray((int)m, k->unew);
    k->dg_aa         = N_VCloneVectorArray((int)m, k->unew);
    k->fold_aa       = N_VClone(k->unew);
    k->gold_aa       = N_VClone(k->unew);
    k->cv            = (sunrealtype*)malloc(sizeof(sunrealtype) * 2 * (m + 1));
    k->Xv            = (N_Vector*)malloc(sizeof(N_Vector) * 2 * (m + 1));
    if (!k->gamma_aa || !k->R_aa || !k->q_aa || !k->df_aa || !k->dg_aa || !k->fold_aa ||
        !k->gold_aa || !k->cv || !k->Xv)
    {
      free_aa(k);
      return KIN_MEM_FAIL;
    }
  }
  return KIN_SUCCESS;
}

static int init_orth(KinMem* k)
{
  /* (re)allocate also when an earlier solve used a method that needed less */
  if (k->m_aa > k->orth_alloc || (k->orth_aa != KIN_ORTH_MGS && !k->vtemp3) ||
      (k->orth_aa == KIN_ORTH_ICWY && !k->T_aa))
  {
    free_orth(k);
    k->orth_alloc = k->m_aa;
    if (k->orth_aa != KIN_ORTH_MGS)
    {
      k->vtemp3 = N_VClone(k->unew);
      if (!k->vtemp3) { return KIN_MEM_FAIL; }
    }
    if (k->orth_aa == KIN_ORTH_ICWY)
    {
      k->T_aa = (sunrealtype*)malloc(sizeof(sunrealtype) * k->m_aa * k->m_aa);
      if (!k->T_aa) { return KIN_MEM_FAIL; }
    }
  }
  /* no vector here offers the single-buffer (local + all-reduce) dot
     products, so the plain ICWY and DCGS2 updates are the ones used */
  k->qr.vtemp = k->vtemp2;
  switch (k->orth_aa)
  {
  case KIN_ORTH_MGS: k->qr_func = (SUNQRAddFn)SUNQRAdd_MGS; break;
  case KIN_ORTH_ICWY:
    k->qr_func       = (SUNQRAddFn)SUNQRAdd_ICWY;
    k->qr.vtemp2     = k->vtemp3;
    k->qr.temp_array = k->T_aa;
    break;
  case KIN_ORTH_CGS2:
    k->qr_func       = (SUNQRAddFn)SUNQRAdd_CGS2;
    k->qr.vtemp2     = k->vtemp3;
    k->qr.temp_array = k->cv;
    break;
  default:
    k->qr_func       = (SUNQRAddFn)SUNQRAdd_DCGS2;
    k->qr.vtemp2     = k->vtemp3;
    k->qr.temp_array = k->cv;
    break;
  }
  return KIN_SUCCESS;
}

/* max norm of f_scale * v */
sunrealtype kin_sc_fnorm(KinMem* k, N_Vector v, N_Vector scale)
{
  if (k->usefused)
  {
    const sunrealtype r = k->fused_sfnorm(v, scale);
    if (r >= ZERO) { return r; }
  }
  N_VProd(scale, v, k->vtemp1);
  return N_VMaxNorm(k->vtemp1);
}

/* the user damping function with the range check of the reference */
static int call_damping(KinMem* k, N_Vector u, N_Vector g, sunrealtype* qt_fn, long int depth,
                        sunrealtype* beta)
{
  if (k->damping_fn(k->nni, u, g, qt_fn, depth, k->user_data, beta)) { return KIN_DAMPING_FN_ERR; }
  if (*beta <= ZERO || *beta > ONE) { return KIN_DAMPING_FN_ERR; }
  return KIN_SUCCESS;
}

/* drop the oldest column of Q R (Givens rotations restore the triangle; the
   rotated Q columns are formed with LinearSums) */
static int qr_delete(KinMem* k, N_Vector* Q, sunrealtype* R, int depth)
{
  for (int i = 0; i < depth - 1; i++)
  {
    sunrealtype a = R[(i + 1) * depth + i];
    sunrealtype b = R[(i + 1) * depth + i + 1];
    const sunrealtype ss = a * a + b * b;
    sunrealtype t        = (ss <= ZERO) ? ZERO : sqrt(ss);
    const sunrealtype c  = a / t, s = b / t;
    R[(i + 1) * depth + i]     = t;
    R[(i + 1) * depth + i + 1] = ZERO;
    for (int j = i + 2; j < depth; j++)
    {
      a                    = R[j * depth + i];
      b                    = R[j * depth + i + 1];
      t                    = c * a + s * b;
      R[j * depth + i + 1] = -s * a + c * b;
      R[j * depth + i]     = t;
    }
    N_VLinearSum(c, Q[i], s, Q[i + 1], k->vtemp2);
    N_VLinearSum(-s, Q[i], c, Q[i + 1], Q[i + 1]);
    N_VScale(ONE, k->vtemp2, Q[i]);
  }
  for (int i = 1; i < depth; i++)
  {
    for (int j = 0; j < depth - 1; j++) { R[(i - 1) * depth + j] = R[i * depth + j]; }
  }
  if (k->orth_aa == KIN_ORTH_ICWY)
  {
    k->T_aa[0] = ONE;
    for (int i = 2; i < depth; i++)
    {
      SUNErrCode e = N_VDotProdMulti(i - 1, Q[i - 1], Q, k->T_aa + (i - 1) * depth);
      if (e) { return KIN_VECTOROP_ERR; }
      k->T_aa[(i - 1) * depth + (i - 1)] = ONE;
    }
  }
  return KIN_SUCCESS;
}

/* move the oldest difference columns to the end (they are overwritten next) */
static void rotate_history(KinMem* k, long int n)
{
  N_Vector g0 = k->dg_aa[0], f0 = k->df_aa[0];
  for (long int i = 1; i < n; i++)
  {
    k->dg_aa[i - 1] = k->dg_aa[i];
    k->df_aa[i - 1] = k->df_aa[i];
  }
  k->dg_aa[n - 1] = g0;
  k->df_aa[n - 1] = f0;
}

/* the first (or a depth-0) iterate: damped or plain fixed-point update */
static int fp_update(KinMem* k, N_Vector gval, N_Vector x, N_Vector xold)
{
  if (k->damping_aa || k->damping_fn)
  {
    if (k->damping_fn)
    {
      int r = call_damping(k, xold, gval, NULL, 0, &k->beta_aa);
      if (r) { return r; }
    }
    N_VLinearSum(ONE - k->beta_aa, xold, k->beta_aa, gval, x);
  }
  else { N_VScale(ONE, gval, x); }
  return KIN_SUCCESS;
}

/* x = the accelerated iterate from G(xold) = gval; fv receives F = gval - xold */
static int anderson(KinMem* k, N_Vector gval, N_Vector fv, N_Vector x, N_Vector xold,
                    long int iter)
{
  const long int mm = k->m_aa;
  sunrealtype* R    = k->R_aa;
  sunrealtype* gam  = k->gamma_aa;
  N_VLinearSum(ONE, gval, -ONE, xold, fv);
  if (iter > 0)
  {
    if (k->current_depth == mm)
    {
      rotate_history(k, mm);
      int r = qr_delete(k, k->q_aa, R, (int)mm);
      if (r) { return r; }
      k->current_depth--;
    }
    N_VLinearSum(ONE, gval, -ONE, k->gold_aa, k->dg_aa[k->current_depth]);
    N_VLinearSum(ONE, fv, -ONE, k->fold_aa, k->df_aa[k->current_depth]);
    k->current_depth++;
  }
  N_VScale(ONE, gval, k->gold_aa);
  N_VScale(ONE, fv, k->fold_aa);
  if (k->current_depth == 0) { return fp_update(k, gval, x, xold); }

  /* append the newest difference column to Q R */
  if (k->current_depth == 1)
  {
    sunrealtype d = N_VDotProd(k->df_aa[0], k->df_aa[0]);
    R[0]          = (d <= ZERO) ? ZERO : sqrt(d);
    N_VScale(ONE / R[0], k->df_aa[0], k->q_aa[0]);
  }
  else
  {
    k->qr_func(k->q_aa, R, k->df_aa[k->current_depth - 1], (int)k->current_depth - 1, (int)mm,
               &k->qr);
  }

  if (k->depth_fn)
  {
    long int nd = k->current_depth;
    if (k->depth_fn(k->nni, xold, gval, fv, k->df_aa, R, k->current_depth, k->user_data, &nd,
                    NULL))
    {
      return KIN_DEPTH_FN_ERR;
    }
    if (nd > k->current_depth) { nd = k->current_depth; }
    if (nd < 0) { nd = 0; }
    if (nd == 0)
    {
      k->current_depth = 0;
      return fp_update(k, gval, x, xold);
    }
    /* columns leave from the left, one at a time */
    while (k->current_depth > nd)
    {
      rotate_history(k, k->current_depth);
      int r = qr_delete(k, k->q_aa, R, (int)k->current_depth);
      if (r) { return r; }
      k->current_depth--;
    }
  }

  /* least squares: gamma = R^-1 Q^T fv, then x = gval - sum gamma_i dg_i
     (- (1 - beta)(fv - sum gamma_i df_i) when damped) in one combination */
  const long int l = k->current_depth;
  if (N_VDotProdMulti((int)l, fv, k->q_aa, gam) != SUN_SUCCESS) { return KIN_VECTOROP_ERR; }
  if (k->damping_fn)
  {
    int r = call_damping(k, xold, gval, gam, l, &k->beta_aa);
    if (r) { return r; }
  }
  sunrealtype* cv = k->cv;
  N_Vector* Xv    = k->Xv;
  int nvec        = 1;
  cv[0]           = ONE;
  Xv[0]           = gval;
  for (long int i = l - 1; i > -1; i--)
  {
    for (long int j = i + 1; j < l; j++) { gam[i] = gam[i] - R[j * mm + i] * gam[j]; }
    gam[i]     = gam[i] / R[i * mm + i];
    cv[nvec]   = -gam[i];
    Xv[nvec++] = k->dg_aa[i];
  }
  if (k->damping_aa || k->damping_fn)
  {
    const sunrealtype omb = ONE - k->beta_aa;
    cv[nvec]              = -omb;
    Xv[nvec++]            = fv;
    for (long int i = l - 1; i > -1; i--)
    {
      cv[nvec]   = omb * gam[i];
      Xv[nvec++] = k->df_aa[i];
    }
  }
  if (N_VLinearCombination(nvec, cv, Xv, x) != SUN_SUCCESS) { return KIN_VECTOROP_ERR; }
  return KIN_SUCCESS;
}

static int kin_fp(KinMem* k)
{
  N_Vector delta = k->vtemp1;
  int ret        = CONTINUE_ITERATIONS;
  sunrealtype tolfac;
  k->nni = 0;
  while (ret == CONTINUE_ITERATIONS)
  {
    k->nni++;
    int r = k->func(k->uu, k->fval, k->user_data);
    k->nfe++;
    if (r < 0)
    {
      ret = KIN_SYSFUNC_FAIL;
      break;
    }
    if (k->m_aa == 0 || k->nni - 1 < k->delay_aa)
    {
      if (k->damping || k->damping_fn)
      {
        if (k->damping_fn)
        {
          r = call_damping(k, k->uu, k->fval, NULL, 0, &k->beta);
          if (r)
          {
            ret = r;
            break;
          }
        }
        N_VLinearSum(ONE - k->beta, k->uu, k->beta, k->fval, k->unew);
        tolfac = k->beta;
      }
      else
      {
        N_VScale(ONE, k->fval, k->unew);
        tolfac = ONE;
      }
    }
    else
    {
      const long int iter_aa = (k->delay_aa > 0) ? k->nni - 1 - k->delay_aa : k->nni - 1;
      r = anderson(k, k->fval, delta, k->unew, k->uu, iter_aa);
      if (r)
      {
        ret = r;
        break;
      }
      tolfac = (iter_aa == 0 && (k->damping_aa || k->damping_fn)) ? k->beta_aa : ONE;
    }
    N_VLinearSum(ONE, k->unew, -ONE, k->uu, delta);
    k->fnorm = kin_sc_fnorm(k, delta, k->fscale);
    if (k->nni >= k->mxiter) { ret = KIN_MAXITER_REACHED; }
    if (k->fnorm <= tolfac * k->fnormtol) { ret = KIN_SUCCESS; }
    if (ret == CONTINUE_ITERATIONS || k->ret_newest) { N_VScale(ONE, k->unew, k->uu); }
  }
  return ret;
}

int KINSol(void* kinmem, N_Vector uu, int strategy, N_Vector u_scale, N_Vector f_scale)
{
  KMEM();
  if (!k->malloc_done) { return KIN_NO_MALLOC; }
  if (!uu) { return KIN_ILL_INPUT; }
  k->uu       = uu;
  k->uscale   = u_scale;
  k->fscale   = f_scale;
  k->strategy = strategy;
  if (strategy == KIN_NONE || strategy == KIN_LINESEARCH) { return kin_newton(k); }
  /* KIN_PICARD is not provided by this library */
  if (strategy != KIN_FP || !f_scale || k->constraints) { return KIN_ILL_INPUT; }
  if (k->m_aa != 0)
  {
    int r = init_aa(k);
    if (r) { return r; }
    r = init_orth(k);
    if (r) { return r; }
  }
  k->nfe = k->nni = 0;
  return kin_fp(k);
}

void KINFree(void** kinmem)
{
  if (!kinmem || !*kinmem) { return; }
  KinMem* k = (KinMem*)*kinmem;
  free_aa(k);
  free_orth(k);
  N_VDestroy(k->unew);
  N_VDestroy(k->fval);
  N_VDestroy(k->vtemp1);
  N_VDestroy(k->vtemp2);
  N_VDestroy(k->pp);
  if (k->constraints) { N_VDestroy(k->constraints); }
  kin_ls_free(k);
  free(k);
  *kinmem = NULL;
}
