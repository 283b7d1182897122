/* -----------------------------------------------------------------------------
 * Private KINSOL solver memory shared by the fixed-point / Anderson file
 * (kinsol_fp.c) and the Newton / line-search / linear-solver-interface file
 * (kinsol_newton.c).  Field meanings follow the reference's KINMemRec
 * (src/kinsol/kinsol_impl.h:60-330) and KINLsMemRec (kinsol_ls_impl.h:40-110);
 * the layout is this library's own and is never seen by callers.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_KINSOL_IMPL_H
#define SUNAMD_KINSOL_IMPL_H

#include "sunamd/kinsol_sunamd.h"
#include "sunamd/sundials_iterative.h"

#define ZERO 0.0
#define HALF 0.5
#define ONE  1.0
#define TWO  2.0

#define MXITER_DEFAULT     200
#define MSBSET_DEFAULT     10
#define MSBSET_SUB_DEFAULT 5
#define MXNBCF_DEFAULT     10
#define OMEGA_MIN_DEFAULT  0.00001
#define OMEGA_MAX_DEFAULT  0.9
#define ONETHIRD           0.3333333333333333
#define TWOTHIRDS          0.6666666666666666
#define UROUND             2.220446049250313e-16

/* internal flags of the iteration (kinsol.c:124-127) */
#define CONTINUE_ITERATIONS (-999)
#define RETRY_ITERATION     (-998)
#define STEP_TOO_SMALL      (-997)
#define CONSTR_VIOLATED     (-996)

typedef struct KinMem KinMem;

/* the attached linear solver (kinsol_ls_impl.h:40-110, direct solvers) */
typedef struct
{
  SUNLinearSolver LS;
  SUNMatrix J;
  KINLsJacFn jac;
  void* J_data;
  int jacDQ;
  long int nje, nfeDQ;
  long int last_flag;
} KinLsMem;

struct KinMem
{
  SUNContext ctx;
  KINSysFn func;
  void* user_data;
  int malloc_done;
  /* options shared by every strategy */
  long int mxiter;
  sunrealtype fnormtol;
  int ret_newest;
  /* fixed-point / Anderson options */
  sunrealtype beta, beta_aa;
  int damping, damping_aa;
  long int m_aa, delay_aa;
  int orth_aa;
  KINDampingFn damping_fn;
  KINDepthFn depth_fn;
  /* Newton options */
  sunrealtype scsteptol, sqrt_relfunc, mxnstepin;
  long int msbset, msbset_sub, mxnbcf;
  int no_init_setup, no_res_mon;
  sunrealtype omega, omega_min, omega_max;
  N_Vector constraints; /* owned copy, NULL when not set */
  /* per-solve state */
  int strategy;
  N_Vector uu, uscale, fscale;
  N_Vector unew, fval, pp, vtemp1, vtemp2, vtemp3;
  long int nni, nfe;
  sunrealtype fnorm, f1norm;
  /* Newton state */
  sunrealtype mxnewtstep, stepl, stepmul, sFdotJp, sJpnorm;
  sunrealtype fnorm_sub, sthrsh;
  long int nnilset, nnilset_sub, nbcf, nbktrk, ncscmx;
  int jac_current, update_fnorm_sub, retry_nni, eval_omega;
  KinLsMem* ls;
  /* fused scaled norms of the MI355X vector (KINSetUseFusedKernels) */
  int usefused;
  sunrealtype (*fused_ssnorm)(N_Vector v, N_Vector u, N_Vector uscale);
  sunrealtype (*fused_sfnorm)(N_Vector v, N_Vector fscale);
  /* Anderson acceleration workspace */
  long int m_aa_alloc, orth_alloc, current_depth;
  sunrealtype *gamma_aa, *R_aa, *T_aa, *cv;
  N_Vector *q_aa, *df_aa, *dg_aa, *Xv;
  N_Vector fold_aa, gold_aa;
  struct _SUNQRData qr;
  SUNQRAddFn qr_func;
};

/* kinsol.c:2038-2042 */
sunrealtype kin_sc_fnorm(KinMem* k, N_Vector v, N_Vector scale);
/* Newton iteration with or without the line search (kinsol_newton.c) */
int kin_newton(KinMem* k);
void kin_ls_free(KinMem* k);

#endif
