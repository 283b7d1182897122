/* -----------------------------------------------------------------------------
 * ARKStep IMEX integrator: additive Runge-Kutta with an explicit and an
 * implicit (ESDIRK) table, adaptive steps, Newton + matrix-free iterative
 * linear solver.  Restatement of the reference's control flow for the
 * configuration this library supports (identity mass matrix, trivial
 * predictor, no relaxation / roots / constraints / forcing):
 *
 *   ARKodeEvolve main loop                arkode.c:657-1180
 *   arkInitialSetup, arkHin, UpperBoundH0,
 *   YddNorm, arkCompleteStep               arkode.c:1978-2270, 2485-2780
 *   arkStopTests (stop-time part)          arkode.c:2272-2480
 *   arkCheckConvergence,
 *   arkCheckTemporalError                  arkode.c:3207-3276, 3335-3400
 *   arkEwtSetSS                            arkode.c:2935-2960
 *   ARKodeSetDefaults                      arkode_io.c:49-120
 *   arkAdapt                               arkode_adapt.c:108-200
 *   I / PI / PID controllers (Soderlind)   sunadaptcontroller_soderlind.c:308-420, 566-700
 *   arkStep_SetDefaults                    arkode_arkstep_io.c:740-795
 *   arkStep_Init                           arkode_arkstep.c:945-1220
 *   arkStep_FullRHS                        arkode_arkstep.c:1297-1620
 *   arkStep_TakeStep_Z                     arkode_arkstep.c:1666-2240
 *   arkStep_Predict (trivial),
 *   arkStep_StageSetup,
 *   arkStep_ComputeSolutions               arkode_arkstep.c:2807-3335
 *   arkStep_Nls, NlsLSetup, NlsLSolve,
 *   NlsResidual_MassIdent, NlsConvTest     arkode_arkstep_nls.c:286-1200
 *   arkLsSetup, arkLsSolve, arkLsATimes,
 *   arkLsPSetup, arkLsPSolve               arkode_ls.c:2319-2470, 3166-3520
 * The ARK3(2)4L[2]SA coefficients are Kennedy & Carpenter's (Appl. Numer.
 * Math. 44, 2003), as tabulated in arkode_butcher_erk.def:218-246 and
 * arkode_butcher_dirk.def:248-276.
 * x^y is the correctly rounded sunamd_pow (SUNRpowerR).
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sunamd/arkode_sunamd.h"
#include "sunamd_pow.h"


#define ZERO SUN_RCONST(0.0)
#define HALF SUN_RCONST(0.5)
#define ONE  SUN_RCONST(1.0)
#define TWO  SUN_RCONST(2.0)
#define FOUR SUN_RCONST(4.0)

#define SUNMAX(a, b) ((a) > (b) ? (a) : (b))
#define SUNMIN(a, b) ((a) < (b) ? (a) : (b))
#define SUNRabs(x)   fabs(x)
#define SUNRsqrt(x)  ((x) <= 0.0 ? 0.0 : sqrt((x)))

/* arkode_impl.h */
#define MXSTEP_DEFAULT 500
#define MAXNEF         7
#define MAXNCF         10
#define MXHNIL         10
#define TINY           1.0e-10
#define FUZZ_FACTOR    100.0
#define H0_LBFACTOR    100.0
#define H0_UBFACTOR    0.1
#define H0_BIAS        HALF
#define H0_ITERS       4
#define ONEPSM         1.000001
#define ONEMSM         0.999999
/* arkode_adapt_impl.h */
#define SAFETY    0.9
#define GROWTH    20.0
#define HFIXED_LB 1.0
#define HFIXED_UB 1.0
#define ETAMX1    10000.0
#define ETAMXF    0.3
#define ETAMIN    0.1
#define ETACF     0.25
#define SMALL_NEF 2
/* arkode_arkstep_impl.h */
#define MAXCOR  3
#define CRDOWN  0.3
#define DGMAX   0.2
#define RDIV    2.3
#define MSBP    20
#define NLSCOEF 0.1
/* arkode_ls_impl.h */
#define ARKLS_MSBJ  51
#define ARKLS_EPLIN 0.05

/* step / convergence flags (arkode_impl.h) */
#define FIRST_CALL     100
#define PREV_CONV_FAIL 101
#define PREV_ERR_FAIL  102
#define PREDICT_AGAIN  103
#define TRY_AGAIN      104
#define CONV_FAIL      105
#define RHSFUNC_RECVR  106

#define ARK_NO_FAILURES 0
#define ARK_FAIL_BAD_J  1
#define ARK_FAIL_OTHER  2

#define ARK_FULLRHS_START 0
#define ARK_FULLRHS_END   1
#define ARK_FULLRHS_OTHER 2

/* largest stage count of the tables provided here (ARK4(3)7L[2]SA) */
#define NSTAGES 7

typedef struct
{
  int s, q, p;
  sunrealtype A[NSTAGES][NSTAGES], b[NSTAGES], d[NSTAGES], c[NSTAGES];
} ButcherTable;

/* Soderlind-family controller state (I, PI and PID are parameter subsets) */
typedef struct
{
  sunrealtype k1, k2, k3, k4, k5, bias, ep, epp, hp, hpp;
  int firststeps, historysize;
} Controller;

typedef struct ARKodeMemRec
{
  SUNContext sunctx;
  ARKRhsFn fe, fi;
  void* user_data;
  ARKStagePredictFn stage_predict; /* user stage predictor or NULL */
  /* tolerances */
  sunrealtype reltol, Sabstol;
  /* vectors */
  N_Vector yn, ycur, ewt, fn, tempv1, tempv2, tempv3, tempv4, tempv5;
  N_Vector Fe[NSTAGES], Fi[NSTAGES], sdata, zpred, zcor;
  /* tables */
  ButcherTable Be, Bi;
  int stages, q, p;
  /* time stepping */
  sunrealtype uround, tcur, tn, tretlast, h, hin, hprime, next_h, eta, hold, h0u, tstop;
  sunrealtype hmin, hmax_inv, tolsf;
  sunbooleantype tstopset, fixedstep, initsetup, firststage, fn_is_current,
    call_fullrhs;
  int interp_type, user_order;
  long int mxstep, nhnil;
  int maxnef, maxncf, mxhnil;
  /* adaptivity (ARKodeHAdaptMem) */
  sunrealtype etamax, etamx1, etamxf, etamin, etacf, safety, growth, lbound, ubound;
  int small_nef, pq, hp_order, hq_order, adjust;
  Controller C;
  int has_controller;
  /* counters */
  long int nst, nst_attempts, ncfn, netf, nfe, nfi, nsetups, nls_iters, nls_fails,
    nst_acc;
  /* ARKStep nonlinear-solve data */
  sunrealtype gamma, gammap, gamrat, crate, delp, eRNrm, nlscoef, crdown, rdiv, dgmax;
  int msbp, maxcor, istage, convfail;
  sunbooleantype linear, linear_timedep, autonomous;
  N_Vector fn_implicit; /* fi(tn, yn) reused by the first residual, or NULL */
  long int nstlp;
  sunbooleantype jcur;
  SUNNonlinearSolver NLS;
  sunbooleantype ownNLS;
  sunrealtype cvals[2 * NSTAGES + 2];
  N_Vector Xvecs[2 * NSTAGES + 2];
  /* linear solver interface (arkode_ls) */
  SUNLinearSolver LS;
  ARKLsPrecSetupFn pset;
  ARKLsPrecSolveFn psolve;
  ARKLsJacTimesSetupFn jtsetup;
  ARKLsJacTimesVecFn jtimes;
  sunrealtype ls_tcur, eplifac, nrmfac;
  N_Vector ls_ycur, ls_fcur, ytemp, x;
  sunbooleantype jbad;
  /* no matrix and no preconditioner setup: the Newton iteration never calls
     a linear-solver setup (arkLsInitialize, arkode_ls.c:3133-3139) */
  sunbooleantype lsetup_disabled;
  long int msbj, nstlj, nje, npe, nli, nps, ncfl, njtsetup, njtimes, nfeDQ;
  sunbooleantype jtimesDQ;
  int ls_last_flag;
} ARKodeMemRec;

typedef ARKodeMemRec* ARKodeMem;

/* ---------------------------------------------------------------------------
 * ARK3(2)4L[2]SA pair
 * -------------------------------------------------------------------------*/
static void ark324_tables(ButcherTable* Be, ButcherTable* Bi)
{
  memset(Be, 0, sizeof(*Be));
  memset(Bi, 0, sizeof(*Bi));
  Be->s = Bi->s = 4;
  Be->q = Bi->q = 3;
  Be->p = Bi->p = 2;
  const sunrealtype g = 0.4358665215084589994160194511935568425293;
  const sunrealtype b[4] = {0.1876410243467238251612921441668043913795,
                            -0.5952974735769549480478230275858851737782,
                            0.9717899277217721234705114322255239398694, g};
  const sunrealtype d[4] = {0.2147402862233891404862383406484193714659,
                            -0.4851622638849390928209050808398155895845,
                            0.86872500252038755116621237682951240796,
                            0.4016969751411624011684543633618838101586};
  const sunrealtype c[4] = {0.0, 0.8717330430169179988320389023871136850586, 0.6, 1.0};
  for (int j = 0; j < 4; j++)
  {
    Be->b[j] = Bi->b[j] = b[j];
    Be->d[j] = Bi->d[j] = d[j];
    Be->c[j] = Bi->c[j] = c[j];
  }
  Be->A[1][0] = 0.8717330430169179988320389023871136850586;
  Be->A[2][0] = 0.52758901197630041156180797140291790433;
  Be->A[2][1] = 0.07241098802369958843819202859708209566999;
  Be->A[3][0] = 0.3990960076760701320627260736092142797856;
  Be->A[3][1] = -0.437557654613519443722846363831022571942;
  Be->A[3][2] = 1.038461646937449311660120290221808292156;
  Bi->A[1][0] = g;
  Bi->A[1][1] = g;
  Bi->A[2][0] = 0.2576482460664272457999960162840797092643;
  Bi->A[2][1] = -0.09351476757488624521601546747763655179361;
  Bi->A[2][2] = g;
  Bi->A[3][0] = b[0];
  Bi->A[3][1] = b[1];
  Bi->A[3][2] = b[2];
  Bi->A[3][3] = g;
}

/* ---------------------------------------------------------------------------
 * ARK4(3)7L[2]SA pair, the reference's default IMEX method of order 4
 * (ARKODE_ARK437L2SA_ERK_7_3_4 / _DIRK_7_3_4, arkode_butcher_erk.def:371 and
 * arkode_butcher_dirk.def:474; Kennedy & Carpenter 2019).  The coefficients
 * are ratios of integers, each formed by one correctly rounded division.
 * -------------------------------------------------------------------------*/
#define R(p, q) ((sunrealtype)(p) / (sunrealtype)(q))
static void ark437_tables(ButcherTable* Be, ButcherTable* Bi)
{
  memset(Be, 0, sizeof(*Be));
  memset(Bi, 0, sizeof(*Bi));
  Be->s = Bi->s = 7;
  Be->q = Bi->q = 4;
  Be->p = Bi->p = 3;
  /* shared weights, embedding and abscissae (b_0 = b_1 = d_0 = d_1 = c_0 = 0) */
  const sunrealtype b[7] = {0.0, 0.0, R(9164257142617.0, 17756377923965.0),
                            R(-10812980402763.0, 74029279521829.0),
                            R(1335994250573.0, 5691609445217.0),
                            R(2273837961795.0, 8368240463276.0), R(247.0, 2000.0)};
  const sunrealtype d[7] = {0.0, 0.0, R(4469248916618.0, 8635866897933.0),
                            R(-621260224600.0, 4094290005349.0),
                            R(696572312987.0, 2942599194819.0),
                            R(1532940081127.0, 5565293938103.0), R(2441.0, 20000.0)};
  const sunrealtype c[7] = {0.0, R(247.0, 1000.0), R(4276536705230.0, 10142255878289.0),
                            R(67.0, 200.0), R(3.0, 40.0), R(7.0, 10.0), 1.0};
  for (int j = 0; j < 7; j++)
  {
    Be->b[j] = b[j];
    Be->d[j] = d[j];
    Be->c[j] = Bi->c[j] = c[j];
  }
  /* explicit table */
  Be->A[1][0] = R(247.0, 1000.0);
  Be->A[2][0] = R(247.0, 4000.0);
  Be->A[2][1] = R(2694949928731.0, 7487940209513.0);
  Be->A[3][0] = R(464650059369.0, 8764239774964.0);
  Be->A[3][1] = R(878889893998.0, 2444806327765.0);
  Be->A[3][2] = R(-952945855348.0, 12294611323341.0);
  Be->A[4][0] = R(476636172619.0, 8159180917465.0);
  Be->A[4][1] = R(-1271469283451.0, 7793814740893.0);
  Be->A[4][2] = R(-859560642026.0, 4356155882851.0);
  Be->A[4][3] = R(1723805262919.0, 4571918432560.0);
  Be->A[5][0] = R(6338158500785.0, 11769362343261.0);
  Be->A[5][1] = R(-4970555480458.0, 10924838743837.0);
  Be->A[5][2] = R(3326578051521.0, 2647936831840.0);
  Be->A[5][3] = R(-880713585975.0, 1841400956686.0);
  Be->A[5][4] = R(-1428733748635.0, 8843423958496.0);
  Be->A[6][0] = R(760814592956.0, 3276306540349.0);
  Be->A[6][1] = R(760814592956.0, 3276306540349.0);
  Be->A[6][2] = R(-47223648122716.0, 6934462133451.0);
  Be->A[6][3] = R(71187472546993.0, 9669769126921.0);
  Be->A[6][4] = R(-13330509492149.0, 9695768672337.0);
  Be->A[6][5] = R(11565764226357.0, 8513123442827.0);
  /* ESDIRK table: diagonal gamma = 0.1235, last row = b (stiffly accurate);
     its b_6 is gamma, not the explicit table's 247/2000 (the same number) */
  const sunrealtype g = R(1235.0, 10000.0);
  for (int i = 1; i < 7; i++) { Bi->A[i][i] = g; }
  Bi->A[1][0] = g;
  Bi->A[2][0] = Bi->A[2][1] = R(624185399699.0, 4186980696204.0);
  Bi->A[3][0] = Bi->A[3][1] = R(1258591069120.0, 10082082980243.0);
  Bi->A[3][2] = R(-322722984531.0, 8455138723562.0);
  Bi->A[4][0] = Bi->A[4][1] = R(-436103496990.0, 5971407786587.0);
  Bi->A[4][2] = R(-2689175662187.0, 11046760208243.0);
  Bi->A[4][3] = R(4431412449334.0, 12995360898505.0);
  Bi->A[5][0] = Bi->A[5][1] = R(-2207373168298.0, 14430576638973.0);
  Bi->A[5][2] = R(242511121179.0, 3358618340039.0);
  Bi->A[5][3] = R(3145666661981.0, 7780404714551.0);
  Bi->A[5][4] = R(5882073923981.0, 14490790706663.0);
  for (int j = 0; j < 7; j++)
  {
    Bi->b[j] = (j == 6) ? g : b[j];
    Bi->d[j] = d[j];
    if (j < 6) { Bi->A[6][j] = b[j]; }
  }
}
#undef R

/* the pair's coefficients for inspection (ARKodeButcherTable_LoadERK /
   LoadDIRK of ARKODE_ARK324L2SA_*_4_2_3): A (4x4 row-major), b, c, d */
/* tables for a method order: 3 gives ARK3(2)4L[2]SA, 4 (and the default,
   order <= 0) ARK4(3)7L[2]SA, as arkStep_SetButcherTables selects the
   ARKSTEP_DEFAULT_ARK_*TABLE_q pair (arkode_arkstep.c:2216-2262) */
static int load_tables(int order, ButcherTable* Be, ButcherTable* Bi)
{
  if (order <= 0 || order == 4) { ark437_tables(Be, Bi); }
  else if (order == 3) { ark324_tables(Be, Bi); }
  else { return -1; }
  return 0;
}

int SUNAMD_ARKTable(int order, int implicit_part, int* s, double* A, double* b, double* c,
                    double* d, int* q, int* p)
{
  ButcherTable Be, Bi;
  if (load_tables(order, &Be, &Bi) != 0) { return -1; }
  const ButcherTable* B = implicit_part ? &Bi : &Be;
  *s = B->s;
  for (int i = 0; i < B->s; i++)
  {
    for (int j = 0; j < B->s; j++) { A[B->s * i + j] = B->A[i][j]; }
    b[i] = B->b[i];
    c[i] = B->c[i];
    d[i] = B->d[i];
  }
  *q = B->q;
  *p = B->p;
  return 0;
}

int SUNAMD_ARK324Table(int implicit_part, double* A, double* b, double* c, double* d, int* q,
                       int* p)
{
  ButcherTable Be, Bi;
  ark324_tables(&Be, &Bi);
  const ButcherTable* B = implicit_part ? &Bi : &Be;
  for (int i = 0; i < 4; i++)
  {
    for (int j = 0; j < 4; j++) { A[4 * i + j] = B->A[i][j]; }
    b[i] = B->b[i];
    c[i] = B->c[i];
    d[i] = B->d[i];
  }
  *q = B->q;
  *p = B->p;
  return 0;
}

/* ARKodeButcherTable_IsStifflyAccurate: b_j == A_{s-1,j} and c_{s-1} == 1 */
static sunbooleantype table_sa(const ButcherTable* B)
{
  for (int j = 0; j < B->s; j++)
  {
    if (SUNRabs(B->b[j] - B->A[B->s - 1][j]) > 100.0 * SUN_UNIT_ROUNDOFF) { return SUNFALSE; }
  }
  return SUNTRUE;
}

/* ---------------------------------------------------------------------------
 * controllers (sunadaptcontroller_soderlind.c)
 * -------------------------------------------------------------------------*/
static void ctrl_set_params(Controller* C, sunrealtype k1, sunrealtype k2, sunrealtype k3,
                            sunrealtype k4, sunrealtype k5)
{
  C->k1 = k1;
  C->k2 = k2;
  C->k3 = k3;
  C->k4 = k4;
  C->k5 = k5;
  if (k5 != ZERO || k3 != ZERO) { C->historysize = 2; }
  else if (k4 != ZERO || k2 != ZERO) { C->historysize = 1; }
  else { C->historysize = 0; }
}

static void ctrl_reset(Controller* C)
{
  C->ep = C->epp = C->hp = C->hpp = ONE;
  C->firststeps = 0;
}

static int ctrl_create(Controller* C, const char* name)
{
  /* Soderlind defaults (H0321 parameters, bias 1), then the subset */
  C->bias = 1.0;
  ctrl_set_params(C, 1.25, 0.5, -0.75, 0.25, 0.75);
  ctrl_reset(C);
  if (strcmp(name, "I") == 0) { ctrl_set_params(C, 1.0, 0.0, 0.0, 0.0, 0.0); }
  else if (strcmp(name, "PI") == 0) { ctrl_set_params(C, 0.8, -0.31, 0.0, 0.0, 0.0); }
  else if (strcmp(name, "PID") == 0) { ctrl_set_params(C, 0.58, -0.21, 0.1, 0.0, 0.0); }
  else { return -1; }
  return 0;
}

static sunrealtype ctrl_estimate(const Controller* C, sunrealtype h, int p, sunrealtype dsm)
{
  const int ord        = p + 1;
  const sunrealtype e1 = C->bias * dsm;
  if (C->firststeps < C->historysize) { return h * sunamd_pow(e1, -ONE / ord); }
  const sunrealtype k1 = -C->k1 / ord;
  sunrealtype hnew     = h * sunamd_pow(e1, k1);
  if (C->historysize > 0)
  {
    const sunrealtype k2    = -C->k2 / ord;
    const sunrealtype hrat1 = h / C->hp;
    hnew *= sunamd_pow(C->ep, k2) * sunamd_pow(hrat1, C->k4);
    if (C->historysize > 1)
    {
      const sunrealtype k3    = -C->k3 / ord;
      const sunrealtype hrat2 = C->hp / C->hpp;
      hnew *= sunamd_pow(C->epp, k3) * sunamd_pow(hrat2, C->k5);
    }
  }
  return hnew;
}

static void ctrl_update(Controller* C, sunrealtype h, sunrealtype dsm)
{
  C->epp = C->ep;
  C->ep  = C->bias * dsm;
  C->hpp = C->hp;
  C->hp  = h;
  if (C->firststeps < C->historysize) { C->firststeps += 1; }
}

/* ---------------------------------------------------------------------------
 * creation and options
 * -------------------------------------------------------------------------*/
static int set_defaults(ARKodeMem ark)
{
  ark->fixedstep = SUNFALSE;
  ark->reltol    = 1.e-4;
  ark->Sabstol   = 1.e-9;
  ark->mxstep    = MXSTEP_DEFAULT;
  ark->mxhnil    = MXHNIL;
  ark->maxnef    = MAXNEF;
  ark->maxncf    = MAXNCF;
  ark->hin       = ZERO;
  ark->hmin      = ZERO;
  ark->hmax_inv  = ZERO;
  ark->tstopset  = SUNFALSE;
  ark->tstop     = ZERO;
  ark->etamx1    = ETAMX1;
  ark->etamxf    = ETAMXF;
  ark->etamin    = ETAMIN;
  ark->small_nef = SMALL_NEF;
  ark->etacf     = ETACF;
  ark->safety    = SAFETY;
  ark->growth    = GROWTH;
  ark->lbound    = HFIXED_LB;
  ark->ubound    = HFIXED_UB;
  ark->pq        = 0;
  ark->hp_order  = 0;
  ark->hq_order  = 0;
  ark->adjust    = 0;
  /* arkStep_SetDefaults */
  ark->user_order = 0;
  ark->maxcor     = MAXCOR;
  ark->nlscoef    = NLSCOEF;
  ark->crdown     = CRDOWN;
  ark->rdiv       = RDIV;
  ark->dgmax      = DGMAX;
  ark->msbp       = MSBP;
  ark->stages     = 0;
  ark->istage     = 0;
  ark->jcur       = SUNFALSE;
  ark->convfail   = ARK_NO_FAILURES;
  ark->linear         = SUNFALSE;
  ark->linear_timedep = SUNTRUE;
  ark->autonomous     = SUNFALSE;
  /* default controller (arkReplaceAdaptController with NULL): I */
  ctrl_create(&ark->C, "I");
  ark->has_controller = 1;
  return ARK_SUCCESS;
}

void* ARKStepCreate(ARKRhsFn fe, ARKRhsFn fi, sunrealtype t0, N_Vector y0,
                    SUNContext sunctx)
{
  if (fe == NULL || fi == NULL || y0 == NULL) { return NULL; }
  ARKodeMem ark = (ARKodeMem)calloc(1, sizeof(ARKodeMemRec));
  if (ark == NULL) { return NULL; }
  ark->sunctx      = sunctx;
  ark->fe          = fe;
  ark->fi          = fi;
  ark->uround      = SUN_UNIT_ROUNDOFF;
  ark->interp_type = ARK_INTERP_HERMITE;
  set_defaults(ark);
  N_Vector* vs[] = {&ark->yn,     &ark->ycur,   &ark->ewt,    &ark->fn,   &ark->tempv1,
                    &ark->tempv2, &ark->tempv3, &ark->tempv4, &ark->tempv5, &ark->sdata,
                    &ark->zpred,  &ark->zcor};
  for (size_t k = 0; k < sizeof(vs) / sizeof(vs[0]); k++)
  {
    *vs[k] = N_VClone(y0);
    if (*vs[k] == NULL)
    {
      ARKodeFree((void**)&ark);
      return NULL;
    }
  }
  for (int j = 0; j < NSTAGES; j++)
  {
    ark->Fe[j] = N_VClone(y0);
    ark->Fi[j] = N_VClone(y0);
  }
  /* arkInit (FIRST_INIT) */
  N_VScale(ONE, y0, ark->yn);
  ark->tcur = ark->tn = ark->tretlast = t0;
  ark->hold = ark->hprime = ark->next_h = ark->h = ZERO;
  ark->h0u = ZERO;
  ark->eta = ONE;
  ark->initsetup = ark->firststage = SUNTRUE;
  ark->fn_is_current = SUNFALSE;
  ark->call_fullrhs  = SUNFALSE;
  ark->etamax        = ark->etamx1;
  /* default nonlinear solver: Newton */
  ark->NLS = SUNNonlinSol_Newton(y0, sunctx);
  if (ark->NLS == NULL)
  {
    ARKodeFree((void**)&ark);
    return NULL;
  }
  ark->ownNLS = SUNTRUE;
  return ark;
}

void ARKodeFree(void** arkode_mem)
{
  if (arkode_mem == NULL || *arkode_mem == NULL) { return; }
  ARKodeMem ark = (ARKodeMem)(*arkode_mem);
  N_Vector vs[] = {ark->yn,     ark->ycur,   ark->ewt,    ark->fn,     ark->tempv1,
                   ark->tempv2, ark->tempv3, ark->tempv4, ark->tempv5, ark->sdata,
                   ark->zpred,  ark->zcor,   ark->ytemp,  ark->x};
  for (size_t k = 0; k < sizeof(vs) / sizeof(vs[0]); k++)
  {
    if (vs[k]) { N_VDestroy(vs[k]); }
  }
  for (int j = 0; j < NSTAGES; j++)
  {
    if (ark->Fe[j]) { N_VDestroy(ark->Fe[j]); }
    if (ark->Fi[j]) { N_VDestroy(ark->Fi[j]); }
  }
  if (ark->NLS && ark->ownNLS) { SUNNonlinSolFree(ark->NLS); }
  free(ark);
  *arkode_mem = NULL;
}

#define ARKMEM_OR_RETURN()                                    \
  if (arkode_mem == NULL) { return ARK_MEM_NULL; }            \
  ARKodeMem ark = (ARKodeMem)arkode_mem

int ARKodeSStolerances(void* arkode_mem, sunrealtype reltol, sunrealtype abstol)
{
  ARKMEM_OR_RETURN();
  if (reltol < ZERO || abstol < ZERO) { return ARK_ILL_INPUT; }
  ark->reltol  = reltol;
  ark->Sabstol = abstol;
  return ARK_SUCCESS;
}

int ARKodeSetUserData(void* arkode_mem, void* user_data)
{
  ARKMEM_OR_RETURN();
  ark->user_data = user_data;
  return ARK_SUCCESS;
}

int ARKodeSetOrder(void* arkode_mem, int ord)
{
  ARKMEM_OR_RETURN();
  /* the order-3 and order-4 ARK pairs are provided (order <= 0: default 4) */
  if (ord > 0 && ord != 3 && ord != 4) { return ARK_ILL_INPUT; }
  ark->user_order = ord;
  return ARK_SUCCESS;
}

int ARKodeSetStopTime(void* arkode_mem, sunrealtype tstop)
{
  ARKMEM_OR_RETURN();
  /* arkode_io.c: tstop must lie ahead of tcur in the direction of h */
  if (ark->nst > 0 && (tstop - ark->tcur) * ark->h < ZERO) { return ARK_ILL_INPUT; }
  ark->tstop    = tstop;
  ark->tstopset = SUNTRUE;
  return ARK_SUCCESS;
}

int ARKodeSetStagePredictFn(void* arkode_mem, ARKStagePredictFn PredictStage)
{
  ARKMEM_OR_RETURN();
  /* arkode_arkstep_io.c:1143-1160 */
  ark->stage_predict = PredictStage;
  return ARK_SUCCESS;
}

int ARKodeSetMaxNumSteps(void* arkode_mem, long int mxsteps)
{
  ARKMEM_OR_RETURN();
  ark->mxstep = (mxsteps == 0) ? MXSTEP_DEFAULT : mxsteps;
  return ARK_SUCCESS;
}

int ARKodeSetFixedStep(void* arkode_mem, sunrealtype hfixed)
{
  ARKMEM_OR_RETURN();
  if (hfixed != ZERO)
  {
    ark->fixedstep = SUNTRUE;
    ark->hin       = hfixed;
  }
  else { ark->fixedstep = SUNFALSE; }
  return ARK_SUCCESS;
}

int ARKodeSetInitStep(void* arkode_mem, sunrealtype hin)
{
  ARKMEM_OR_RETURN();
  ark->hin = hin;
  ark->h0u = ZERO;
  return ARK_SUCCESS;
}

int ARKodeSetInterpolantType(void* arkode_mem, int itype)
{
  ARKMEM_OR_RETURN();
  if (itype != ARK_INTERP_NONE && itype != ARK_INTERP_HERMITE) { return ARK_ILL_INPUT; }
  ark->interp_type = itype;
  return ARK_SUCCESS;
}

int ARKodeSetAdaptControllerByName(void* arkode_mem, const char* cname)
{
  ARKMEM_OR_RETURN();
  Controller C;
  if (ctrl_create(&C, cname) != 0) { return ARK_ILL_INPUT; }
  ark->C              = C;
  ark->has_controller = 1;
  return ARK_SUCCESS;
}

int ARKodeSetErrorBias(void* arkode_mem, sunrealtype bias)
{
  ARKMEM_OR_RETURN();
  ark->C.bias = (bias < ONE) ? 1.0 : bias;
  return ARK_SUCCESS;
}

int ARKodeSetMaxNonlinIters(void* arkode_mem, int maxcor)
{
  ARKMEM_OR_RETURN();
  ark->maxcor = (maxcor <= 0) ? MAXCOR : maxcor;
  return ARK_SUCCESS;
}

int ARKodeSetNonlinConvCoef(void* arkode_mem, sunrealtype nlscoef)
{
  ARKMEM_OR_RETURN();
  ark->nlscoef = (nlscoef <= ZERO) ? NLSCOEF : nlscoef;
  return ARK_SUCCESS;
}

int ARKodeSetLSetupFrequency(void* arkode_mem, int msbp)
{
  ARKMEM_OR_RETURN();
  ark->msbp = (msbp == 0) ? MSBP : msbp;
  return ARK_SUCCESS;
}

int ARKodeSetLinear(void* arkode_mem, int timedepend)
{
  ARKMEM_OR_RETURN();
  if (timedepend && ark->autonomous) { return ARK_ILL_INPUT; }
  ark->linear         = SUNTRUE;
  ark->linear_timedep = (timedepend == 1);
  ark->dgmax          = 100.0 * SUN_UNIT_ROUNDOFF;
  return ARK_SUCCESS;
}

int ARKodeSetNonlinear(void* arkode_mem)
{
  ARKMEM_OR_RETURN();
  ark->linear         = SUNFALSE;
  ark->linear_timedep = SUNTRUE;
  ark->dgmax          = DGMAX;
  return ARK_SUCCESS;
}

int ARKodeSetAutonomous(void* arkode_mem, sunbooleantype autonomous)
{
  ARKMEM_OR_RETURN();
  if (autonomous && ark->linear && ark->linear_timedep) { return ARK_ILL_INPUT; }
  ark->autonomous = autonomous;
  return ARK_SUCCESS;
}

int ARKodeSetNonlinearSolver(void* arkode_mem, SUNNonlinearSolver NLS)
{
  ARKMEM_OR_RETURN();
  if (NLS == NULL) { return ARK_ILL_INPUT; }
  if (ark->NLS && ark->ownNLS) { SUNNonlinSolFree(ark->NLS); }
  ark->NLS    = NLS;
  ark->ownNLS = SUNFALSE;
  return ARK_SUCCESS;
}

static int arkLsATimes(void* arkode_mem, N_Vector v, N_Vector z);
static int arkLsPSetup(void* arkode_mem);
static int arkLsPSolve(void* arkode_mem, N_Vector r, N_Vector z, sunrealtype tol, int lr);

int ARKodeSetLinearSolver(void* arkode_mem, SUNLinearSolver LS, SUNMatrix A)
{
  ARKMEM_OR_RETURN();
  if (LS == NULL) { return ARKLS_ILL_INPUT; }
  /* this library's interface is the matrix-free iterative one */
  if (A != NULL || SUNLinSolGetType(LS) != SUNLINEARSOLVER_ITERATIVE) { return ARKLS_ILL_INPUT; }
  ark->LS      = LS;
  ark->pset    = NULL;
  ark->psolve  = NULL;
  ark->jtsetup  = NULL;
  ark->jtimes   = NULL;
  ark->jtimesDQ = SUNTRUE;
  ark->msbj     = ARKLS_MSBJ;
  ark->eplifac = ARKLS_EPLIN;
  if (SUNLinSolSetATimes(LS, ark, arkLsATimes) != SUN_SUCCESS) { return ARKLS_ILL_INPUT; }
  if (SUNLinSolSetPreconditioner(LS, ark, NULL, NULL) != SUN_SUCCESS) { return ARKLS_ILL_INPUT; }
  if (ark->ytemp == NULL) { ark->ytemp = N_VClone(ark->yn); }
  if (ark->x == NULL) { ark->x = N_VClone(ark->yn); }
  if (ark->ytemp == NULL || ark->x == NULL) { return ARKLS_MEM_FAIL; }
  ark->nrmfac = SUNRsqrt((sunrealtype)N_VGetLength(ark->x));
  return ARKLS_SUCCESS;
}

int ARKodeSetPreconditioner(void* arkode_mem, ARKLsPrecSetupFn psetup,
                            ARKLsPrecSolveFn psolve)
{
  ARKMEM_OR_RETURN();
  if (ark->LS == NULL) { return ARKLS_LMEM_NULL; }
  ark->pset   = psetup;
  ark->psolve = psolve;
  if (SUNLinSolSetPreconditioner(ark->LS, ark, psetup ? arkLsPSetup : NULL,
                                 psolve ? arkLsPSolve : NULL) != SUN_SUCCESS)
  {
    return ARKLS_ILL_INPUT;
  }
  return ARKLS_SUCCESS;
}

int ARKodeSetJacTimes(void* arkode_mem, ARKLsJacTimesSetupFn jtsetup,
                      ARKLsJacTimesVecFn jtimes)
{
  ARKMEM_OR_RETURN();
  if (ark->LS == NULL) { return ARKLS_LMEM_NULL; }
  /* a NULL product selects the difference quotient (arkode_ls.c:929-944) */
  ark->jtimesDQ = (jtimes == NULL);
  ark->jtsetup  = (jtimes == NULL) ? NULL : jtsetup;
  ark->jtimes   = jtimes;
  return ARKLS_SUCCESS;
}

int ARKodeSetEpsLin(void* arkode_mem, sunrealtype eplifac)
{
  ARKMEM_OR_RETURN();
  ark->eplifac = (eplifac <= ZERO) ? ARKLS_EPLIN : eplifac;
  return ARKLS_SUCCESS;
}

int ARKodeSetJacEvalFrequency(void* arkode_mem, long int msbj)
{
  ARKMEM_OR_RETURN();
  ark->msbj = (msbj <= 0) ? ARKLS_MSBJ : msbj;
  return ARKLS_SUCCESS;
}

/* ---------------------------------------------------------------------------
 * error weights (arkEwtSetSS)
 * -------------------------------------------------------------------------*/
static int ewt_set(ARKodeMem ark, N_Vector ycur, N_Vector weight)
{
  N_VAbs(ycur, ark->tempv1);
  N_VScale(ark->reltol, ark->tempv1, ark->tempv1);
  N_VAddConst(ark->tempv1, ark->Sabstol, ark->tempv1);
  if (ark->Sabstol == ZERO)
  {
    if (N_VMin(ark->tempv1) <= ZERO) { return -1; }
  }
  N_VInv(ark->tempv1, weight);
  return 0;
}

/* ---------------------------------------------------------------------------
 * arkStep_FullRHS
 * -------------------------------------------------------------------------*/
static int full_rhs(ARKodeMem ark, sunrealtype t, N_Vector y, N_Vector f, int mode)
{
  int retval;
  const sunbooleantype sa = table_sa(&ark->Be) && table_sa(&ark->Bi);
  switch (mode)
  {
  case ARK_FULLRHS_START:
    if (!ark->fn_is_current)
    {
      retval = ark->fe(t, y, ark->Fe[0], ark->user_data);
      ark->nfe++;
      if (retval != 0) { return ARK_RHSFUNC_FAIL; }
      retval = ark->fi(t, y, ark->Fi[0], ark->user_data);
      ark->nfi++;
      if (retval != 0) { return ARK_RHSFUNC_FAIL; }
    }
    N_VLinearSum(ONE, ark->Fi[0], ONE, ark->Fe[0], f);
    break;
  case ARK_FULLRHS_END:
    if (!ark->fn_is_current)
    {
      if (!sa)
      {
        retval = ark->fe(t, y, ark->Fe[0], ark->user_data);
        ark->nfe++;
        if (retval != 0) { return ARK_RHSFUNC_FAIL; }
        retval = ark->fi(t, y, ark->Fi[0], ark->user_data);
        ark->nfi++;
        if (retval != 0) { return ARK_RHSFUNC_FAIL; }
      }
      else
      {
        N_VScale(ONE, ark->Fe[ark->stages - 1], ark->Fe[0]);
        N_VScale(ONE, ark->Fi[ark->stages - 1], ark->Fi[0]);
      }
    }
    N_VLinearSum(ONE, ark->Fi[0], ONE, ark->Fe[0], f);
    break;
  case ARK_FULLRHS_OTHER:
    retval = ark->fe(t, y, ark->tempv2, ark->user_data);
    ark->nfe++;
    if (retval != 0) { return ARK_RHSFUNC_FAIL; }
    retval = ark->fi(t, y, ark->sdata, ark->user_data);
    ark->nfi++;
    if (retval != 0) { return ARK_RHSFUNC_FAIL; }
    N_VLinearSum(ONE, ark->sdata, ONE, ark->tempv2, f);
    break;
  default: return ARK_RHSFUNC_FAIL;
  }
  return ARK_SUCCESS;
}

/* ---------------------------------------------------------------------------
 * nonlinear-solver callbacks (arkode_arkstep_nls.c)
 * -------------------------------------------------------------------------*/
static int ls_setup(ARKodeMem ark, int convfail, sunrealtype tpred, N_Vector ypred,
                    N_Vector fpred, sunbooleantype* jcurPtr);
static int ls_solve(ARKodeMem ark, N_Vector b, sunrealtype tnow, N_Vector ynow,
                    N_Vector fnow, sunrealtype eRNrm, int mnewt);

/* arkStep_NlsResidual_MassIdent and, with the trivial predictor on an
   autonomous problem, its _TrivialPredAutonomous variant (the first
   iteration reuses fi(tn, yn)) */
static int nls_residual(N_Vector zcor, N_Vector r, void* arkode_mem)
{
  ARKodeMem ark = (ARKodeMem)arkode_mem;
  N_VLinearSum(ONE, ark->zpred, ONE, zcor, ark->ycur);
  int nls_iter = -1;
  if (ark->autonomous)
  {
    if (SUNNonlinSolGetCurIter(ark->NLS, &nls_iter) != SUN_SUCCESS) { return ARK_NLS_OP_ERR; }
  }
  if (ark->autonomous && nls_iter == 0 && ark->fn_implicit)
  {
    N_VScale(ONE, ark->fn_implicit, ark->Fi[ark->istage]);
  }
  else
  {
    int retval = ark->fi(ark->tcur, ark->ycur, ark->Fi[ark->istage], ark->user_data);
    ark->nfi++;
    if (retval < 0) { return ARK_RHSFUNC_FAIL; }
    if (retval > 0) { return RHSFUNC_RECVR; }
  }
  sunrealtype c[3];
  N_Vector X[3];
  c[0] = ONE;
  X[0] = zcor;
  c[1] = -ONE;
  X[1] = ark->sdata;
  c[2] = -ark->gamma;
  X[2] = ark->Fi[ark->istage];
  if (N_VLinearCombination(3, c, X, r) != 0) { return ARK_VECTOROP_ERR; }
  return ARK_SUCCESS;
}

static int nls_lsetup(sunbooleantype jbad, sunbooleantype* jcur, void* arkode_mem)
{
  ARKodeMem ark = (ARKodeMem)arkode_mem;
  if (jbad) { ark->convfail = ARK_FAIL_BAD_J; }
  ark->nsetups++;
  int retval = ls_setup(ark, ark->convfail, ark->tcur, ark->ycur, ark->Fi[ark->istage],
                        &ark->jcur);
  *jcur           = ark->jcur;
  ark->firststage = SUNFALSE;
  ark->gamrat = ark->crate = ONE;
  ark->gammap = ark->gamma;
  ark->nstlp  = ark->nst;
  if (retval < 0) { return ARK_LSETUP_FAIL; }
  if (retval > 0) { return CONV_FAIL; }
  return ARK_SUCCESS;
}

static int nls_lsolve(N_Vector b, void* arkode_mem)
{
  ARKodeMem ark = (ARKodeMem)arkode_mem;
  int nonlin_iter;
  if (SUNNonlinSolGetCurIter(ark->NLS, &nonlin_iter) != SUN_SUCCESS) { return ARK_NLS_OP_ERR; }
  int retval = ls_solve(ark, b, ark->tcur, ark->ycur, ark->Fi[ark->istage], ark->eRNrm,
                        nonlin_iter);
  if (retval < 0) { return ARK_LSOLVE_FAIL; }
  if (retval > 0) { return CONV_FAIL; }
  return ARK_SUCCESS;
}

static int nls_convtest(SUNNonlinearSolver NLS, N_Vector y, N_Vector del, sunrealtype tol,
                        N_Vector ewt, void* arkode_mem)
{
  (void)y;
  ARKodeMem ark = (ARKodeMem)arkode_mem;
  if (ark->linear) { return SUN_SUCCESS; }
  int m;
  sunrealtype delnrm = N_VWrmsNorm(del, ewt);
  if (SUNNonlinSolGetCurIter(NLS, &m) != SUN_SUCCESS) { return ARK_MEM_NULL; }
  if (m > 0) { ark->crate = SUNMAX(ark->crdown * ark->crate, delnrm / ark->delp); }
  sunrealtype dcon = SUNMIN(ark->crate, ONE) * delnrm / tol;
  if (dcon <= ONE) { return SUN_SUCCESS; }
  if ((m >= 1) && (delnrm > ark->rdiv * ark->delp)) { return SUN_NLS_CONV_RECVR; }
  ark->delp = delnrm;
  return SUN_NLS_CONTINUE;
}

static int nls_init(ARKodeMem ark)
{
  ark->nls_iters = 0;
  ark->nls_fails = 0;
  if (SUNNonlinSolSetConvTestFn(ark->NLS, nls_convtest, ark) != SUN_SUCCESS) { return ARK_NLS_INIT_FAIL; }
  if (SUNNonlinSolSetMaxIters(ark->NLS, ark->maxcor) != SUN_SUCCESS) { return ARK_NLS_INIT_FAIL; }
  const sunbooleantype lsetup = ark->LS && !ark->lsetup_disabled;
  if (SUNNonlinSolSetLSetupFn(ark->NLS, lsetup ? nls_lsetup : NULL) != SUN_SUCCESS) { return ARK_NLS_INIT_FAIL; }
  if (SUNNonlinSolSetLSolveFn(ark->NLS, ark->LS ? nls_lsolve : NULL) != SUN_SUCCESS) { return ARK_NLS_INIT_FAIL; }
  if (SUNNonlinSolSetSysFn(ark->NLS, nls_residual) != SUN_SUCCESS) { return ARK_NLS_INIT_FAIL; }
  if (SUNNonlinSolInitialize(ark->NLS) != SUN_SUCCESS) { return ARK_NLS_INIT_FAIL; }
  return ARK_SUCCESS;
}

/* arkStep_Nls */
static int stage_nls(ARKodeMem ark, int nflag)
{
  sunbooleantype callLSetup;
  long int nls_iters_inc = 0, nls_fails_inc = 0;
  if (ark->LS && !ark->lsetup_disabled)
  {
    if (ark->linear)
    {
      ark->convfail = (nflag == FIRST_CALL) ? ARK_NO_FAILURES : ARK_FAIL_OTHER;
    }
    else
    {
      ark->convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? ARK_NO_FAILURES
                                                                          : ARK_FAIL_OTHER;
    }
    callLSetup = (ark->firststage) || (ark->msbp < 0) ||
                 (SUNRabs(ark->gamrat - ONE) > ark->dgmax);
    if (ark->linear) { callLSetup = callLSetup || (ark->linear_timedep); }
    else
    {
      callLSetup = callLSetup || (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) ||
                   (ark->nst >= ark->nstlp + labs((long)ark->msbp));
    }
  }
  else
  {
    ark->crate = ONE;
    callLSetup = SUNFALSE;
  }
  N_VConst(ZERO, ark->zcor);
  ark->eRNrm = 0.1 * ark->nlscoef;
  int retval = SUNNonlinSolSolve(ark->NLS, ark->zpred, ark->zcor, ark->ewt, ark->nlscoef,
                                 callLSetup, ark);
  (void)SUNNonlinSolGetNumIters(ark->NLS, &nls_iters_inc);
  ark->nls_iters += nls_iters_inc;
  (void)SUNNonlinSolGetNumConvFails(ark->NLS, &nls_fails_inc);
  ark->nls_fails += nls_fails_inc;
  if (retval == SUN_SUCCESS)
  {
    ark->jcur = SUNFALSE;
    N_VLinearSum(ONE, ark->zcor, ONE, ark->zpred, ark->ycur);
    return ARK_SUCCESS;
  }
  if (retval == SUN_NLS_CONV_RECVR) { return CONV_FAIL; }
  return retval;
}

/* ---------------------------------------------------------------------------
 * linear-solver interface (arkode_ls.c, matrix-free iterative case)
 * -------------------------------------------------------------------------*/
static int ls_setup(ARKodeMem ark, int convfail, sunrealtype tpred, N_Vector ypred,
                    N_Vector fpred, sunbooleantype* jcurPtr)
{
  ark->ls_tcur = tpred;
  ark->ls_ycur = ypred;
  ark->ls_fcur = fpred;
  const sunbooleantype dgamma_fail = (SUNRabs(ark->gamrat - ONE) >= ark->dgmax);
  ark->jbad = (ark->initsetup) || (ark->nst >= ark->nstlj + ark->msbj) ||
              ((convfail == ARK_FAIL_BAD_J) && (!dgamma_fail)) || (convfail == ARK_FAIL_OTHER);
  *jcurPtr          = ark->jbad;
  ark->ls_last_flag = SUNLinSolSetup(ark->LS, NULL);
  if (*jcurPtr)
  {
    ark->npe++;
    ark->nstlj = ark->nst;
  }
  if (ark->jbad) { *jcurPtr = SUNTRUE; }
  return ark->ls_last_flag;
}

static int ls_solve(ARKodeMem ark, N_Vector b, sunrealtype tnow, N_Vector ynow,
                    N_Vector fnow, sunrealtype eRNrm, int mnewt)
{
  ark->ls_tcur = tnow;
  ark->ls_ycur = ynow;
  ark->ls_fcur = fnow;
  const sunrealtype deltar = ark->eplifac * eRNrm;
  const sunrealtype bnorm  = N_VWrmsNorm(b, ark->ewt);
  if (bnorm <= deltar)
  {
    if (mnewt > 0) { N_VConst(ZERO, b); }
    ark->ls_last_flag = ARKLS_SUCCESS;
    return ark->ls_last_flag;
  }
  const sunrealtype delta = deltar * ark->nrmfac;
  if (SUNLinSolSetScalingVectors(ark->LS, ark->ewt, ark->ewt) != SUN_SUCCESS)
  {
    ark->ls_last_flag = -9;
    return ark->ls_last_flag;
  }
  N_VConst(ZERO, ark->x);
  if (SUNLinSolSetZeroGuess(ark->LS, SUNTRUE) != SUN_SUCCESS) { return -1; }
  if (ark->jtsetup)
  {
    ark->ls_last_flag = ark->jtsetup(tnow, ynow, fnow, ark->user_data);
    ark->njtsetup++;
    if (ark->ls_last_flag) { return ark->ls_last_flag; }
  }
  int retval = SUNLinSolSolve(ark->LS, NULL, ark->x, b, delta);
  N_VScale(ONE, ark->x, b);
  int nli_inc = SUNLinSolNumIters(ark->LS);
  ark->nli += nli_inc;
  if (retval != SUN_SUCCESS) { ark->ncfl++; }
  ark->ls_last_flag = retval;
  switch (retval)
  {
  case SUN_SUCCESS: return 0;
  case SUNLS_RES_REDUCED: return (mnewt == 0) ? 0 : 1;
  case SUNLS_CONV_FAIL:
  case SUNLS_ATIMES_FAIL_REC:
  case SUNLS_PSOLVE_FAIL_REC:
  case SUNLS_PACKAGE_FAIL_REC:
  case SUNLS_QRFACT_FAIL:
  case SUNLS_LUFACT_FAIL: return 1;
  case SUNLS_GS_FAIL:
  case SUNLS_QRSOL_FAIL:
  case SUNLS_ATIMES_FAIL_UNREC:
  case SUNLS_PSOLVE_FAIL_UNREC: return -1;
  default: return (retval < 0) ? -1 : 0;
  }
}

/* arkLsDQJtimes (arkode_ls.c:2842-2880): Jv ~ [fi(t, y + sig v) - fi(t, y)] / sig
   with sig = 1 / ||v||_WRMS, shrunk by 1/4 while fi fails recoverably */
#define MAX_DQITERS 3
static int dq_jtimes(ARKodeMem ark, N_Vector v, N_Vector Jv, sunrealtype t, N_Vector y,
                     N_Vector fy, N_Vector work)
{
  sunrealtype sig = ONE / N_VWrmsNorm(v, ark->ewt);
  int retval      = 0;
  for (int iter = 0; iter < MAX_DQITERS; iter++)
  {
    N_VLinearSum(sig, v, ONE, y, work);
    retval = ark->fi(t, work, Jv, ark->user_data);
    ark->nfeDQ++;
    if (retval == 0) { break; }
    if (retval < 0) { return -1; }
    sig *= 0.25;
  }
  if (retval > 0) { return +1; }
  const sunrealtype siginv = ONE / sig;
  N_VLinearSum(siginv, Jv, -siginv, fy, Jv);
  return 0;
}

static int arkLsATimes(void* arkode_mem, N_Vector v, N_Vector z)
{
  ARKodeMem ark = (ARKodeMem)arkode_mem;
  int retval    = ark->jtimesDQ
                    ? dq_jtimes(ark, v, z, ark->ls_tcur, ark->ls_ycur, ark->ls_fcur, ark->ytemp)
                    : ark->jtimes(v, z, ark->ls_tcur, ark->ls_ycur, ark->ls_fcur,
                                  ark->user_data, ark->ytemp);
  ark->njtimes++;
  if (retval != 0) { return retval; }
  N_VLinearSum(ONE, v, -ark->gamma, z, z);
  return 0;
}

static int arkLsPSetup(void* arkode_mem)
{
  ARKodeMem ark = (ARKodeMem)arkode_mem;
  return ark->pset(ark->ls_tcur, ark->ls_ycur, ark->ls_fcur, !(ark->jbad), &ark->jcur,
                   ark->gamma, ark->user_data);
}

static int arkLsPSolve(void* arkode_mem, N_Vector r, N_Vector z, sunrealtype tol, int lr)
{
  ARKodeMem ark = (ARKodeMem)arkode_mem;
  int retval = ark->psolve(ark->ls_tcur, ark->ls_ycur, ark->ls_fcur, r, z, ark->gamma, tol, lr,
                           ark->user_data);
  ark->nps++;
  return retval;
}

/* ---------------------------------------------------------------------------
 * stage predictor (trivial), stage setup and step solution
 * -------------------------------------------------------------------------*/
static int stage_setup(ARKodeMem ark, sunbooleantype implicit)
{
  const int i = ark->istage;
  int nvec    = 0;
  if (implicit)
  {
    ark->gamma = ark->h * ark->Bi.A[i][i];
    if (ark->firststage) { ark->gammap = ark->gamma; }
    ark->gamrat = (ark->firststage) ? ONE : ark->gamma / ark->gammap;
  }
  if (implicit)
  {
    N_VLinearSum(ONE, ark->yn, -ONE, ark->zpred, ark->sdata);
    ark->cvals[0] = ONE;
    ark->Xvecs[0] = ark->sdata;
    nvec          = 1;
  }
  for (int j = 0; j < i; j++)
  {
    ark->cvals[nvec] = ark->h * ark->Be.A[i][j];
    ark->Xvecs[nvec] = ark->Fe[j];
    nvec += 1;
  }
  for (int j = 0; j < i; j++)
  {
    ark->cvals[nvec] = ark->h * ark->Bi.A[i][j];
    ark->Xvecs[nvec] = ark->Fi[j];
    nvec += 1;
  }
  if (N_VLinearCombination(nvec, ark->cvals, ark->Xvecs, ark->sdata) != 0) { return ARK_VECTOROP_ERR; }
  return ARK_SUCCESS;
}

static int compute_solutions(ARKodeMem ark, sunrealtype* dsmPtr)
{
  N_Vector y    = ark->ycur;
  N_Vector yerr = ark->tempv1;
  int nvec;
  *dsmPtr = ZERO;
  const sunbooleantype sa = table_sa(&ark->Be) && table_sa(&ark->Bi);
  if (!sa)
  {
    ark->cvals[0] = ONE;
    ark->Xvecs[0] = ark->yn;
    nvec          = 1;
    for (int j = 0; j < ark->stages; j++)
    {
      ark->cvals[nvec] = ark->h * ark->Be.b[j];
      ark->Xvecs[nvec] = ark->Fe[j];
      nvec += 1;
      ark->cvals[nvec] = ark->h * ark->Bi.b[j];
      ark->Xvecs[nvec] = ark->Fi[j];
      nvec += 1;
    }
    if (N_VLinearCombination(nvec, ark->cvals, ark->Xvecs, y) != 0) { return ARK_VECTOROP_ERR; }
  }
  if (!ark->fixedstep)
  {
    nvec = 0;
    for (int j = 0; j < ark->stages; j++)
    {
      ark->cvals[nvec] = ark->h * (ark->Be.b[j] - ark->Be.d[j]);
      ark->Xvecs[nvec] = ark->Fe[j];
      nvec += 1;
      ark->cvals[nvec] = ark->h * (ark->Bi.b[j] - ark->Bi.d[j]);
      ark->Xvecs[nvec] = ark->Fi[j];
      nvec += 1;
    }
    if (N_VLinearCombination(nvec, ark->cvals, ark->Xvecs, yerr) != 0) { return ARK_VECTOROP_ERR; }
    *dsmPtr = N_VWrmsNorm(yerr, ark->ewt);
  }
  return ARK_SUCCESS;
}

/* arkStep_TakeStep_Z for the IMEX pair (explicit first stage) */
static int take_step(ARKodeMem ark, sunrealtype* dsmPtr, int* nflagPtr)
{
  int retval;
  /* NLS setup: the Newton solver has none */
  sunbooleantype implicit_stage = SUNFALSE;
  int is_start                  = 1;
  if (SUNRabs(ark->Bi.A[0][0]) > TINY)
  {
    implicit_stage = SUNTRUE;
    is_start       = 0;
  }
  const sunbooleantype save_fn_for_interp = (ark->interp_type == ARK_INTERP_HERMITE);
  /* trivial predictor (the only one here) on an autonomous problem: fi(tn,
     yn) is reused by each stage's first residual */
  const sunbooleantype save_fn_for_residual = ark->autonomous;
  const sunbooleantype eval_rhs = !implicit_stage || save_fn_for_interp || save_fn_for_residual;
  if (!ark->fn_is_current && eval_rhs)
  {
    /* an IMEX pair: the full RHS is evaluated unless the first stage is
       implicit and no interpolant needs it */
    if (!implicit_stage || save_fn_for_interp)
    {
      const int mode = (ark->initsetup) ? ARK_FULLRHS_START : ARK_FULLRHS_END;
      retval         = full_rhs(ark, ark->tn, ark->yn, ark->fn, mode);
      if (retval) { return ARK_RHSFUNC_FAIL; }
      ark->fn_is_current = SUNTRUE;
    }
    else
    {
      retval = ark->fi(ark->tn, ark->yn, ark->Fi[0], ark->user_data);
      ark->nfi++;
      if (retval < 0) { return ARK_RHSFUNC_FAIL; }
      if (retval > 0) { return ARK_UNREC_RHSFUNC_ERR; }
    }
  }
  ark->fn_implicit = NULL;
  if (save_fn_for_residual)
  {
    if (!implicit_stage) { ark->fn_implicit = ark->Fi[0]; }
    else
    {
      N_VScale(ONE, ark->Fi[0], ark->tempv5);
      ark->fn_implicit = ark->tempv5;
    }
  }
  for (int is = is_start; is < ark->stages; is++)
  {
    ark->istage    = is;
    implicit_stage = (SUNRabs(ark->Bi.A[is][is]) > TINY);
    ark->tcur      = ark->tn + ark->Bi.c[is] * ark->h;
    if (implicit_stage)
    {
      /* arkStep_Predict, trivial predictor: the step's initial state */
      N_VScale(ONE, ark->yn, ark->zpred);
      /* the user's stage predictor after the built-in one
         (arkode_arkstep.c:1943-1955): it may refine zpred; a positive
         return retries the step attempt (the caller's convergence-failure
         path with nflag unchanged), a negative one fails the step */
      if (ark->stage_predict)
      {
        retval = ark->stage_predict(ark->tcur, ark->zpred, ark->user_data);
        if (retval < 0) { return ARK_USER_PREDICT_FAIL; }
        if (retval > 0) { return TRY_AGAIN; }
      }
    }
    retval = stage_setup(ark, implicit_stage);
    if (retval != ARK_SUCCESS) { return retval; }
    if (implicit_stage)
    {
      *nflagPtr = stage_nls(ark, *nflagPtr);
      if (*nflagPtr != ARK_SUCCESS) { return TRY_AGAIN; }
    }
    else { N_VLinearSum(ONE, ark->yn, ONE, ark->sdata, ark->ycur); }
    retval = ark->fi(ark->tcur, ark->ycur, ark->Fi[is], ark->user_data);
    ark->nfi++;
    if (retval < 0) { return ARK_RHSFUNC_FAIL; }
    if (retval > 0) { return ARK_UNREC_RHSFUNC_ERR; }
    retval = ark->fe(ark->tn + ark->Be.c[is] * ark->h, ark->ycur, ark->Fe[is], ark->user_data);
    ark->nfe++;
    if (retval < 0) { return ARK_RHSFUNC_FAIL; }
    if (retval > 0) { return ARK_UNREC_RHSFUNC_ERR; }
  }
  *nflagPtr = compute_solutions(ark, dsmPtr);
  if (*nflagPtr < 0) { return *nflagPtr; }
  if (*nflagPtr > 0) { return TRY_AGAIN; }
  return ARK_SUCCESS;
}

/* ---------------------------------------------------------------------------
 * initial step (arkHin, arkUpperBoundH0, arkYddNorm)
 * -------------------------------------------------------------------------*/
static sunrealtype upper_bound_h0(ARKodeMem ark, sunrealtype tdist)
{
  N_Vector temp1 = ark->tempv1, temp2 = ark->tempv2;
  N_VAbs(ark->yn, temp2);
  ewt_set(ark, ark->yn, temp1);
  N_VInv(temp1, temp1);
  N_VLinearSum(H0_UBFACTOR, temp2, ONE, temp1, temp1);
  N_VAbs(ark->fn, temp2);
  N_VDiv(temp2, temp1, temp1);
  const sunrealtype hub_inv = N_VMaxNorm(temp1);
  sunrealtype hub           = H0_UBFACTOR * tdist;
  if (hub * hub_inv > ONE) { hub = ONE / hub_inv; }
  return hub;
}

static int ydd_norm(ARKodeMem ark, sunrealtype hg, sunrealtype* yddnrm)
{
  N_VLinearSum(hg, ark->fn, ONE, ark->yn, ark->ycur);
  int retval = full_rhs(ark, ark->tcur + hg, ark->ycur, ark->tempv1, ARK_FULLRHS_OTHER);
  if (retval != 0) { return ARK_RHSFUNC_FAIL; }
  N_VLinearSum(ONE / hg, ark->tempv1, -ONE / hg, ark->fn, ark->tempv1);
  N_VScale(ONE, ark->yn, ark->ycur);
  *yddnrm = N_VWrmsNorm(ark->tempv1, ark->ewt);
  return ARK_SUCCESS;
}

static int ark_hin(ARKodeMem ark, sunrealtype tout)
{
  sunrealtype tdiff, tdist, tround, hlb, hub, hg, hgs, hs, hnew = ZERO, hrat, h0, yddnrm;
  int retval, sign, count1, count2;
  sunbooleantype hgOK;
  if ((tdiff = tout - ark->tcur) == ZERO) { return ARK_TOO_CLOSE; }
  sign   = (tdiff > ZERO) ? 1 : -1;
  tdist  = SUNRabs(tdiff);
  tround = ark->uround * SUNMAX(SUNRabs(ark->tcur), SUNRabs(tout));
  if (tdist < TWO * tround) { return ARK_TOO_CLOSE; }
  if (!ark->fn_is_current)
  {
    retval = full_rhs(ark, ark->tn, ark->yn, ark->fn, ARK_FULLRHS_START);
    if (retval) { return ARK_RHSFUNC_FAIL; }
    ark->fn_is_current = SUNTRUE;
  }
  hlb = H0_LBFACTOR * tround;
  hub = upper_bound_h0(ark, tdist);
  hg  = SUNRsqrt(hlb * hub);
  if (hub < hlb)
  {
    ark->h = (sign == -1) ? -hg : hg;
    return ARK_SUCCESS;
  }
  hs = hg;
  for (count1 = 1; count1 <= H0_ITERS; count1++)
  {
    hgOK = SUNFALSE;
    for (count2 = 1; count2 <= H0_ITERS; count2++)
    {
      hgs    = hg * sign;
      retval = ydd_norm(ark, hgs, &yddnrm);
      if (retval < 0) { return ARK_RHSFUNC_FAIL; }
      if (retval == ARK_SUCCESS)
      {
        hgOK = SUNTRUE;
        break;
      }
      hg *= 0.2;
    }
    if (!hgOK)
    {
      if (count1 <= 2) { return ARK_REPTD_RHSFUNC_ERR; }
      hnew = hs;
      break;
    }
    hs   = hg;
    hnew = (yddnrm * hub * hub > TWO) ? SUNRsqrt(TWO / yddnrm) : SUNRsqrt(hg * hub);
    if (count1 == H0_ITERS) { break; }
    hrat = hnew / hg;
    if ((hrat > HALF) && (hrat < TWO)) { break; }
    if ((count1 > 1) && (hrat > TWO))
    {
      hnew = hg;
      break;
    }
    hg = hnew;
  }
  h0 = H0_BIAS * hnew;
  if (h0 < hlb) { h0 = hlb; }
  if (h0 > hub) { h0 = hub; }
  if (sign == -1) { h0 = -h0; }
  ark->h = h0;
  return ARK_SUCCESS;
}

/* ---------------------------------------------------------------------------
 * arkStep_Init + arkInitialSetup
 * -------------------------------------------------------------------------*/
static int initial_setup(ARKodeMem ark, sunrealtype tout)
{
  int retval;
  if (ark->fixedstep && ark->hin == ZERO) { return ARK_ILL_INPUT; }
  if (ark->tstopset)
  {
    const sunrealtype htmp = (ark->h == ZERO) ? tout - ark->tcur : ark->h;
    if ((ark->tstop - ark->tcur) * htmp <= ZERO) { return ARK_ILL_INPUT; }
  }
  if (ewt_set(ark, ark->yn, ark->ewt) != 0) { return ARK_ILL_INPUT; }
  /* arkStep_Init (FIRST_INIT) */
  if (load_tables(ark->user_order, &ark->Be, &ark->Bi) != 0) { return ARK_ILL_INPUT; }
  ark->stages = ark->Bi.s;
  ark->q = ark->hq_order = ark->Bi.q;
  ark->p = ark->hp_order = ark->Bi.p;
  if (!ark->fixedstep && ark->p <= 0) { return ARK_ILL_INPUT; }
  if (ark->LS)
  {
    /* arkLsInitialize: counters, then the solver */
    ark->nje = ark->npe = ark->nli = ark->nps = ark->ncfl = ark->njtsetup = ark->njtimes = 0;
    ark->nfeDQ = 0;
    ark->nstlj = 0;
    ark->lsetup_disabled = (ark->pset == NULL);
    if (SUNLinSolInitialize(ark->LS) != SUN_SUCCESS) { return ARK_LINIT_FAIL; }
  }
  if (ark->NLS)
  {
    retval = nls_init(ark);
    if (retval != ARK_SUCCESS) { return ARK_NLS_INIT_FAIL; }
  }
  ark->call_fullrhs = SUNTRUE;
  /* Hermite interpolation module: it needs f(tn, yn) at every step */
  if (ark->interp_type == ARK_INTERP_HERMITE) { ark->call_fullrhs = SUNTRUE; }
  /* set initial step size */
  if (ark->h0u == ZERO)
  {
    ark->h = ark->hin;
    if ((ark->h != ZERO) && ((tout - ark->tcur) * ark->h < ZERO)) { return ARK_ILL_INPUT; }
    if (ark->h == ZERO)
    {
      ark->h = SUNRabs(tout - ark->tcur);
      if (ark->h == ZERO) { ark->h = ONE; }
      sunrealtype tout_hin = tout;
      if (ark->tstopset && (tout - ark->tcur) * (tout - ark->tstop) > ZERO) { tout_hin = ark->tstop; }
      int hflag = ark_hin(ark, tout_hin);
      if (hflag != ARK_SUCCESS) { return hflag; }
      ark->etamax = ark->etamx1;
    }
    else if (ark->nst == 0) { ark->etamax = ark->etamx1; }
    else { ark->etamax = ark->growth; }
    sunrealtype rh = SUNRabs(ark->h) * ark->hmax_inv;
    if (rh > ONE) { ark->h /= rh; }
    if (SUNRabs(ark->h) < ark->hmin) { ark->h *= ark->hmin / SUNRabs(ark->h); }
    if (ark->tstopset)
    {
      if ((ark->tcur + ark->h - ark->tstop) * ark->h > ZERO)
      {
        ark->h = (ark->tstop - ark->tcur) * (ONE - FOUR * ark->uround);
      }
    }
    ark->h0u    = ark->h;
    ark->eta    = ONE;
    ark->hprime = ark->h;
  }
  else
  {
    if (ark->tstopset)
    {
      if ((ark->tcur + ark->hprime - ark->tstop) * ark->h > ZERO)
      {
        ark->hprime = (ark->tstop - ark->tcur) * (ONE - FOUR * ark->uround);
        ark->eta    = ark->hprime / ark->h;
      }
    }
  }
  return ARK_SUCCESS;
}

/* ---------------------------------------------------------------------------
 * step outcome handling
 * -------------------------------------------------------------------------*/
static int check_convergence(ARKodeMem ark, int* nflagPtr, int* ncfPtr)
{
  if (*nflagPtr == ARK_SUCCESS) { return ARK_SUCCESS; }
  ark->ncfn++;
  if (ark->fixedstep) { return ARK_CONV_FAILURE; }
  if (*nflagPtr < 0)
  {
    if (*nflagPtr == ARK_LSETUP_FAIL) { return ARK_LSETUP_FAIL; }
    if (*nflagPtr == ARK_LSOLVE_FAIL) { return ARK_LSOLVE_FAIL; }
    if (*nflagPtr == ARK_RHSFUNC_FAIL) { return ARK_RHSFUNC_FAIL; }
    return ARK_NLS_OP_ERR;
  }
  (*ncfPtr)++;
  ark->etamax = ONE;
  if ((*ncfPtr == ark->maxncf) || (SUNRabs(ark->h) <= ark->hmin * ONEPSM))
  {
    if (*nflagPtr == CONV_FAIL) { return ARK_CONV_FAILURE; }
    if (*nflagPtr == RHSFUNC_RECVR) { return ARK_REPTD_RHSFUNC_ERR; }
  }
  ark->eta  = ark->etacf;
  *nflagPtr = PREV_CONV_FAIL;
  return PREDICT_AGAIN;
}

/* arkAdapt */
static int adapt(ARKodeMem ark, sunrealtype hcur, sunrealtype dsm)
{
  if (!ark->has_controller)
  {
    ark->eta = ONE;
    return ARK_SUCCESS;
  }
  int controller_order;
  if (ark->pq == 0) { controller_order = ark->hp_order + ark->adjust; }
  else if (ark->pq == 1) { controller_order = ark->hq_order + ark->adjust; }
  else { controller_order = SUNMIN(ark->hp_order, ark->hq_order) + ark->adjust; }
  sunrealtype h_acc = ctrl_estimate(&ark->C, hcur, controller_order, dsm);
  h_acc *= ark->safety;
  h_acc = SUNMIN(SUNRabs(h_acc), SUNRabs(ark->etamax * hcur));
  h_acc = SUNMAX(h_acc, SUNRabs(ark->etamin * hcur));
  ark->nst_acc++;
  if (dsm <= ONE)
  {
    if ((h_acc > SUNRabs(hcur * ark->lbound * ONEMSM)) &&
        (h_acc < SUNRabs(hcur * ark->ubound * ONEPSM)))
    {
      h_acc = hcur;
    }
  }
  h_acc    = copysign(h_acc, hcur);
  ark->eta = h_acc / hcur;
  ark->eta = SUNMAX(ark->eta, ark->hmin / SUNRabs(hcur));
  ark->eta /= SUNMAX(ONE, SUNRabs(hcur) * ark->hmax_inv * ark->eta);
  return ARK_SUCCESS;
}

static int check_temporal_error(ARKodeMem ark, int* nflagPtr, int* nefPtr, sunrealtype dsm)
{
  if (adapt(ark, ark->h, dsm) != ARK_SUCCESS) { return ARK_ERR_FAILURE; }
  ark->eta = SUNMIN(ark->eta, ark->etamax);
  ark->eta = SUNMAX(ark->eta, ark->hmin / SUNRabs(ark->h));
  ark->eta /= SUNMAX(ONE, SUNRabs(ark->h) * ark->hmax_inv * ark->eta);
  if (dsm <= ONE) { return ARK_SUCCESS; }
  (*nefPtr)++;
  ark->netf++;
  *nflagPtr = PREV_ERR_FAIL;
  if (*nefPtr == ark->maxnef) { return ARK_ERR_FAILURE; }
  ark->etamax = ONE;
  if (*nefPtr >= ark->small_nef) { ark->eta = SUNMIN(ark->eta, ark->etamxf); }
  ark->eta = SUNMIN(ark->eta, ark->etamax);
  ark->eta = SUNMAX(ark->eta, ark->hmin / SUNRabs(ark->h));
  ark->eta /= SUNMAX(ONE, SUNRabs(ark->h) * ark->hmax_inv * ark->eta);
  return TRY_AGAIN;
}

static int complete_step(ARKodeMem ark, sunrealtype dsm)
{
  ark->tcur = ark->tn + ark->h;
  if (ark->tstopset)
  {
    const sunrealtype troundoff = FUZZ_FACTOR * ark->uround * (SUNRabs(ark->tcur) + SUNRabs(ark->h));
    if (SUNRabs(ark->tcur - ark->tstop) <= troundoff) { ark->tcur = ark->tstop; }
  }
  /* Hermite interpolation update: needs f(tn, yn) of the step's start */
  if (ark->interp_type == ARK_INTERP_HERMITE && !ark->fn_is_current)
  {
    int retval = full_rhs(ark, ark->tn, ark->yn, ark->fn, ARK_FULLRHS_START);
    if (retval) { return ARK_RHSFUNC_FAIL; }
    ark->fn_is_current = SUNTRUE;
  }
  N_VScale(ONE, ark->ycur, ark->yn);
  ark->fn_is_current = SUNFALSE;
  if (ark->has_controller) { ctrl_update(&ark->C, ark->h, dsm); }
  ark->nst++;
  ark->hold   = ark->h;
  ark->tn     = ark->tcur;
  ark->hprime = ark->h * ark->eta;
  ark->etamax = ark->growth;
  ark->initsetup  = SUNFALSE;
  ark->firststage = SUNFALSE;
  return ARK_SUCCESS;
}

/* ---------------------------------------------------------------------------
 * ARKodeEvolve
 * -------------------------------------------------------------------------*/
int ARKodeEvolve(void* arkode_mem, sunrealtype tout, N_Vector yout, sunrealtype* tret,
                 int itask)
{
  ARKMEM_OR_RETURN();
  int retval, kflag, istate = ARK_SUCCESS, nflag, ncf, nef;
  long int nstloc;
  sunrealtype troundoff, nrm, dsm;
  if (yout == NULL || tret == NULL) { return ARK_ILL_INPUT; }
  if ((itask != ARK_NORMAL) && (itask != ARK_ONE_STEP)) { return ARK_ILL_INPUT; }
  if (ark->initsetup)
  {
    ark->tretlast = *tret = ark->tcur;
    retval                = initial_setup(ark, tout);
    if (retval != ARK_SUCCESS) { return retval; }
  }
  else
  {
    /* arkStopTests */
    troundoff = FUZZ_FACTOR * ark->uround * (SUNRabs(ark->tcur) + SUNRabs(ark->h));
    if (ark->tstopset)
    {
      if (SUNRabs(ark->tcur - ark->tstop) <= troundoff)
      {
        if ((tout - ark->tstop) * ark->h >= ZERO || SUNRabs(tout - ark->tstop) <= troundoff)
        {
          N_VScale(ONE, ark->yn, yout);
          ark->tretlast = *tret = ark->tstop;
          ark->tstopset         = SUNFALSE;
          return ARK_TSTOP_RETURN;
        }
      }
      else if ((ark->tcur + ark->hprime - ark->tstop) * ark->h > ZERO)
      {
        ark->hprime = (ark->tstop - ark->tcur) * (ONE - FOUR * ark->uround);
        ark->eta    = ark->hprime / ark->h;
      }
    }
    if ((itask == ARK_NORMAL) && ((ark->tcur - tout) * ark->h >= ZERO))
    {
      /* dense output is not part of this library: return the step's state */
      N_VScale(ONE, ark->yn, yout);
      ark->tretlast = *tret = ark->tcur;
      return ARK_SUCCESS;
    }
  }

  nstloc = 0;
  for (;;)
  {
    ark->next_h = ark->h;
    if (!ark->initsetup)
    {
      if (ewt_set(ark, ark->yn, ark->ewt) != 0)
      {
        istate        = ARK_ILL_INPUT;
        ark->tretlast = *tret = ark->tcur;
        N_VScale(ONE, ark->yn, yout);
        break;
      }
    }
    if ((ark->mxstep > 0) && (nstloc >= ark->mxstep))
    {
      istate        = ARK_TOO_MUCH_WORK;
      ark->tretlast = *tret = ark->tcur;
      N_VScale(ONE, ark->yn, yout);
      break;
    }
    nrm        = N_VWrmsNorm(ark->yn, ark->ewt);
    ark->tolsf = ark->uround * nrm;
    if (ark->tolsf > ONE && !ark->fixedstep)
    {
      istate        = ARK_TOO_MUCH_ACC;
      ark->tretlast = *tret = ark->tcur;
      N_VScale(ONE, ark->yn, yout);
      ark->tolsf *= TWO;
      break;
    }
    else { ark->tolsf = ONE; }
    if (ark->tcur + ark->h == ark->tcur) { ark->nhnil++; }
    if (ark->hprime != ark->h)
    {
      ark->h      = ark->h * ark->eta;
      ark->next_h = ark->h;
    }
    if (ark->fixedstep)
    {
      ark->h      = ark->hin;
      ark->next_h = ark->h;
      if (ark->tstopset)
      {
        if ((ark->tcur + ark->h - ark->tstop) * ark->h > ZERO)
        {
          ark->h = (ark->tstop - ark->tcur) * (ONE - FOUR * ark->uround);
        }
      }
    }
    dsm   = ZERO;
    kflag = ARK_SUCCESS;
    ncf = nef = 0;
    nflag     = FIRST_CALL;
    for (;;)
    {
      ark->nst_attempts++;
      kflag = take_step(ark, &dsm, &nflag);
      if (kflag < 0) { break; }
      kflag = check_convergence(ark, &nflag, &ncf);
      if (kflag < 0) { break; }
      if (ark->fixedstep)
      {
        ark->eta = ONE;
        break;
      }
      if (kflag == ARK_SUCCESS)
      {
        kflag = check_temporal_error(ark, &nflag, &nef, dsm);
        if (kflag < 0) { break; }
      }
      if (kflag == ARK_SUCCESS) { break; }
      if (SUNRabs(ark->h) <= ark->hmin * ONEPSM) { return ARK_ERR_FAILURE; }
      ark->h *= ark->eta;
      ark->next_h = ark->hprime = ark->h;
    }
    if (kflag == ARK_SUCCESS) { kflag = complete_step(ark, dsm); }
    if (kflag != ARK_SUCCESS)
    {
      istate        = kflag;
      ark->tretlast = *tret = ark->tcur;
      N_VScale(ONE, ark->yn, yout);
      break;
    }
    nstloc++;
    if (ark->tstopset)
    {
      troundoff = FUZZ_FACTOR * ark->uround * (SUNRabs(ark->tcur) + SUNRabs(ark->h));
      if (SUNRabs(ark->tcur - ark->tstop) <= troundoff)
      {
        if ((tout - ark->tstop) * ark->h >= ZERO || SUNRabs(tout - ark->tstop) <= troundoff)
        {
          N_VScale(ONE, ark->yn, yout);
          ark->tretlast = *tret = ark->tstop;
          ark->tstopset         = SUNFALSE;
          istate                = ARK_TSTOP_RETURN;
          break;
        }
      }
      else if ((ark->tcur + ark->hprime - ark->tstop) * ark->h > ZERO)
      {
        ark->hprime = (ark->tstop - ark->tcur) * (ONE - FOUR * ark->uround);
        ark->eta    = ark->hprime / ark->h;
      }
    }
    if ((itask == ARK_NORMAL) && (ark->tcur - tout) * ark->h >= ZERO)
    {
      /* no dense output: the step's state at tcur */
      N_VScale(ONE, ark->yn, yout);
      ark->tretlast = *tret = ark->tcur;
      ark->next_h           = ark->hprime;
      istate                = ARK_SUCCESS;
      break;
    }
    if (itask == ARK_ONE_STEP)
    {
      ark->tretlast = *tret = ark->tcur;
      N_VScale(ONE, ark->yn, yout);
      ark->next_h = ark->hprime;
      istate      = ARK_SUCCESS;
      break;
    }
  }
  return istate;
}

/* ---------------------------------------------------------------------------
 * statistics
 * -------------------------------------------------------------------------*/
#define GETTER(name, type, expr)                 \
  int name(void* arkode_mem, type* out)          \
  {                                              \
    ARKMEM_OR_RETURN();                          \
    *out = (expr);                               \
    return ARK_SUCCESS;                          \
  }

GETTER(ARKodeGetNumSteps, long int, ark->nst)
GETTER(ARKodeGetNumStepAttempts, long int, ark->nst_attempts)
GETTER(ARKodeGetNumErrTestFails, long int, ark->netf)
GETTER(ARKodeGetNumConvFails, long int, ark->ncfn)
GETTER(ARKodeGetNumNonlinSolvIters, long int, ark->nls_iters)
GETTER(ARKodeGetNumNonlinSolvConvFails, long int, ark->nls_fails)
GETTER(ARKodeGetNumLinSolvSetups, long int, ark->nsetups)
GETTER(ARKodeGetNumLinIters, long int, ark->nli)
GETTER(ARKodeGetNumLinConvFails, long int, ark->ncfl)
GETTER(ARKodeGetNumPrecEvals, long int, ark->npe)
GETTER(ARKodeGetNumPrecSolves, long int, ark->nps)
GETTER(ARKodeGetNumJtimesEvals, long int, ark->njtimes)
GETTER(ARKodeGetNumLinRhsEvals, long int, ark->nfeDQ)
GETTER(ARKodeGetLastStep, sunrealtype, ark->hold)
GETTER(ARKodeGetActualInitStep, sunrealtype, ark->h0u)
GETTER(ARKodeGetCurrentStep, sunrealtype, ark->next_h)
GETTER(ARKodeGetCurrentTime, sunrealtype, ark->tcur)

int ARKStepGetNumRhsEvals(void* arkode_mem, long int* nfe_evals, long int* nfi_evals)
{
  ARKMEM_OR_RETURN();
  *nfe_evals = ark->nfe;
  *nfi_evals = ark->nfi;
  return ARK_SUCCESS;
}
