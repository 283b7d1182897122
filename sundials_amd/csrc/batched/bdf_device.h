// -----------------------------------------------------------------------------
// Batched per-system BDF on the MI355X (device part): one persistent launch per size bucket,
// one group of G = pow2ceil(m) lanes per system (lane = species), groups pull
// systems from a device work queue until the bucket is done.
//
// Every decision follows the reference CVODE exactly as the host integrator
// (cvode_bdf.c) restates it for a serial vector of length m and a dense LU:
//   CVode first call / loop        cvode.c:1037-1670
//   cvHin, cvUpperBoundH0, cvYddNorm cvode.c:2161-2340
//   cvStep, cvAdjustParams/Order, cvIncreaseBDF, cvDecreaseBDF, cvRescale,
//   cvPredict, cvSet/SetBDF/SetTqBDF  cvode.c:2367-3053
//   cvNls, cvHandleNFlag, cvRestore, cvDoErrorTest, cvCompleteStep,
//   cvPrepareNextStep, cvSetEta, cvComputeEtaqm1/qp1, cvChooseEta
//                                  cvode.c:3068-3711
//   cvEwtSetSS                     cvode.c:4791-4817
//   residual / conv. test / setup  cvode_nls.c:256-399
//   J reuse, A = I - gamma J, solve scaling  cvode_ls.c:1311-1850
//   Newton loop                    sunnonlinsol_newton.c:179-324
//   dense GETRF / GETRS            sundials_dense.c:85-186
// Vector operations are per-lane arithmetic with the serial vector's
// per-element formulas (LinearSum's three rounding forms), norms are summed in
// species order through LDS, so in FP64 each system's result equals the host
// integrator's bit for bit.
// -----------------------------------------------------------------------------
#ifndef SUNAMD_BDF_DEVICE_H
#define SUNAMD_BDF_DEVICE_H

#include <hip/hip_runtime.h>

#include <climits>

#include "../core/lane_ops.h"
#include "../core/sunamd_pow.h"
#include "bdf_args.h"
#include "sunamd/batched_bdf.h"

namespace sunamd_bdf
{

// ---- constants (cvode_impl.h, cvode.c, cvode_ls_impl.h, cvode_nls.c) ------
constexpr int QMAX             = 5;
constexpr double ETA_MIN_FX    = 0.0;
constexpr double ETA_MAX_FX    = 1.5;
constexpr double ETA_MAX_FS    = 10000.0;
constexpr double ETA_MAX_ES    = 10.0;
constexpr double ETA_MAX_GS    = 10.0;
constexpr double ETA_MIN       = 0.1;
constexpr double ETA_MAX_EF    = 0.2;
constexpr double ETA_MIN_EF    = 0.1;
constexpr double ETA_CF        = 0.25;
constexpr long SMALL_NST       = 10;
constexpr int SMALL_NEF        = 2;
constexpr double ONEPSM        = 1.000001;
constexpr double ADDON         = 0.000001;
constexpr double BIAS1         = 6.0;
constexpr double BIAS2         = 6.0;
constexpr double BIAS3         = 10.0;
constexpr int LONG_WAIT        = 10;
constexpr int MXNCF            = 10;
constexpr int MXNEF            = 7;
constexpr int MXNEF1           = 3;
constexpr long MSBP            = 20;
constexpr double DGMAX_LSETUP  = 0.3;
constexpr long MSBJ            = 51;
constexpr double DGMAX_JBAD    = 0.2;
constexpr double CORTES        = 0.1;
constexpr int NLS_MAXCOR       = 3;
constexpr double CRDOWN        = 0.3;
constexpr double RDIV          = 2.0;
constexpr double HLB_FACTOR    = 100.0;
constexpr double HUB_FACTOR    = 0.1;
constexpr double H_BIAS        = 0.5;
constexpr int MAX_ITERS        = 4;
constexpr double UROUND        = 2.2204460492503131e-16;

// flags (cvode.h / cvode_impl.h / sundials_nonlinearsolver.h)
constexpr int CV_SUCCESS          = 0;
constexpr int CV_TOO_MUCH_WORK    = -1;
constexpr int CV_TOO_MUCH_ACC     = -2;
constexpr int CV_ERR_FAILURE      = -3;
constexpr int CV_CONV_FAILURE     = -4;
constexpr int CV_LSETUP_FAIL      = -6;
constexpr int CV_TOO_CLOSE        = -27;
constexpr int DO_ERROR_TEST       = 2;
constexpr int PREDICT_AGAIN       = 3;
constexpr int TRY_AGAIN           = 5;
constexpr int FIRST_CALL          = 6;
constexpr int PREV_CONV_FAIL      = 7;
constexpr int PREV_ERR_FAIL       = 9;
constexpr int NLS_CONTINUE        = 901;
constexpr int NLS_CONV_RECVR      = 902;
constexpr int CV_NO_FAILURES      = 0;
constexpr int CV_FAIL_BAD_J       = 1;
constexpr int CV_FAIL_OTHER       = 2;
constexpr int LUFACT_FAIL         = 808;


// serial-vector LinearSum rounding: its nine branches reduce to these three
__device__ __forceinline__ double lsum(double a, double x, double b, double y)
{
  if (a == b) { return a * (x + y); }
  if (a == -b) { return a * (x - y); }
  return (a * x) + (b * y);
}

__device__ __forceinline__ double sqrt0(double v) { return (v <= 0.0) ? 0.0 : sqrt(v); }

// wavefront-level ordering point for a group's LDS exchange
__device__ __forceinline__ void gsync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// per-group scalar arrays live in LDS (indexed by the order q at run time)
struct GroupScalars
{
  double tau[QMAX + 2];
  double tq[6];
  double l[QMAX + 1];
  double cvals[QMAX + 2];
#ifdef SUNAMD_BDF_PHASES
  unsigned long long ph_t;                  // time stamp of the last phase change
  unsigned long long ph_acc[SUNAMD_BDF_NPHASES]; // clock ticks spent in each phase
  unsigned int ph_cnt[SUNAMD_BDF_NPHASES];  // entries into each phase
  int ph_cur;                               // current phase
#endif
};

template <int M, typename T>
struct Sys
{
  // lane-owned vector elements (species i of the group's system)
  double zn[QMAX + 1];
  double ewt, acor, tempv, ftemp, y, delta, ysav;
  // Newton-matrix factor row held by this lane, and its physical row
  T lu[M];
  int pos;
  float rinv; // FP32 factors: 1 / this lane's diagonal entry
  // integrator scalars (uniform across the group)
  double tn, h, hscale, hprime, eta, hu, h0u;
  double rl1, gamma, gammap, gamrat, crate, delp, acnrm;
  double etaqm1, etaq, etaqp1, etamax, saved_tq5;
  int q, qprime, qwait, L, indx_acor, convfail, qu;
  int nst, nfe, ncfn, netf, nni, nnf, nsetups, nstlp, nje, nstlj;
  bool nls_jcur, acnrmcur;
  // learned initial iterate: this species' rate feature, and counts of the
  // steps where a guess was tried / kept
  float ratef;
  int ltried, lused;
};

template <int M, typename T>
struct Grp
{
  int i;       // lane within group (= species when valid)
  int base;    // first lane of the group in the wavefront
  bool valid;  // i < M
  double* red; // G doubles of LDS
  T* prow;     // G entries of LDS
  int* inv;    // G ints of LDS
  GroupScalars* sc;
  float* mlp;         // MLP_LOG_FLOATS of LDS staging a training-log row, or nullptr
  unsigned short* wx; // the wavefront's network area (MLP_WAVE_BYTES of LDS) or nullptr
  int row;            // this group's row in the wavefront's network inputs
  bool learn;         // kernel built with the learned-start / training-log paths
  const BucketArgs* a;
  // the bucket's network, staged in LDS once per workgroup
  const int* idx;  // 4R reaction indices
  const int* soff; // M+1 term-list offsets
  const int* sent; // 4R term codes
  double* rates;   // this group's 2R rate constants (k_f then k_b), staged per system
  T* jrow;         // this lane's Jacobian-row accumulator (M entries, LDS)
  double* wbuf;    // the group's reaction rates during an RHS (aliases the rows)
  long cell;
  const double* kf; // this system's k_f, k_b (the LDS slot above)
  const double* kb;
};

// ---- phase clock (instrumented build only) ----------------------------------
// Built with SUNAMD_BDF_PHASES, every group accumulates the shader clock
// (s_memtime) between phase changes into its LDS slot, counts the entries into
// each phase, and adds the totals to BucketArgs::phase when its wavefront
// leaves the kernel.  A phase change is a scalar clock read and, on the
// group's first lane, a few LDS updates; the arithmetic of the integrator is
// untouched.  Each group also times a run of back-to-back phase changes at
// its start, so the cost of one change can be taken off every phase it
// entered.  Phase numbers and the output layout: batched_bdf.h.
#ifdef SUNAMD_BDF_PHASES
template <int M, typename T>
__device__ __forceinline__ int ph_set(const Grp<M, T>& g, int p)
{
  // no instruction is scheduled across the clock read
  __builtin_amdgcn_sched_barrier(0);
  const unsigned long long t = __builtin_amdgcn_s_memtime();
  __builtin_amdgcn_sched_barrier(0);
  const int prev             = g.sc->ph_cur;
  if (g.i == 0)
  {
    g.sc->ph_acc[prev] += t - g.sc->ph_t;
    g.sc->ph_cnt[p] += 1u;
    g.sc->ph_t   = t;
    g.sc->ph_cur = p;
  }
  return prev;
}
#define BDF_PH_PUSH(g, p) const int ph_prev_ = ph_set(g, p)
#define BDF_PH_POP(g) ph_set(g, ph_prev_)
#define BDF_PH_SET(g, p) ph_set(g, p)
// a phase's result is complete before the clock is read at its end (the
// compiler may otherwise sink the work past the read)
#define BDF_PH_FENCE(v) asm volatile("" ::"v"(v))
#else
#define BDF_PH_PUSH(g, p) ((void)0)
#define BDF_PH_POP(g) ((void)0)
#define BDF_PH_SET(g, p) ((void)0)
#define BDF_PH_FENCE(v) ((void)0)
#endif

// The product form (single-precision factors) is free of the FP64 form's
// bit-for-bit contract with the host integrator, so it takes the fast forms:
// fused multiply-adds and no zero tests in the factorisation, a pivot search
// that is one unsigned maximum per column, (for groups of 32 or 64 lanes) a
// single-precision solve, WRMS norms summed as a butterfly
// instead of in species order, and x^y in the step-ratio formulas through the
// hardware log2 / exp2 (relative error ~1e-7, immaterial in a ratio that is
// clamped to [0.1, 10]).  Residuals, updates and the norms themselves stay
// FP64.  The double-precision form keeps the dense GETRF/GETRS operation
// order, species-order norms and the correctly rounded x^y.
template <typename T>
inline constexpr bool kFastFactor = (sizeof(T) == 4);

using sunamd_lane::dpp_i;
using sunamd_lane::dpp_v;
using sunamd_lane::readlane_d;

// Sum over a G-lane group as a butterfly: quad swaps, then half-row and row
// mirrors (each lane adds the lane its mirror pairs it with, which holds the
// other half's partial sum), then the four 16-lane rows by lane reads.  Every
// addition has the same two operands in both lanes of a pair, so every lane of
// the group ends with the same bits.
template <int G>
__device__ __forceinline__ double g_tree_sum(double v)
{
  if constexpr (G >= 2) { v = v + dpp_v<0xB1>(v); }  // quad_perm [1,0,3,2]
  if constexpr (G >= 4) { v = v + dpp_v<0x4E>(v); }  // quad_perm [2,3,0,1]
  if constexpr (G >= 8) { v = v + dpp_v<0x141>(v); } // row_half_mirror
  if constexpr (G >= 16) { v = v + dpp_v<0x140>(v); } // row_mirror
  if constexpr (G >= 32)
  {
    const double r0 = readlane_d(v, 0), r1 = readlane_d(v, 16);
    const double r2 = readlane_d(v, 32), r3 = readlane_d(v, 48);
    const double lo = r0 + r1, hi = r2 + r3;
    if constexpr (G == 64) { v = lo + hi; }
    else { v = (__lane_id() >= 32) ? hi : lo; }
  }
  return v;
}

// x^y for the step-ratio formulas of the product form: x = 2^e * mant exactly
// (frexp), log2 of the mantissa and exp2 of the fraction on the hardware's
// single-precision units, the powers of two exact (ldexp).  x >= 0, 0 < y <= 1.
__device__ __forceinline__ double fast_pow(double x, double y)
{
  if (!(x > 0.0)) { return 0.0; }
  if (!(x < __builtin_huge_val())) { return x; }
  int e           = 0;
  const double mt = frexp(x, &e);
  const double t  = y * ((double)e + (double)__builtin_amdgcn_logf((float)mt));
  const double fl = floor(t);
  const double r  = (double)__builtin_amdgcn_exp2f((float)(t - fl));
  return ldexp(r, (int)fl);
}

// ---- group reductions (species order = the serial vector's order) --------
template <int M, typename T>
__device__ __forceinline__ double g_wrms(const Grp<M, T>& g, double x, double w)
{
  BDF_PH_PUSH(g, SUNAMD_PH_NORM);
  double p = x * w;
  if constexpr (kFastFactor<T>)
  {
    const double s = g_tree_sum<pow2ceil(M)>(g.valid ? p * p : 0.0);
    BDF_PH_FENCE(s);
    BDF_PH_POP(g);
    return sqrt0(s / (double)M);
  }
  gsync();
  if (g.valid) { g.red[g.i] = p * p; }
  gsync();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < M; k++) { s += g.red[k]; }
  gsync();
  BDF_PH_FENCE(s);
  BDF_PH_POP(g);
  return sqrt0(s / (double)M);
}

template <int M, typename T>
__device__ __forceinline__ double g_maxabs(const Grp<M, T>& g, double x)
{
  gsync();
  if (g.valid) { g.red[g.i] = fabs(x); }
  gsync();
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < M; k++)
  {
    double v = g.red[k];
    if (v > mx) { mx = v; }
  }
  gsync();
  return mx;
}

// ---- mass-action RHS: f_i accumulated in reaction order ------------------
// Each reaction's net rate w_r = (k_f y_a) y_b - (k_b y_c) y_d is evaluated
// once (reactions spread over the group's lanes) into LDS; species i then
// adds its terms -w (a, b roles) and +w (c, d roles) in reaction and role
// order, the serial RHS's order of updates of f_i.
template <int M, typename T>
__device__ __forceinline__ double g_rhs(const Grp<M, T>& g, double yi)
{
  BDF_PH_PUSH(g, SUNAMD_PH_RHS);
  constexpr int G = pow2ceil(M);
  gsync();
  if (g.valid) { g.red[g.i] = yi; }
  gsync();
  const int R = g.a->R;
  for (int r = g.i; r < R; r += G)
  {
    const int4 q   = *reinterpret_cast<const int4*>(g.idx + 4 * r);
    const double wf = (g.kf[r] * g.red[q.x]) * g.red[q.y];
    const double wb = (g.kb[r] * g.red[q.z]) * g.red[q.w];
    g.wbuf[r]       = wf - wb;
  }
  gsync();
  double f = 0.0;
  if (g.valid)
  {
    // four terms per pass: their codes, then their rates, are read together
    // (independent LDS reads in flight), then added in term order
    const int e1 = g.soff[g.i + 1];
    for (int e = g.soff[g.i]; e < e1; e += 4)
    {
      int code[4];
      double w[4];
#pragma unroll
      for (int q = 0; q < 4; q++) { code[q] = (e + q < e1) ? g.sent[e + q] : 0; }
#pragma unroll
      for (int q = 0; q < 4; q++) { w[q] = g.wbuf[code[q] >> 2]; }
#pragma unroll
      for (int q = 0; q < 4; q++)
      {
        if (e + q < e1) { f = ((code[q] & 3) < 2) ? f - w[q] : f + w[q]; }
      }
    }
  }
  gsync();
  BDF_PH_FENCE(f);
  BDF_PH_POP(g);
  return f;
}

// ---- Jacobian row i in reaction / target / column order, then the Newton
// matrix row (I - gamma*J)_i with the dense ScaleAddI rounding --------------
template <int M, typename T>
__device__ __forceinline__ void g_newton_matrix(const Grp<M, T>& g, double yj, double gamma,
                                                T (&a)[M])
{
  gsync();
  if (g.valid) { g.red[g.i] = yj; }
  gsync();
  // the row is accumulated in LDS so that a term's column is a plain address
  // (the same additions in the same order as a register row); the FP32 form
  // accumulates in single precision (half the LDS, more wavefronts per CU)
  T* row         = g.jrow;
  const int* idx = g.idx;
  if (g.valid)
  {
#pragma unroll
    for (int j = 0; j < M; j++) { row[j] = (T)0; }
    const int e1 = g.soff[g.i + 1];
    for (int e = g.soff[g.i]; e < e1; e++)
    {
      const int code = g.sent[e];
      const int r = code >> 2, t = code & 3;
      int ix[4]    = {idx[4 * r], idx[4 * r + 1], idx[4 * r + 2], idx[4 * r + 3]};
      double kfv   = g.kf[r], kbv = g.kb[r];
      double dw[4] = {kfv * g.red[ix[1]], kfv * g.red[ix[0]], -(kbv * g.red[ix[3]]),
                      -(kbv * g.red[ix[2]])};
#pragma unroll
      for (int q = 0; q < 4; q++)
      {
        const T v  = (T)dw[q];
        row[ix[q]] = (t < 2) ? (row[ix[q]] - v) : (row[ix[q]] + v);
      }
    }
  }
  gsync();
  // (I - gamma J)_ij in FP64, rounded once to the factor's precision
  const double c = -gamma;
#pragma unroll
  for (int j = 0; j < M; j++)
  {
    double v = g.valid ? (double)row[j] * c : 0.0;
    if (j == g.i) { v = v + 1.0; }
    a[j] = (T)v;
  }
}

// b of lane base+k of this lane's group, for groups of 64 or 32 lanes
template <int G>
__device__ __forceinline__ double g_lane_value(double b, int k)
{
  if constexpr (G == 64) { return readlane_d(b, k); }
  else
  {
    const double b0 = readlane_d(b, k);
    const double b1 = readlane_d(b, 32 + k);
    return (__lane_id() < 32) ? b0 : b1;
  }
}

template <int G>
__device__ __forceinline__ float g_lane_value(float b, int k)
{
  if constexpr (G == 64) { return readlane_d(b, k); }
  else
  {
    const float b0 = readlane_d(b, k);
    const float b1 = readlane_d(b, 32 + k);
    return (__lane_id() < 32) ? b0 : b1;
  }
}

// pivot search: (value, row) maximum over the group, ties to the smaller row
using sunamd_lane::g_argmax;


// unsigned maximum over a G-lane group (every lane ends with it)
template <int G>
__device__ __forceinline__ unsigned g_max_u32(unsigned v)
{
  auto mx = [](unsigned a, unsigned b) { return a > b ? a : b; };
  if constexpr (G >= 2) { v = mx(v, (unsigned)dpp_i<0xB1>((int)v)); }
  if constexpr (G >= 4) { v = mx(v, (unsigned)dpp_i<0x4E>((int)v)); }
  if constexpr (G >= 8) { v = mx(v, (unsigned)dpp_i<0x141>((int)v)); }
  if constexpr (G >= 16) { v = mx(v, (unsigned)dpp_i<0x140>((int)v)); }
  if constexpr (G >= 32)
  {
    const unsigned r0 = (unsigned)__builtin_amdgcn_readlane((int)v, 0);
    const unsigned r1 = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
    const unsigned r2 = (unsigned)__builtin_amdgcn_readlane((int)v, 32);
    const unsigned r3 = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
    const unsigned lo = mx(r0, r1), hi = mx(r2, r3);
    if constexpr (G == 64) { v = mx(lo, hi); }
    else { v = (__lane_id() >= 32) ? hi : lo; }
  }
  return v;
}

// ---- LU of the group's matrix (rows in lanes, implicit pivoting) ---------
template <int M, typename T>
__device__ __forceinline__ int g_getrf(const Grp<M, T>& g, Sys<M, T>& S)
{
  constexpr int G = pow2ceil(M);
  int pos  = g.i;
  int fail = 0;
#pragma unroll
  for (int k = 0; k < M; k++)
  {
    T v;
    int bp;
    if constexpr (kFastFactor<T>)
    {
      // one unsigned maximum on |a_ik|'s bits with the low six replaced by
      // 63 - row: pivots within 2^-17 of the largest tie to the smallest row
      const unsigned key = (g.valid && pos >= k)
                             ? ((__float_as_uint(fabsf((float)S.lu[k])) & ~63u) | (unsigned)(63 - pos))
                             : 0u;
      const unsigned mk = g_max_u32<G>(key);
      v  = (T)__uint_as_float(mk & ~63u);
      bp = 63 - (int)(mk & 63u);
    }
    else
    {
      v  = (g.valid && pos >= k) ? (T)fabs((double)S.lu[k]) : (T)-1;
      bp = g.valid ? pos : INT_MAX;
      g_argmax<G>(v, bp);
    }
    if (v == (T)0 && fail == 0) { fail = k + 1; }
    if (pos == k) { pos = bp; }
    else if (pos == bp) { pos = k; }
    if constexpr (G == 64)
    {
      // one system per wavefront: the pivot row's lane is wavefront-uniform,
      // its entries are read straight from that lane's registers
      const unsigned long long holder = __ballot(g.valid && pos == k);
      const int pl  = __builtin_amdgcn_readfirstlane(__ffsll((long long)holder) - 1) & 63;
      const bool up = g.valid && pos > k && fail == 0;
      const T pk    = readlane_d(S.lu[k], pl);
      if constexpr (kFastFactor<T>)
      {
        // single precision: multiplier 0 on the rows not updated, fused
        // multiply-add, no zero test
        const T l = up ? S.lu[k] * ((T)1 / pk) : (T)0;
        if (up) { S.lu[k] = l; }
#pragma unroll
        for (int j = k + 1; j < M; j++) { S.lu[j] = __builtin_fmaf(-readlane_d(S.lu[j], pl), l, S.lu[j]); }
      }
      else
      {
        T l = (T)0;
        if (up)
        {
          const T mult = (T)1 / pk;
          l            = S.lu[k] * mult;
          S.lu[k]      = l;
        }
#pragma unroll
        for (int j = k + 1; j < M; j++)
        {
          const T pj = readlane_d(S.lu[j], pl);
          if (up && pj != (T)0) { S.lu[j] = S.lu[j] - pj * l; }
        }
      }
    }
    else
    {
      gsync();
      if (g.valid && pos == k)
      {
#pragma unroll
        for (int j = k; j < M; j++) { g.prow[j] = S.lu[j]; }
      }
      gsync();
      const bool up = g.valid && pos > k && fail == 0;
      if constexpr (kFastFactor<T>)
      {
        const T l = up ? S.lu[k] * ((T)1 / g.prow[k]) : (T)0;
        if (up) { S.lu[k] = l; }
#pragma unroll
        for (int j = k + 1; j < M; j++) { S.lu[j] = __builtin_fmaf(-g.prow[j], l, S.lu[j]); }
      }
      else if (up)
      {
        const T mult = (T)1 / g.prow[k];
        const T l    = S.lu[k] * mult;
        S.lu[k]      = l;
#pragma unroll
        for (int j = k + 1; j < M; j++)
        {
          const T pj = g.prow[j];
          if (pj != (T)0) { S.lu[j] = S.lu[j] - pj * l; }
        }
      }
    }
  }
  gsync();
  if (g.valid) { g.inv[pos] = g.i; }
  gsync();
  S.pos = pos;
  if constexpr (G >= 32)
  {
    // one or two systems per wavefront: move physical row j to lane base+j,
    // so the solves read row k's element from a fixed lane
    const int src = g.base + g.inv[g.valid ? g.i : 0];
#pragma unroll
    for (int j = 0; j < M; j++) { S.lu[j] = __shfl(S.lu[j], src, 64); }
    S.pos = g.i;
    if constexpr (kFastFactor<T>)
    {
      // single precision: keep 1/u_ii per lane and store the strict upper
      // part column-scaled, u_ij / u_jj, so the backward solve's chain is a
      // lane read and a fused multiply-add per column
      float d = S.lu[0];
#pragma unroll
      for (int j = 1; j < M; j++) { d = (j == g.i) ? S.lu[j] : d; }
      S.rinv = 1.0f / d;
#pragma unroll
      for (int j = 1; j < M; j++)
      {
        const float rj = g_lane_value<G>(S.rinv, j);
        S.lu[j]        = (j > g.i) ? S.lu[j] * rj : S.lu[j];
      }
    }
  }
  return fail;
}

// x = (LU)^{-1} P b for b held as one element per lane
template <int M, typename T>
__device__ __forceinline__ double g_getrs(const Grp<M, T>& g, const Sys<M, T>& S, double b)
{
  // the row swaps carried logical row r to physical row pos_r, and b's
  // element r to the same position: lane r already holds b'[pos_r]
  const int pos = S.pos;
  constexpr int G = pow2ceil(M);
  if constexpr (G >= 32)
  {
    // rows sit in lane order (physical row j in lane base+j): gather b' (b's
    // element of row j is held by the lane the row came from), then read
    // row k's element from lane base+k; x_i ends in lane base+i
    b = __shfl(b, g.base + g.inv[g.valid ? g.i : 0], 64);
    if constexpr (kFastFactor<T>)
    {
      // single-precision solve: multipliers selected off the dependency
      // chain (0 on lanes a column does not update), U column-scaled
      float x = (float)b;
#pragma unroll
      for (int k = 0; k < M - 1; k++)
      {
        const float lk = (g.i > k) ? S.lu[k] : 0.0f;
        x              = __builtin_fmaf(-lk, g_lane_value<G>(x, k), x);
      }
#pragma unroll
      for (int k = M - 1; k > 0; k--)
      {
        const float uk = (g.i < k) ? S.lu[k] : 0.0f;
        x              = __builtin_fmaf(-uk, g_lane_value<G>(x, k), x);
      }
      return (double)(x * S.rinv);
    }
#pragma unroll
    for (int k = 0; k < M - 1; k++)
    {
      const double bk = g_lane_value<G>(b, k);
      if (g.i > k) { b = b - (double)S.lu[k] * bk; }
    }
#pragma unroll
    for (int k = M - 1; k > 0; k--)
    {
      if (g.i == k) { b = b / (double)S.lu[k]; }
      const double bk = g_lane_value<G>(b, k);
      if (g.i < k) { b = b - (double)S.lu[k] * bk; }
    }
    if (g.i == 0) { b = b / (double)S.lu[0]; }
    return b;
  }
  else
  {
#pragma unroll
    for (int k = 0; k < M - 1; k++)
    {
      double bk = __shfl(b, g.base + g.inv[k], 64);
      if (pos > k) { b = b - (double)S.lu[k] * bk; }
    }
#pragma unroll
    for (int k = M - 1; k > 0; k--)
    {
      if (pos == k) { b = b / (double)S.lu[k]; }
      double bk = __shfl(b, g.base + g.inv[k], 64);
      if (pos < k) { b = b - (double)S.lu[k] * bk; }
    }
  }
  if (pos == 0) { b = b / (double)S.lu[0]; }
  // variable i sits in the lane holding physical row i
  return __shfl(b, g.base + g.inv[g.valid ? g.i : 0], 64);
}

// ---- integrator pieces -----------------------------------------------------
template <int M, typename T>
__device__ __forceinline__ double g_ewt(const Grp<M, T>& g, double yi)
{
  double t = fabs(yi);
  t        = g.a->rtol * t;
  t        = t + g.a->atol;
  return 1.0 / t;
}

// The Nordsieck columns stay in registers only while every access uses a
// constant index.  A runtime column index is resolved by value: each column is
// passed through an empty asm so the optimiser cannot turn the choice back
// into an indexed load from the state struct (which would move the whole
// integrator state to scratch memory).
__device__ __forceinline__ double opaque(double x)
{
  asm("" : "+v"(x));
  return x;
}

template <int M, typename T>
__device__ __forceinline__ double zget(const Sys<M, T>& S, int j)
{
  double v = opaque(S.zn[0]);
#pragma unroll
  for (int k = 1; k <= QMAX; k++)
  {
    const double zk = opaque(S.zn[k]);
    v               = (j == k) ? zk : v;
  }
  return v;
}

template <int M, typename T>
__device__ __forceinline__ void zset(Sys<M, T>& S, int j, double v)
{
#pragma unroll
  for (int k = 0; k <= QMAX; k++) { S.zn[k] = (j == k) ? v : S.zn[k]; }
}

template <int M, typename T>
__device__ __forceinline__ void g_rescale(const Grp<M, T>& g, Sys<M, T>& S)
{
  double c = S.eta;
#pragma unroll
  for (int j = 1; j <= QMAX; j++)
  {
    if (j <= S.q) { S.zn[j] = c * S.zn[j]; }
    c = S.eta * c;
  }
  S.h      = S.hscale * S.eta;
  S.hscale = S.h;
}

template <int M, typename T>
__device__ __forceinline__ void g_increase_bdf(const Grp<M, T>& g, Sys<M, T>& S)
{
  double* l = g.sc->l;
  for (int i = 0; i <= QMAX; i++) { l[i] = 0.0; }
  double alpha0, alpha1, prod, xi, xiold, hsum, A1;
  l[2] = alpha1 = prod = xiold = 1.0;
  alpha0                       = -1.0;
  hsum                         = S.hscale;
  if (S.q > 1)
  {
    for (int j = 1; j < S.q; j++)
    {
      hsum += g.sc->tau[j + 1];
      xi = hsum / S.hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; i--) { l[i] = l[i] * xiold + l[i - 1]; }
      xiold = xi;
    }
  }
  A1         = (-alpha0 - alpha1) / prod;
  double zL  = A1 * zget(S, S.indx_acor);
  zset(S, S.L, zL);
#pragma unroll
  for (int j = 2; j <= QMAX; j++)
  {
    if (j <= S.q) { S.zn[j] = lsum(l[j], zL, 1.0, S.zn[j]); }
  }
}

template <int M, typename T>
__device__ __forceinline__ void g_decrease_bdf(const Grp<M, T>& g, Sys<M, T>& S)
{
  double* l = g.sc->l;
  for (int i = 0; i <= QMAX; i++) { l[i] = 0.0; }
  l[2]        = 1.0;
  double hsum = 0.0;
  for (int j = 1; j <= S.q - 2; j++)
  {
    hsum += g.sc->tau[j];
    double xi = hsum / S.hscale;
    for (int i = j + 2; i >= 2; i--) { l[i] = l[i] * xi + l[i - 1]; }
  }
  const double zq = zget(S, S.q);
#pragma unroll
  for (int j = 2; j <= QMAX; j++)
  {
    if (j < S.q) { S.zn[j] = lsum(-l[j], zq, 1.0, S.zn[j]); }
  }
}

template <int M, typename T>
__device__ __forceinline__ void g_adjust_order(const Grp<M, T>& g, Sys<M, T>& S, int deltaq)
{
  if ((S.q == 2) && (deltaq != 1)) { return; }
  if (deltaq == 1) { g_increase_bdf(g, S); }
  else if (deltaq == -1) { g_decrease_bdf(g, S); }
}

template <int M, typename T>
__device__ __forceinline__ void g_predict(Sys<M, T>& S)
{
  S.tn += S.h;
#pragma unroll
  for (int k = 1; k <= QMAX; k++)
  {
#pragma unroll
    for (int j = QMAX; j >= 1; j--)
    {
      if (k <= S.q && j >= k && j <= S.q) { S.zn[j - 1] = S.zn[j - 1] + S.zn[j]; }
    }
  }
}

template <int M, typename T>
__device__ __forceinline__ void g_restore(Sys<M, T>& S, double saved_t)
{
  S.tn = saved_t;
#pragma unroll
  for (int k = 1; k <= QMAX; k++)
  {
#pragma unroll
    for (int j = QMAX; j >= 1; j--)
    {
      if (k <= S.q && j >= k && j <= S.q) { S.zn[j - 1] = S.zn[j - 1] - S.zn[j]; }
    }
  }
}

// element j of a register array indexed at run time (0 <= j <= N), chosen by
// value (see opaque) so the array stays in registers
template <int N>
__device__ __forceinline__ double rsel(const double (&a)[N + 1], int j)
{
  double v = opaque(a[0]);
#pragma unroll
  for (int k = 1; k <= N; k++)
  {
    const double ak = opaque(a[k]);
    v               = (j == k) ? ak : v;
  }
  return v;
}

// cvSetBDF + cvSetTqBDF + the gamma update of cvSet.  The step history and
// the coefficients are computed in registers (loops unrolled to QMAX and
// predicated on q, the same operations in the same order) and only the
// results go to the group's LDS slots.
template <int M, typename T>
__device__ __forceinline__ void g_set(const Grp<M, T>& g, Sys<M, T>& S, long nst)
{
  double tau[QMAX + 1];
#pragma unroll
  for (int k = 0; k <= QMAX; k++) { tau[k] = g.sc->tau[k]; }
  double l[QMAX + 1];
  double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
  l[0] = l[1] = xi_inv = xistar_inv = 1.0;
#pragma unroll
  for (int i = 2; i <= QMAX; i++) { l[i] = 0.0; }
  alpha0 = alpha0_hat = -1.0;
  hsum                = S.h;
  if (S.q > 1)
  {
#pragma unroll
    for (int j = 2; j < QMAX; j++)
    {
      if (j < S.q)
      {
        hsum += tau[j - 1];
        xi_inv = S.h / hsum;
        alpha0 -= 1.0 / j;
#pragma unroll
        for (int i = j; i >= 1; i--) { l[i] += l[i - 1] * xi_inv; }
      }
    }
    alpha0 -= 1.0 / S.q;
    xistar_inv = -l[1] - alpha0;
    hsum += rsel<QMAX>(tau, S.q - 1);
    xi_inv     = S.h / hsum;
    alpha0_hat = -l[1] - xi_inv;
#pragma unroll
    for (int i = QMAX; i >= 1; i--)
    {
      if (i <= S.q) { l[i] += l[i - 1] * xistar_inv; }
    }
  }
  // cvSetTqBDF
  double* tq = g.sc->tq;
  const double lq = rsel<QMAX>(l, S.q);
  double A1 = 1.0 - alpha0_hat + alpha0;
  double A2 = 1.0 + S.q * A1;
  tq[2]     = fabs(A1 / (alpha0 * A2));
  tq[5]     = fabs(A2 * xistar_inv / (lq * xi_inv));
  if (S.qwait == 1)
  {
    if (S.q > 1)
    {
      double C     = xistar_inv / lq;
      double A3    = alpha0 + 1.0 / S.q;
      double A4    = alpha0_hat + xi_inv;
      double Cpinv = (1.0 - A4 + A3) / A3;
      tq[1]        = fabs(C * Cpinv);
    }
    else { tq[1] = 1.0; }
    hsum += rsel<QMAX>(tau, S.q);
    xi_inv        = S.h / hsum;
    double A5     = alpha0 - (1.0 / (S.q + 1));
    double A6     = alpha0_hat - xi_inv;
    double Cppinv = (1.0 - A6 + A5) / A2;
    tq[3]         = fabs(Cppinv / (xi_inv * (S.q + 2) * A5));
  }
  tq[4] = CORTES / tq[2];
#pragma unroll
  for (int i = 0; i <= QMAX; i++) { g.sc->l[i] = l[i]; }
  // cvSet
  S.rl1   = 1.0 / l[1];
  S.gamma = S.h * S.rl1;
  if (nst == 0) { S.gammap = S.gamma; }
  S.gamrat = (nst > 0) ? S.gamma / S.gammap : 1.0;
}

// ---- Newton pieces ------------------------------------------------------------
// residual: y = zn0 + ycor; ftemp = f(y); delta = rl1*zn1 + ycor - gamma*ftemp
template <int M, typename T>
__device__ __forceinline__ void g_residual(const Grp<M, T>& g, Sys<M, T>& S)
{
  S.y     = lsum(1.0, S.zn[0], 1.0, S.acor);
  S.ftemp = g_rhs(g, S.y);
  S.nfe++;
  double res = lsum(S.rl1, S.zn[1], 1.0, S.acor);
  S.delta    = lsum(-S.gamma, S.ftemp, 1.0, res);
}

// linear-solver setup (cvNlsLSetup + cvLsSetup + cvLsLinSys + GETRF)
template <int M, typename T>
__device__ __forceinline__ int g_lsetup(const Grp<M, T>& g, Sys<M, T>& S, bool jbad_nls)
{
  if (jbad_nls) { S.convfail = CV_FAIL_BAD_J; }
  double dgamma = fabs((S.gamma / S.gammap) - 1.0);
  bool ls_jbad  = (S.nst == 0) || (S.nst >= S.nstlj + MSBJ) ||
                 ((S.convfail == CV_FAIL_BAD_J) && (dgamma < DGMAX_JBAD)) ||
                 (S.convfail == CV_FAIL_OTHER);
  bool jcur;
  if (!ls_jbad) { jcur = false; } // saved J rebuilt from the state it was taken at
  else
  {
    jcur   = true;
    S.ysav = S.y;
  }
  // the Newton-matrix row is built in the factor registers themselves
  BDF_PH_PUSH(g, SUNAMD_PH_JAC);
  g_newton_matrix(g, S.ysav, S.gamma, S.lu);
#pragma unroll
  for (int j = 0; j < M; j++) { BDF_PH_FENCE(S.lu[j]); }
  if (jcur)
  {
    S.nje++;
    S.nstlj = S.nst;
  }
  BDF_PH_SET(g, SUNAMD_PH_GETRF);
  int fail = g_getrf(g, S);
#pragma unroll
  for (int j = 0; j < M; j++) { BDF_PH_FENCE(S.lu[j]); }
  BDF_PH_FENCE(fail);
  BDF_PH_POP(g);
  S.nsetups++;
  S.nls_jcur = jcur;
  S.gamrat   = 1.0;
  S.gammap   = S.gamma;
  S.crate    = 1.0;
  S.nstlp    = S.nst;
  return fail ? NLS_CONV_RECVR : CV_SUCCESS;
}

template <int M, typename T>
__device__ __forceinline__ void g_lsolve(const Grp<M, T>& g, Sys<M, T>& S)
{
  BDF_PH_PUSH(g, SUNAMD_PH_GETRS);
  double x = g_getrs(g, S, S.delta);
  BDF_PH_FENCE(x);
  BDF_PH_POP(g);
  S.delta  = x;
  if (S.gamrat != 1.0) { S.delta = (2.0 / (1.0 + S.gamrat)) * S.delta; }
}

template <int M, typename T>
__device__ __forceinline__ int g_ctest(const Grp<M, T>& g, Sys<M, T>& S, int m, double tol,
                                       bool nonzero_start)
{
  double del = g_wrms(g, S.delta, S.ewt);
  if (m > 0)
  {
    double r = del / S.delp;
    S.crate  = (CRDOWN * S.crate > r) ? CRDOWN * S.crate : r;
  }
  double dcon = del * ((1.0 < S.crate) ? 1.0 : S.crate) / tol;
  if (dcon <= 1.0)
  {
    // after one iteration from a zero start the correction is the update
    // itself; from a learned start its norm has to be taken
    S.acnrm    = (m == 0 && !nonzero_start) ? del : g_wrms(g, S.acor, S.ewt);
    S.acnrmcur = true;
    return CV_SUCCESS;
  }
  if ((m >= 1) && (del > RDIV * S.delp)) { return NLS_CONV_RECVR; }
  S.delp = del;
  return NLS_CONTINUE;
}

// have_res: the residual of the current correction is already in S.delta
// (and S.y, S.ftemp); nonzero_start: the correction starts from a guess
template <int M, typename T>
__device__ __forceinline__ int g_newton(const Grp<M, T>& g, Sys<M, T>& S, bool callLSetup,
                                        bool have_res, bool nonzero_start)
{
  bool jbad        = false;
  long niters      = 0;
  long nconvfails  = 0;
  int retval       = CV_SUCCESS;
  const double tol = g.sc->tq[4];
  for (;;)
  {
    int curiter = 0;
    if (!have_res) { g_residual(g, S); }
    have_res = false;
    if (callLSetup)
    {
      retval = g_lsetup(g, S, jbad);
      if (retval != CV_SUCCESS) { break; }
    }
    for (;;)
    {
      niters++;
      S.delta = -1.0 * S.delta;
      g_lsolve(g, S);
      S.acor = lsum(1.0, S.acor, 1.0, S.delta);
      retval = g_ctest(g, S, curiter, tol, nonzero_start);
      curiter++;
      if (retval == CV_SUCCESS)
      {
        S.nls_jcur = false;
        S.nni += niters;
        S.nnf += nconvfails;
        return CV_SUCCESS;
      }
      if (retval != NLS_CONTINUE) { break; }
      if (curiter >= NLS_MAXCOR)
      {
        retval = NLS_CONV_RECVR;
        break;
      }
      g_residual(g, S);
    }
    if ((retval > 0) && !S.nls_jcur)
    {
      nconvfails++;
      callLSetup    = true;
      jbad          = true;
      S.acor        = 0.0;
      nonzero_start = false;
      continue;
    }
    break;
  }
  nconvfails++;
  S.nni += niters;
  S.nnf += nconvfails;
  return retval;
}

// ---- learned initial iterate ------------------------------------------------
__device__ __forceinline__ float bf16_round(float f)
{
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) { return f; }
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}

__device__ __forceinline__ float bf16_at(const unsigned short* w, long i)
{
  return __uint_as_float(((unsigned int)w[i]) << 16);
}

// the group's MLP inputs (unrounded) into xs[0..143]; each index has exactly
// one writing lane
template <int M, typename T>
__device__ __forceinline__ void g_features(const Grp<M, T>& g, const Sys<M, T>& S, float* xs)
{
  constexpr int G = pow2ceil(M);
  gsync();
  if (g.valid)
  {
    double a = fabs(S.zn[0]);
    xs[g.i]      = (float)(log10(a > 1e-30 ? a : 1e-30) * 0.1);
    xs[65 + g.i] = S.ratef;
  }
  for (int j = g.i; j < 144; j += G)
  {
    const bool feat = (j < M) || (j >= 65 && j < 65 + M);
    if (j == 64) { xs[64] = (float)(log10(S.gamma) * 0.1); }
    else if (!feat) { xs[j] = 0.0f; }
  }
  gsync();
}

// bf16 bits of f, rounded to nearest even (the training kernels' rounding)
__device__ __forceinline__ unsigned short bf16_bits(float f)
{
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) { return (unsigned short)(u >> 16); }
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// the group's network inputs, rounded to bf16, into its row of the
// wavefront's input block (columns 144.. stay zero)
template <int M, typename T>
__device__ __forceinline__ void g_features_bf16(const Grp<M, T>& g, const Sys<M, T>& S,
                                                unsigned short* xrow)
{
  constexpr int G = pow2ceil(M);
  if (g.valid)
  {
    const double a   = fabs(S.zn[0]);
    xrow[g.i]        = bf16_bits((float)(log10(a > 1e-30 ? a : 1e-30) * 0.1));
    xrow[65 + g.i]   = bf16_bits(S.ratef);
  }
  for (int j = g.i; j < 144; j += G)
  {
    const bool feat = (j < M) || (j >= 65 && j < 65 + M);
    if (j == 64) { xrow[64] = bf16_bits((float)(log10(S.gamma) * 0.1)); }
    else if (!feat) { xrow[j] = 0; }
  }
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// out = act(in W^T + b) for the wavefront's 16 rows.  in: [16 x K] bf16 in
// LDS (row stride ls); W: [N x KW] bf16 row-major in global memory (k >= KW
// reads as zero); b: [N] FP32.  ACT: SiLU, rounded to bf16 into out_h (row
// stride lh); otherwise the FP32 sums into out_f ([16 x 64]).  One
// 16 x 16 x 32 bf16 MFMA per (16 output columns, 32-deep slice): lane l
// supplies row l%16 of `in` and row n0 + l%16 of W for the 8 consecutive k
// of its quarter l/16 (one 16-byte load each) and receives rows
// 4(l/16) .. 4(l/16)+3 of column n0 + l%16; FP32 accumulation.
template <int K, int N, int KW, bool ACT>
__device__ __forceinline__ void w_layer(const unsigned short* in, int ls, const unsigned short* W,
                                        const float* b, unsigned short* out_h, int lh,
                                        float* out_f)
{
  const int l = __lane_id();
  const int r = l & 15, q = l >> 4;
  // the lane's weight row and k offset; each slice's address is this plus a
  // constant (the zero-padded slice reads the row's start and is then
  // zeroed), so no per-slice address has to stay live across the layer
  const unsigned short* wl = W + r * KW + 8 * q;
#pragma unroll
  for (int n0 = 0; n0 < N; n0 += 16)
  {
    // all slices' operands are requested before the first product (the
    // scheduling barrier keeps the compiler from interleaving load, wait and
    // product slice by slice)
    constexpr int NS = K / 32;
    bf16x8_t av[NS], wv[NS];
#pragma unroll
    for (int t = 0; t < NS; t++)
    {
      const int k0 = 32 * t;
      const int k  = k0 + 8 * q;
      av[t]        = *reinterpret_cast<const bf16x8_t*>(in + r * ls + k);
      if (k0 + 32 <= KW)
      {
        wv[t] = *reinterpret_cast<const bf16x8_t*>(wl + n0 * KW + k0);
      }
      else
      {
        wv[t] = *reinterpret_cast<const bf16x8_t*>(wl + n0 * KW + (k < KW ? k0 : -8 * q));
        if (k >= KW) { wv[t] = bf16x8_t{}; }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < NS; t++) { acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], wv[t], acc, 0, 0, 0); }
    const float bias = b[n0 + r];
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
      const float z = acc[i] + bias;
      if constexpr (ACT) { out_h[(4 * q + i) * lh + n0 + r] = bf16_bits(z / (1.0f + expf(-z))); }
      else { out_f[(4 * q + i) * 64 + n0 + r] = z; }
    }
  }
}

// The network on the 16 input rows of X (a wavefront's MLP_WAVE_BYTES network
// area in LDS; called with all 64 lanes active): 144 inputs (zero-padded to
// 160) -> 128 -> 128 -> 64, bf16 weights read as MFMA operands straight from
// the (L2-resident) weight block W (W1 W2 W3, then the FP32 biases), FP32
// accumulation, SiLU on the hidden layers.  Returns the [16 x 64] FP32
// outputs, in the area of the first hidden layer.
__device__ __forceinline__ const float* w_forward16(const unsigned short* W, const float* b,
                                                    unsigned short* X)
{
  const unsigned short* W2 = W + 128 * 144;
  const unsigned short* W3 = W2 + 128 * 128;
  unsigned short* H1       = X + 16 * MLP_XS;
  unsigned short* H2       = H1 + 16 * MLP_HS;
  float* Y                 = reinterpret_cast<float*>(H1); // layer 1's outputs are dead by then
  gsync();
  w_layer<160, 128, 144, true>(X, MLP_XS, W, b, H1, MLP_HS, nullptr);
  gsync();
  w_layer<128, 128, 128, true>(H1, MLP_HS, W2, b + 128, H2, MLP_HS, nullptr);
  gsync();
  w_layer<128, 64, 128, false>(H2, MLP_HS, W3, b + 256, nullptr, 0, Y);
  gsync();
  return Y;
}

// The wavefront's network evaluation on the rows its groups wrote (a group
// that did not write its row ignores its output); this lane's output
// (species g.i of its group), in error-weight units.
template <int M, typename T>
__device__ __forceinline__ float w_network(const Grp<M, T>& g, bool want)
{
  const float* Y  = w_forward16(g.a->mlp_w, g.a->mlp_b, g.wx);
  const float out = (want && g.valid) ? Y[g.row * 64 + g.i] : 0.0f;
  gsync();
  return out;
}

// one training row: inputs, the converged correction in error-weight units,
// and the species mask
template <int M, typename T>
__device__ __forceinline__ void g_log_pair(const Grp<M, T>& g, const Sys<M, T>& S)
{
  constexpr int G      = pow2ceil(M);
  const BucketArgs& A  = *g.a;
  unsigned long long r = 0;
  if (g.i == 0) { r = atomicAdd(A.log_next, 1ULL); }
  r = __shfl(r, g.base, 64);
  if (r >= (unsigned long long)A.log_cap) { return; }
  g_features(g, S, g.mlp);
  float* x = A.log_x + (long)r * 144;
  for (int j = g.i; j < 144; j += G) { x[j] = g.mlp[j]; }
  float* t = A.log_t + (long)r * 64;
  float* w = A.log_m + (long)r * 64;
  for (int j = g.i; j < 64; j += G)
  {
    t[j] = (j == g.i && g.valid) ? (float)(S.acor * S.ewt) : 0.0f;
    w[j] = (j < M) ? 1.0f : 0.0f;
  }
}

// have_guess: the wavefront's network proposed guess (error-weight units) for
// this step's first attempt
template <int M, typename T>
__device__ __forceinline__ int g_nls(const Grp<M, T>& g, Sys<M, T>& S, int nflag, bool have_guess,
                                     double guess)
{
  S.convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES : CV_FAIL_OTHER;
  bool callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (S.nst == 0) ||
                   (S.nst >= S.nstlp + MSBP) || (fabs(S.gamrat - 1.0) > DGMAX_LSETUP);
  S.acor        = 0.0;
  bool have_res = false, learned = false;
  if (have_guess && nflag == FIRST_CALL && S.nst > 0)
  {
    // learned start, kept only if its residual is smaller than the zero
    // start's (the plain predictor is the fallback)
    BDF_PH_PUSH(g, SUNAMD_PH_LEARN);
    const double gs = guess / S.ewt;
    g_residual(g, S);
    const double r0  = g_wrms(g, S.delta, S.ewt);
    const double d0 = S.delta, y0 = S.y, f0 = S.ftemp;
    S.acor           = g.valid ? gs : 0.0;
    g_residual(g, S);
    const double r1 = g_wrms(g, S.delta, S.ewt);
    S.ltried++;
    if (r1 < r0)
    {
      learned = true;
      S.lused++;
    }
    else
    {
      S.acor  = 0.0;
      S.delta = d0;
      S.y     = y0;
      S.ftemp = f0;
    }
    have_res = true;
    BDF_PH_POP(g);
  }
  int flag = g_newton(g, S, callSetup, have_res, learned);
  if (flag != CV_SUCCESS) { return flag; }
  if (g.learn && g.a->log_x != nullptr && nflag == FIRST_CALL && (S.nst % g.a->log_stride) == 0)
  {
    g_log_pair(g, S);
  }
  S.y = lsum(1.0, S.zn[0], 1.0, S.acor);
  if (!S.acnrmcur) { S.acnrm = g_wrms(g, S.acor, S.ewt); }
  return flag;
}

template <int M, typename T>
__device__ __forceinline__ int g_handle_nflag(const Grp<M, T>& g, Sys<M, T>& S, int& nflag, double saved_t,
                              int& ncf)
{
  if (nflag == CV_SUCCESS) { return DO_ERROR_TEST; }
  S.ncfn++;
  g_restore(S, saved_t);
  if (nflag < 0) { return CV_LSETUP_FAIL; }
  ncf++;
  S.etamax = 1.0;
  if ((fabs(S.h) <= 0.0 * ONEPSM) || (ncf == MXNCF))
  {
    if (nflag == NLS_CONV_RECVR) { return CV_CONV_FAILURE; }
  }
  S.eta = (ETA_CF > 0.0) ? ETA_CF : 0.0;
  nflag = PREV_CONV_FAIL;
  g_rescale(g, S);
  return PREDICT_AGAIN;
}

// x^y of the eta formulas (SUNRpowerR)
template <int M, typename T>
__device__ __forceinline__ double g_pow(const Grp<M, T>& g, double x, double y)
{
  BDF_PH_PUSH(g, SUNAMD_PH_POW);
  const double r = kFastFactor<T> ? fast_pow(x, y) : sunamd_pow(x, y);
  BDF_PH_FENCE(r);
  BDF_PH_POP(g);
  return r;
}

template <int M, typename T>
__device__ __forceinline__ int g_error_test(const Grp<M, T>& g, Sys<M, T>& S, int& nflag, double saved_t,
                            int& nef, double& dsm_out)
{
  double dsm = S.acnrm * g.sc->tq[2];
  dsm_out    = dsm;
  if (dsm <= 1.0) { return CV_SUCCESS; }
  nef++;
  S.netf++;
  nflag = PREV_ERR_FAIL;
  g_restore(S, saved_t);
  if ((fabs(S.h) <= 0.0 * ONEPSM) || (nef == MXNEF)) { return CV_ERR_FAILURE; }
  S.etamax = 1.0;
  if (nef <= MXNEF1)
  {
    S.eta = 1.0 / (g_pow(g, BIAS2 * dsm, 1.0 / S.L) + ADDON);
    double e2 = (S.eta > 0.0) ? S.eta : 0.0;
    S.eta     = (ETA_MIN_EF > e2) ? ETA_MIN_EF : e2;
    if (nef >= SMALL_NEF) { S.eta = (S.eta < ETA_MAX_EF) ? S.eta : ETA_MAX_EF; }
    g_rescale(g, S);
    return TRY_AGAIN;
  }
  if (S.q > 1)
  {
    S.eta = (ETA_MIN_EF > 0.0) ? ETA_MIN_EF : 0.0;
    g_adjust_order(g, S, -1);
    S.L = S.q;
    S.q--;
    S.qwait = S.L;
    g_rescale(g, S);
    return TRY_AGAIN;
  }
  S.eta    = (ETA_MIN_EF > 0.0) ? ETA_MIN_EF : 0.0;
  S.h *= S.eta;
  S.hscale = S.h;
  S.qwait  = LONG_WAIT;
  S.tempv  = g_rhs(g, S.zn[0]);
  S.nfe++;
  S.zn[1] = S.h * S.tempv;
  return TRY_AGAIN;
}

template <int M, typename T>
__device__ __forceinline__ void g_complete_step(const Grp<M, T>& g, Sys<M, T>& S)
{
  S.nst++;
  S.hu       = S.h;
  S.qu       = S.q;
  // step-size history shift from register copies of the old values
  double* tau = g.sc->tau;
  double t[QMAX + 1];
#pragma unroll
  for (int k = 0; k <= QMAX; k++) { t[k] = tau[k]; }
#pragma unroll
  for (int i = QMAX; i >= 2; i--)
  {
    if (i <= S.q) { tau[i] = t[i - 1]; }
  }
  if ((S.q == 1) && (S.nst > 1)) { tau[2] = t[1]; }
  tau[1] = S.h;
  const double* l = g.sc->l;
#pragma unroll
  for (int j = 0; j <= QMAX; j++)
  {
    if (j <= S.q) { S.zn[j] = lsum(l[j], S.acor, 1.0, S.zn[j]); }
  }
  S.qwait--;
  if ((S.qwait == 1) && (S.q != QMAX))
  {
    S.zn[QMAX]  = S.acor;
    S.saved_tq5 = g.sc->tq[5];
    S.indx_acor = QMAX;
  }
}

template <int M, typename T>
__device__ __forceinline__ void g_set_eta(Sys<M, T>& S)
{
  if ((S.eta > ETA_MIN_FX) && (S.eta < ETA_MAX_FX))
  {
    S.eta    = 1.0;
    S.hprime = S.h;
  }
  else
  {
    if (S.eta >= ETA_MAX_FX)
    {
      S.eta = (S.eta < S.etamax) ? S.eta : S.etamax;
      double d = fabs(S.h) * 0.0 * S.eta;
      S.eta /= (1.0 > d) ? 1.0 : d;
    }
    else
    {
      S.eta = (S.eta > ETA_MIN) ? S.eta : ETA_MIN;
      S.eta = (S.eta > 0.0) ? S.eta : 0.0;
    }
    S.hprime = S.h * S.eta;
  }
}

template <int M, typename T>
__device__ __forceinline__ void g_prepare_next(const Grp<M, T>& g, Sys<M, T>& S, double dsm)
{
  if (S.etamax == 1.0)
  {
    S.qwait  = (S.qwait > 2) ? S.qwait : 2;
    S.qprime = S.q;
    S.hprime = S.h;
    S.eta    = 1.0;
    return;
  }
  S.etaq = 1.0 / (g_pow(g, BIAS2 * dsm, 1.0 / S.L) + ADDON);
  if (S.qwait != 0)
  {
    S.eta    = S.etaq;
    S.qprime = S.q;
    g_set_eta(S);
    return;
  }
  S.qwait = 2;
  // cvComputeEtaqm1
  S.etaqm1 = 0.0;
  if (S.q > 1)
  {
    double ddn = g_wrms(g, zget(S, S.q), S.ewt) * g.sc->tq[1];
    S.etaqm1   = 1.0 / (g_pow(g, BIAS1 * ddn, 1.0 / S.q) + ADDON);
  }
  // cvComputeEtaqp1
  S.etaqp1 = 0.0;
  if (S.q != QMAX && S.saved_tq5 != 0.0)
  {
    double hr = S.h / g.sc->tau[2];
    double pw = 1.0;
    for (int i = 1; i <= S.L; i++) { pw *= hr; } // SUNRpowerI
    double cquot = (g.sc->tq[5] / S.saved_tq5) * pw;
    S.tempv      = lsum(-cquot, S.zn[QMAX], 1.0, S.acor);
    double dup   = g_wrms(g, S.tempv, S.ewt) * g.sc->tq[3];
    S.etaqp1     = 1.0 / (g_pow(g, BIAS3 * dup, 1.0 / (S.L + 1)) + ADDON);
  }
  // cvChooseEta
  double em   = (S.etaq > S.etaqp1) ? S.etaq : S.etaqp1;
  double etam = (S.etaqm1 > em) ? S.etaqm1 : em;
  if ((etam > ETA_MIN_FX) && (etam < ETA_MAX_FX))
  {
    S.eta    = 1.0;
    S.qprime = S.q;
  }
  else if (etam == S.etaq)
  {
    S.eta    = S.etaq;
    S.qprime = S.q;
  }
  else if (etam == S.etaqm1)
  {
    S.eta    = S.etaqm1;
    S.qprime = S.q - 1;
  }
  else
  {
    S.eta      = S.etaqp1;
    S.qprime   = S.q + 1;
    S.zn[QMAX] = S.acor;
  }
  g_set_eta(S);
}

// initial step (cvHin with cvUpperBoundH0 and cvYddNorm)
template <int M, typename T>
__device__ __forceinline__ int g_hin(const Grp<M, T>& g, Sys<M, T>& S, double tout)
{
  double tdiff = tout - S.tn;
  if (tdiff == 0.0) { return CV_TOO_CLOSE; }
  int sign      = (tdiff > 0.0) ? 1 : -1;
  double tdist  = fabs(tdiff);
  double tround = UROUND * ((fabs(S.tn) > fabs(tout)) ? fabs(S.tn) : fabs(tout));
  if (tdist < 2.0 * tround) { return CV_TOO_CLOSE; }
  double hlb = HLB_FACTOR * tround;
  // cvUpperBoundH0
  double temp2  = fabs(S.zn[0]);
  double temp1  = g_ewt(g, S.zn[0]);
  temp1         = 1.0 / temp1;
  temp1         = lsum(HUB_FACTOR, temp2, 1.0, temp1);
  temp2         = fabs(S.zn[1]);
  temp1         = temp2 / temp1;
  double hubinv = g_maxabs(g, temp1);
  double hub    = HUB_FACTOR * tdist;
  if (hub * hubinv > 1.0) { hub = 1.0 / hubinv; }
  double hg = sqrt0(hlb * hub);
  if (hub < hlb)
  {
    S.h = (sign == -1) ? -hg : hg;
    return CV_SUCCESS;
  }
  double hs = hg, hnew = 0.0, yddnrm = 0.0;
  for (int count1 = 1; count1 <= MAX_ITERS; count1++)
  {
    // cvYddNorm (the RHS never fails here)
    double hgs = hg * sign;
    S.y        = lsum(hgs, S.zn[1], 1.0, S.zn[0]);
    S.tempv    = g_rhs(g, S.y);
    S.nfe++;
    S.tempv = lsum(1.0 / hgs, S.tempv, -1.0 / hgs, S.zn[1]);
    yddnrm  = g_wrms(g, S.tempv, S.ewt);
    hs      = hg;
    hnew    = (yddnrm * hub * hub > 2.0) ? sqrt0(2.0 / yddnrm) : sqrt0(hg * hub);
    if (count1 == MAX_ITERS) { break; }
    double hrat = hnew / hg;
    if ((hrat > 0.5) && (hrat < 2.0)) { break; }
    if ((count1 > 1) && (hrat > 2.0))
    {
      hnew = hg;
      break;
    }
    hg = hnew;
  }
  (void)hs;
  double h0 = H_BIAS * hnew;
  if (h0 < hlb) { h0 = hlb; }
  if (h0 > hub) { h0 = hub; }
  if (sign == -1) { h0 = -h0; }
  S.h = h0;
  return CV_SUCCESS;
}

// results of a finished system: state at tout (or where it stopped) and counters
template <int M, typename T>
__device__ __forceinline__ void g_store(const Grp<M, T>& g, const Sys<M, T>& S, double yout, int flag)
{
  const BucketArgs& A = *g.a;
  if (g.valid) { A.yout[g.cell * M + g.i] = yout; }
  if (g.i == 0)
  {
    long* st              = A.stats + g.cell * SUNAMD_BS_COUNT;
    st[SUNAMD_BS_NST]     = S.nst;
    st[SUNAMD_BS_NFE]     = S.nfe;
    st[SUNAMD_BS_NSETUPS] = S.nsetups;
    st[SUNAMD_BS_NETF]    = S.netf;
    st[SUNAMD_BS_NNI]     = S.nni;
    st[SUNAMD_BS_NNF]     = S.nnf;
    st[SUNAMD_BS_NCFN]    = S.ncfn;
    st[SUNAMD_BS_NJE]     = S.nje;
    st[SUNAMD_BS_QLAST]   = S.qu;
    st[SUNAMD_BS_FLAG]    = flag;
    if (A.lstats != nullptr && S.ltried > 0)
    {
      atomicAdd(A.lstats, (unsigned long long)S.ltried);
      atomicAdd(A.lstats + 1, (unsigned long long)S.lused);
    }
  }
}

// Take the next system from the bucket's queue and run CVode's first-call
// work on it (CVodeInit state, error weights, f(t0,y0), initial step).
// Returns false once the queue is empty.
template <int M, typename T>
__device__ __forceinline__ bool g_acquire(Grp<M, T>& g, Sys<M, T>& S, double tout)
{
  const BucketArgs& A = *g.a;
  for (;;)
  {
    unsigned long long ticket = 0;
    if (g.i == 0) { ticket = atomicAdd(A.next, 1ULL); }
    ticket = __shfl(ticket, g.base, 64);
    if (ticket >= (unsigned long long)A.ncells) { return false; }
    g.cell = (long)ticket;
    {
      // this system's rate constants into the group's LDS slot
      constexpr int G = pow2ceil(M);
      gsync();
      const double* kf = A.kf + g.cell * A.R;
      const double* kb = A.kb + g.cell * A.R;
      for (int r = g.i; r < A.R; r += G)
      {
        g.rates[r]       = kf[r];
        g.rates[A.R + r] = kb[r];
      }
      gsync();
    }

    S.zn[0] = g.valid ? A.y0[g.cell * M + g.i] : 0.0;
#pragma unroll
    for (int j = 1; j <= QMAX; j++) { S.zn[j] = 0.0; }
    S.acor = S.tempv = S.ftemp = S.y = S.delta = S.ysav = 0.0;
    S.tn       = 0.0;
    S.h        = 0.0;
    S.q        = 1;
    S.L        = 2;
    S.qwait    = S.L;
    S.etamax   = ETA_MAX_FS;
    S.qu       = 0;
    S.hu       = 0.0;
    S.nst = S.nfe = S.ncfn = S.netf = S.nni = S.nnf = S.nsetups = S.nstlp = 0;
    S.nje = S.nstlj = 0;
    S.saved_tq5 = 0.0;
    S.indx_acor = 0;
    S.acnrmcur  = false;
    S.nls_jcur  = false;
    S.crate = S.delp = S.acnrm = S.gamma = S.gammap = S.gamrat = S.rl1 = 0.0;
    S.etaq = S.etaqm1 = S.etaqp1 = S.eta = S.hprime = 0.0;
    S.qprime   = 1;
    S.convfail = CV_NO_FAILURES;
    S.pos      = g.i;
    S.ltried = S.lused = 0;
    {
      // learned-predictor input: log10(1 + sum of k_f + k_b over the
      // reactions touching this species) / 10
      // (each touching reaction once, however many roles the species has)
      double ks = 0.0;
      if (g.valid)
      {
        int last = -1;
        for (int e = g.soff[g.i]; e < g.soff[g.i + 1]; e++)
        {
          const int r = g.sent[e] >> 2;
          if (r != last) { ks += g.kf[r] + g.kb[r]; }
          last = r;
        }
      }
      S.ratef = g.valid ? (float)(log10(1.0 + ks) * 0.1) : 0.0f;
    }
    for (int k = 0; k < QMAX + 2; k++) { g.sc->tau[k] = 0.0; }
    if (A.bad != nullptr && *A.bad != 0)
    {
      // the reaction table failed validation (CV_ILL_INPUT): nothing is integrated
      g_store(g, S, S.zn[0], -22);
      continue;
    }

    S.ewt   = g_ewt(g, S.zn[0]);
    S.zn[1] = g_rhs(g, S.zn[0]);
    S.nfe++;
    int flag = g_hin(g, S, tout);
    if (flag == CV_SUCCESS)
    {
      S.hscale = S.h;
      S.h0u    = S.h;
      S.hprime = S.h;
      S.zn[1]  = S.h * S.zn[1];
      return true;
    }
    g_store(g, S, S.zn[0], flag); // initial-step failure: y is still y0
  }
}

// what the two halves of one pass of CVode's main loop share
struct StepCtx
{
  double saved_t, dsm;
  int ncf, nef, nflag;
};

// One pass of CVode's main loop (CV_NORMAL), first half: error weights, the
// work and accuracy guards, then cvStep up to its first prediction -- the
// order / step-size change, the Nordsieck predictor and the BDF coefficients.
// Returns true when the system is finished (yout and flag set).
template <int M, typename T>
__device__ __forceinline__ bool g_advance_head(const Grp<M, T>& g, Sys<M, T>& S, double& yout,
                                               int& flag, StepCtx& c)
{
  BDF_PH_SET(g, SUNAMD_PH_GUARD);
  if (S.nst > 0) { S.ewt = g_ewt(g, S.zn[0]); }
  // one CVode call: the steps taken in it are nst; mxstep < 0 disables the check
  if ((g.a->mxstep > 0) && (S.nst >= g.a->mxstep))
  {
    flag = CV_TOO_MUCH_WORK;
    yout = S.zn[0];
    return true;
  }
  double nrm = g_wrms(g, S.zn[0], S.ewt);
  if (UROUND * nrm > 1.0)
  {
    flag = CV_TOO_MUCH_ACC;
    yout = S.zn[0];
    return true;
  }
  c.ncf = c.nef = 0;
  c.dsm         = 0.0;
  BDF_PH_SET(g, SUNAMD_PH_PREDICT);
  if ((S.nst > 0) && (S.hprime != S.h))
  {
    if (S.qprime != S.q)
    {
      g_adjust_order(g, S, S.qprime - S.q);
      S.q     = S.qprime;
      S.L     = S.q + 1;
      S.qwait = S.L;
    }
    g_rescale(g, S);
  }
  c.saved_t = S.tn;
  c.nflag   = FIRST_CALL;
  g_predict(S);
  g_set(g, S, S.nst);
  return false;
}

// Second half: the nonlinear solve (from the learned guess when one is
// given), convergence-failure handling, the error test and the retries
// (each a new prediction), the completed step with the next step size and
// order, and the tout check with interpolation.
template <int M, typename T>
__device__ __forceinline__ bool g_advance_tail(const Grp<M, T>& g, Sys<M, T>& S, double tout,
                                               double& yout, int& flag, StepCtx& c,
                                               bool have_guess, double guess)
{
  int kflag  = CV_SUCCESS;
  bool again = false;
  for (;;)
  {
    if (again)
    {
      BDF_PH_SET(g, SUNAMD_PH_PREDICT);
      g_predict(S);
      g_set(g, S, S.nst);
    }
    again = true;
    BDF_PH_SET(g, SUNAMD_PH_NEWTON);
    c.nflag = g_nls(g, S, c.nflag, have_guess, guess);
    BDF_PH_SET(g, SUNAMD_PH_ERRTEST);
    kflag = g_handle_nflag(g, S, c.nflag, c.saved_t, c.ncf);
    if (kflag == PREDICT_AGAIN) { continue; }
    if (kflag != DO_ERROR_TEST) { break; }
    kflag = g_error_test(g, S, c.nflag, c.saved_t, c.nef, c.dsm);
    if (kflag == TRY_AGAIN) { continue; }
    break;
  }
  if (kflag != CV_SUCCESS)
  {
    BDF_PH_SET(g, SUNAMD_PH_OUTPUT);
    flag = kflag;
    yout = S.zn[0];
    return true;
  }
  BDF_PH_SET(g, SUNAMD_PH_COMPLETE);
  g_complete_step(g, S);
  BDF_PH_SET(g, SUNAMD_PH_PREPARE);
  g_prepare_next(g, S, c.dsm);
  S.etamax = (S.nst <= SMALL_NST) ? ETA_MAX_ES : ETA_MAX_GS;
  S.acor   = g.sc->tq[2] * S.acor;
  BDF_PH_SET(g, SUNAMD_PH_OUTPUT);
  if ((S.tn - tout) * S.h >= 0.0)
  {
    // CVodeGetDky(tout, 0) through the unfused LinearCombination:
    // z = c_q zn[q], then z = c_j zn[j] + z for j = q-1 .. 0, c_j = s^j
    double s = (tout - S.tn) / S.h;
    double z = 0.0;
    for (int j = S.q; j >= 0; j--)
    {
      double cj = 1.0;
      for (int i = 0; i < j; i++) { cj *= s; }
      double zj = zget(S, j);
      z         = (j == S.q) ? cj * zj : lsum(cj, zj, 1.0, z);
    }
    yout = z;
    flag = CV_SUCCESS;
    return true;
  }
  return false;
}

// One wavefront runs 64/G groups.  A group that finishes its system takes
// the next one from the queue between steps, so the wavefront's groups stay
// busy while their neighbours' systems need more steps.
// W: minimum wavefronts per SIMD the register allocation must allow (1 = no
// limit; 2, 3 and 4 cap it at 256, 168 and 128 registers, spilling the rest).
// LEARN: the learned start and the training log are compiled in; without them
// the kernel needs fewer registers (no spills at 2 wavefronts per SIMD for
// most sizes).
template <int M, typename T, int W, bool LEARN>
__global__ void __launch_bounds__(64, W) k_batched_bdf(BucketArgs A)
{
  constexpr int G   = pow2ceil(M);
  constexpr int GPW = 64 / G;
  __shared__ double s_red[64];
  __shared__ T s_prow[64];
  __shared__ int s_inv[64];
  __shared__ GroupScalars s_sc[GPW];
  // dynamic LDS: per-group rate constants (GPW x 2R doubles), the network
  // tables (4R + 4R + M+1 ints), per-group Jacobian rows / reaction rates
  // (GPW x rows_stride doubles), then the wavefront's network area
  // (MLP_WAVE_BYTES) with the learned start, then GPW x MLP_LOG_FLOATS with
  // the training log; 16-byte alignment for the int4 reads of the reaction
  // table and the network's 16-byte operand reads
  extern __shared__ __align__(16) double s_dyn[];
  const long JS   = group_rows_stride(M, A.R, (int)sizeof(T));
  double* s_rates = s_dyn;
  int* s_idx      = (int*)(s_dyn + (long)GPW * 2 * A.R);
  int* s_sent     = s_idx + 4 * A.R;
  int* s_soff     = s_sent + 4 * A.R;
  double* s_jrow  = (double*)(s_soff + ((M + 2) & ~1));
  const bool learned = LEARN && A.mlp_w != nullptr;
  unsigned short* s_wx = (unsigned short*)(s_jrow + (long)GPW * JS);
  float* s_log         = (float*)((char*)s_wx + (learned ? MLP_WAVE_BYTES : 0));

  const int lane = threadIdx.x;
  for (int t = lane; t < 4 * A.R; t += 64)
  {
    s_idx[t]  = A.idx[t];
    s_sent[t] = A.sent[t];
  }
  for (int t = lane; t <= M; t += 64) { s_soff[t] = A.soff[t]; }
  __syncthreads();
  const int grp  = lane / G;
  Grp<M, T> g;
  g.i     = lane % G;
  g.base  = grp * G;
  g.valid = g.i < M;
  g.red   = s_red + g.base;
  g.prow  = s_prow + g.base;
  g.inv   = s_inv + g.base;
  g.sc    = &s_sc[grp];
  g.mlp   = (LEARN && A.log_x != nullptr) ? s_log + grp * MLP_LOG_FLOATS : nullptr;
  g.wx    = learned ? s_wx : nullptr;
  g.row   = grp;
  if (learned)
  {
    for (int t = lane; t < MLP_WAVE_BYTES / 2; t += 64) { s_wx[t] = 0; }
    __syncthreads();
  }
  g.learn = LEARN;
  g.a     = &A;
  g.idx   = s_idx;
  g.soff  = s_soff;
  g.sent  = s_sent;
  g.rates = s_rates + (long)grp * 2 * A.R;
  g.jrow  = (T*)(s_jrow + (long)grp * JS) + (long)(g.valid ? g.i : 0) * (M + 1);
  g.wbuf  = s_jrow + (long)grp * JS;
  g.cell  = 0;
  // the rate constants are always read from the group's LDS slot (a pointer
  // with a single, LDS, origin keeps every access a ds_read)
  g.kf    = g.rates;
  g.kb    = g.rates + A.R;
  Sys<M, T> S;
#ifdef SUNAMD_BDF_PHASES
  constexpr int CAL_MARKS = 32;
  if (g.i == 0)
  {
    for (int p = 0; p < SUNAMD_BDF_NPHASES; p++)
    {
      g.sc->ph_acc[p] = 0;
      g.sc->ph_cnt[p] = 0;
    }
    g.sc->ph_cur = SUNAMD_PH_IDLE;
    g.sc->ph_t   = __builtin_amdgcn_s_memtime();
  }
  gsync();
  // calibration: CAL_MARKS back-to-back changes, timed into the idle slot
  for (int c = 0; c < CAL_MARKS; c++) { ph_set(g, SUNAMD_PH_IDLE); }
  unsigned long long cal_ticks = 0;
  if (g.i == 0)
  {
    cal_ticks           = g.sc->ph_acc[SUNAMD_PH_IDLE];
    g.sc->ph_acc[SUNAMD_PH_IDLE] = 0;
    g.sc->ph_cnt[SUNAMD_PH_IDLE] = 0;
  }
  ph_set(g, SUNAMD_PH_ACQUIRE);
#endif

  bool busy = g_acquire(g, S, A.tout);
  if (!busy) { BDF_PH_SET(g, SUNAMD_PH_IDLE); }
  if (!learned)
  {
    while (busy)
    {
      double yout = 0.0;
      int flag    = 0;
      StepCtx c;
      bool fin = g_advance_head(g, S, yout, flag, c);
      if (!fin) { fin = g_advance_tail(g, S, A.tout, yout, flag, c, false, 0.0); }
      if (fin)
      {
        g_store(g, S, yout, flag);
        BDF_PH_SET(g, SUNAMD_PH_ACQUIRE);
        busy = g_acquire(g, S, A.tout);
        // a group without a system waits for its wavefront's other groups
        if (!busy) { BDF_PH_SET(g, SUNAMD_PH_IDLE); }
      }
    }
  }
  else
  {
    // the network runs once per pass for the whole wavefront, between the
    // halves of every group's pass: the loop is wavefront-uniform, so all 64
    // lanes are active at the evaluation (its MFMA operands come from every
    // lane)
    while (__any(busy))
    {
      double yout = 0.0;
      int flag    = 0;
      StepCtx c;
      bool fin = false, want = false;
      if (busy)
      {
        fin  = g_advance_head(g, S, yout, flag, c);
        want = !fin && S.nst > 0;
        if (want) { g_features_bf16(g, S, g.wx + g.row * MLP_XS); }
      }
      double guess = 0.0;
      if (__any(want))
      {
        BDF_PH_PUSH(g, SUNAMD_PH_LEARN);
        guess = (double)w_network(g, want);
        BDF_PH_POP(g);
      }
      if (busy && !fin) { fin = g_advance_tail(g, S, A.tout, yout, flag, c, want, guess); }
      if (busy && fin)
      {
        g_store(g, S, yout, flag);
        BDF_PH_SET(g, SUNAMD_PH_ACQUIRE);
        busy = g_acquire(g, S, A.tout);
        if (!busy) { BDF_PH_SET(g, SUNAMD_PH_IDLE); }
      }
    }
  }
#ifdef SUNAMD_BDF_PHASES
  BDF_PH_SET(g, SUNAMD_PH_IDLE);
  if (g.i == 0 && A.phase != nullptr)
  {
    for (int p = 0; p < SUNAMD_BDF_NPHASES; p++)
    {
      atomicAdd(A.phase + p, g.sc->ph_acc[p]);
      atomicAdd(A.phase + SUNAMD_BDF_NPHASES + p, (unsigned long long)g.sc->ph_cnt[p]);
    }
    atomicAdd(A.phase + 2 * SUNAMD_BDF_NPHASES, cal_ticks);
    atomicAdd(A.phase + 2 * SUNAMD_BDF_NPHASES + 1, (unsigned long long)CAL_MARKS);
  }
#endif
}


// launcher, instantiated per (m, precision) in the bdf_inst_*.hip units
template <int M, typename T>
void launch(const BucketArgs& A, int grid, hipStream_t stream, int waves)
{
  constexpr int GPW = 64 / pow2ceil(M);
  const bool learn  = A.mlp_w != nullptr || A.log_x != nullptr;
  const size_t lds  = bucket_lds_bytes(M, A.R, GPW, A.mlp_w != nullptr, A.log_x != nullptr,
                                       (int)sizeof(T));
#ifdef SUNAMD_BDF_PHASES
  // instrumented build: the two-wavefront register budget only
  (void)waves;
  if (learn) { k_batched_bdf<M, T, 2, true><<<grid, 64, lds, stream>>>(A); }
  else { k_batched_bdf<M, T, 2, false><<<grid, 64, lds, stream>>>(A); }
#else
  if (waves >= 4) { k_batched_bdf<M, T, 4, true><<<grid, 64, lds, stream>>>(A); }
  else if (waves == 3 && !learn) { k_batched_bdf<M, T, 3, false><<<grid, 64, lds, stream>>>(A); }
  else if (waves >= 2)
  {
    if (learn) { k_batched_bdf<M, T, 2, true><<<grid, 64, lds, stream>>>(A); }
    else { k_batched_bdf<M, T, 2, false><<<grid, 64, lds, stream>>>(A); }
  }
  else
  {
    if (learn) { k_batched_bdf<M, T, 1, true><<<grid, 64, lds, stream>>>(A); }
    else { k_batched_bdf<M, T, 1, false><<<grid, 64, lds, stream>>>(A); }
  }
#endif
}

} // namespace sunamd_bdf

#define SUNAMD_BDF_INSTANTIATE(MM)                                                              \
  template void sunamd_bdf::launch<MM, double>(const sunamd_bdf::BucketArgs&, int, hipStream_t,  \
                                               int);                                            \
  template void sunamd_bdf::launch<MM, float>(const sunamd_bdf::BucketArgs&, int, hipStream_t, int);

#endif
