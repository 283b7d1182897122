// -----------------------------------------------------------------------------
// Device-stepped coupled BDF: one cooperative launch integrates a coupled
// mass-action system (ncells cells of m species in one state vector,
// block-diagonal Newton matrix) from t = 0 to tout.  Every decision of the
// integrator -- Newton convergence, Jacobian reuse, error test, step size
// and order -- is taken on the device; the host launches once and reads the
// result.  The algorithm is the one the host integrator (core/cvode_bdf.c)
// runs over the N_Vector and block-LU drop-ins (cvode.c cvStep / cvNls /
// cvDoErrorTest / cvCompleteStep / cvPrepareNextStep / cvHin, cvode_nls.c,
// cvode_ls.c, sunnonlinsol_newton.c), with the same per-element roundings:
//
//   - every workgroup holds a copy of the integrator's scalars and computes
//     the same updates from the same inputs, so the copies stay identical;
//   - the state (Nordsieck array, error weights, correction) lives in HBM;
//     each wavefront owns a contiguous range of cells and streams them
//     through one fused pass per Newton iteration (predictor and history
//     updates, residual with the right-hand side, Newton matrix and LU when
//     a setup is due, solve, update, norm terms) and one per accepted step
//     (history update, the order-selection norms, the next error weights);
//   - a WRMS norm is each workgroup's partial sum in a fixed order, then a
//     grid barrier, then every workgroup adds the partials in workgroup
//     order -- the same bits everywhere; in sequential mode the terms are
//     kept and added in element order, the serial vector's order, which
//     makes the run bit-identical to the host integrator over the serial
//     vector (the oracle);
//   - per cell, the right-hand side, the Newton matrix, the register-row LU
//     and the solve are the batched integrator's (bdf_device.h), bit-identical
//     to the dense GETRF / GETRS and the CPU model in FP64;
//   - the saved Jacobian is kept as the state it was evaluated at (ysav) and
//     rebuilt at each setup that reuses it (the same values as a stored copy).
// -----------------------------------------------------------------------------
#include "bdf_device.h"
#include "coupled_device.h"
#include "sunamd/sunmatrix_mi355block.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace sunamd_cpl
{
using namespace sunamd_bdf;

constexpr int NRED = 4; // per reduction: three sums and one maximum

// wavefronts per workgroup: the per-group LDS areas (rate constants, the
// Newton-matrix rows) of a workgroup stay within 64 KiB for R <= 2m
template <int M>
__host__ __device__ constexpr int cpl_waves()
{
  return M <= 16 ? 4 : (M <= 32 ? 2 : 1);
}
constexpr int CW_MAX = 4;

// LDS of a wavefront.  With FP64 rows, groups of 32 lanes (two per
// wavefront) share one Newton-matrix rows area, used only by setup passes,
// and build their matrices one after the other: the 8.4 KiB a group's rows
// take would otherwise leave room for three workgroups per CU instead of
// four.  Every group has its own rate constants and reaction-rate buffer.
// Other cases keep one rows area per group (which the reaction rates alias);
// FP32 rows are half the size and fit four workgroups either way (sharing
// measured slower there).  In doubles:
template <int M, typename T>
__host__ __device__ constexpr bool cpl_shared_rows()
{
  return pow2ceil(M) == 32 && sizeof(T) == 8;
}
template <int M, typename T>
__host__ __device__ inline long cpl_rows_doubles(int R)
{
  return ((long)M * (M + 1) * (long)sizeof(T) + 7) / 8;
}
template <int M, typename T>
__host__ __device__ inline long cpl_group_doubles(int R)
{
  return cpl_shared_rows<M, T>() ? 3L * R : 2L * R + group_rows_stride(M, R, (int)sizeof(T));
}
template <int M, typename T>
__host__ __device__ inline long cpl_wave_doubles(int R)
{
  constexpr int GPW = 64 / pow2ceil(M);
  return GPW * cpl_group_doubles<M, T>(R) + (cpl_shared_rows<M, T>() ? cpl_rows_doubles<M, T>(R) : 0);
}

// results written by workgroup 0 at the end
enum Out
{
  O_NST,
  O_NFE,
  O_NSETUPS,
  O_NETF,
  O_NNI,
  O_NNF,
  O_NCFN,
  O_NJE,
  O_QU,
  O_NEXTQ,
  O_H0U,
  O_HU,
  O_NEXTH,
  O_TN,
  O_FLAG,
  O_BARRIERS,
  O_COUNT
};

// Global-memory view of an argument-block pointer.  The kernel reads its
// pointers from the argument block in device memory, so without this the
// compiler cannot tell global from LDS addresses and emits flat accesses,
// which count against the LDS wait counter too: every LDS wait of a pass
// would then also wait for the cell's streaming loads and stores.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr(T* p)
{
  return (__attribute__((address_space(1))) T*)p;
}

struct CplArgs
{
  BucketArgs ba; // ba.R: the reaction count the shared RHS code reads
  int m, R;
  long ncells, N, nchunks;
  int nwg;
  const int* idx;
  const int* soff;
  const int* sent;
  const double* kf; // cell-major (cell * R + r)
  const double* kb;
  const double* y0;
  double* zn; // QMAX + 1 vectors of N
  double* ewt;
  double* acor;
  double* ysav;
  void* lu;     // ncells * M * M factor entries: (cell * M + column) * M + row lane
  int* pos;     // ncells * M
  float* rinv;  // ncells * M (FP32 factors)
  // Two buffers each, used alternately by consecutive reductions: a
  // workgroup that has passed reduction k may write reduction k+1's partials
  // while another still reads reduction k's; it cannot reach reduction k+2
  // before every workgroup has arrived at barrier k+1, i.e. finished reading.
  double* part; // 2 * nwg * NRED
  double* seq;  // 2 * 3 * N norm terms (sequential mode)
  unsigned* bar;
  int* abort;
  double* yout;
  double* out;
  double rtol, atol, tout;
  long mxstep;
  int sequential;
};

// ---- grid barrier ------------------------------------------------------------
// Counter barrier over the cooperative grid (all workgroups resident): each
// workgroup adds one and waits for gen * nwg.  The wait is bounded; a barrier
// that does not complete sets the abort flag.  Thread 0 decides, for its
// whole workgroup, whether the grid is still whole (the answer is shared
// through LDS, so a workgroup's threads never disagree); a workgroup that
// sees the flag stops integrating, and the others see it in their own waits,
// so the grid always drains.
// longest wait at one barrier, in ticks of the 100 MHz constant clock (2 s)
constexpr unsigned long long WAIT_LIMIT = 200000000ull;

__device__ __forceinline__ bool grid_barrier(const CplArgs& A, unsigned& gen, int* s_dead)
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  gen += 1u;
  if (threadIdx.x == 0)
  {
    atomicAdd(A.bar, 1u);
    const unsigned target = gen * (unsigned)A.nwg;
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    int dead                    = 0;
    for (;;)
    {
      const unsigned v = __hip_atomic_load(A.bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      if (v >= target) { break; }
      if (__hip_atomic_load(A.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
      {
        dead = 1;
        break;
      }
      if (__builtin_amdgcn_s_memrealtime() - t0 > WAIT_LIMIT)
      {
        atomicExch(A.abort, 1);
        dead = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    if (__hip_atomic_load(A.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { dead = 1; }
    *s_dead = dead;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return *s_dead != 0;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
  {
    const double o = __shfl_xor(v, off, 64);
    v              = (o > v) ? o : v;
  }
  return v;
}

// the buffer of the next reduction: its parity is that of the barriers
// passed so far, which every workgroup counts alike
__device__ __forceinline__ long buf_of(unsigned gen) { return (long)(gen & 1u); }

// The reduction after a pass: acc[k] are this lane's sums of its terms (tree
// mode), mx its maximum.  Returns the three global sums and the maximum,
// identical in every workgroup.
struct Red
{
  double s[3];
  double mx;
  bool dead; // the grid barrier failed: stop integrating
};

template <int CW>
__device__ __forceinline__ Red reduce(const CplArgs& A, const double (&acc)[3], double mx, unsigned& gen,
                      double (*s_w)[NRED], int* s_dead)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double w[NRED];
#pragma unroll
  for (int k = 0; k < 3; k++) { w[k] = g_tree_sum<64>(A.sequential ? 0.0 : acc[k]); }
  w[3] = wave_max(mx);
  if (lane == 0)
  {
#pragma unroll
    for (int k = 0; k < NRED; k++) { s_w[wave][k] = w[k]; }
  }
  __syncthreads();
  const long pb = buf_of(gen) * (long)A.nwg * NRED; // this reduction's buffers
  const long sb = buf_of(gen) * 3 * A.N;
  if (threadIdx.x == 0)
  {
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
      double t = s_w[0][k];
      for (int v = 1; v < CW; v++) { t = t + s_w[v][k]; }
      gptr(A.part)[pb + (long)blockIdx.x * NRED + k] = t;
    }
    double t = s_w[0][3];
    for (int v = 1; v < CW; v++) { t = (s_w[v][3] > t) ? s_w[v][3] : t; }
    gptr(A.part)[pb + (long)blockIdx.x * NRED + 3] = t;
  }
  Red r;
  r.dead = grid_barrier(A, gen, s_dead);
  // every wavefront folds the workgroups' partials in the same fixed order
  double s[3] = {0.0, 0.0, 0.0}, m = -1.0;
  for (int b = lane; b < A.nwg; b += 64)
  {
#pragma unroll
    for (int k = 0; k < 3; k++) { s[k] = s[k] + gptr(A.part)[pb + (long)b * NRED + k]; }
    const double v = gptr(A.part)[pb + (long)b * NRED + 3];
    m              = (v > m) ? v : m;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) { r.s[k] = g_tree_sum<64>(s[k]); }
  r.mx = wave_max(m);
  if (A.sequential)
  {
    // the serial vector's order: term 0, 1, 2, ... added one at a time
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
      double t = 0.0;
      for (long b = 0; b < A.N; b += 64)
      {
        const long i  = b + lane;
        const double v = (i < A.N) ? gptr(A.seq)[sb + (long)k * A.N + i] : 0.0;
        const int cnt  = (A.N - b < 64) ? (int)(A.N - b) : 64;
        for (int j = 0; j < cnt; j++) { t = t + readlane_d(v, j); }
      }
      r.s[k] = t;
    }
  }
  return r;
}

// ---- the integrator's scalars (every lane holds the same copy) -------------
struct CS
{
  double tn, h, hscale, hprime, eta, etamax, hu, h0u, next_h;
  double rl1, gamma, gammap, gamrat, crate, delp, acnrm, saved_tq5;
  double etaq, etaqm1, etaqp1;
  double tq[6], l[QMAX + 1], tau[QMAX + 2];
  int q, qprime, qwait, L, indx_acor, qu, next_q;
  long nst, nfe, ncfn, netf, nni, nnf, nsetups, nje, nstlp, nstlj, nscon;
  bool nls_jcur; // the Newton solver's record that its matrix is current
  bool jcur;
};

// the vector work of one pass over the cells, applied per element in this order
struct PassD
{
  double scale1;    // != 0: zn1 = scale1 * zn1 (h0, first step)
  int restore_q;    // > 0: undo the predictor of this order
  int save_acor;    // zn[QMAX] = 1.0 * acor (order increase chosen)
  int inc_q;        // > 0: order increase: zn[L] = A1 zn[ia]; zn[j] = l[j] zn[L] + zn[j], 2 <= j <= inc_q
  int inc_ia, inc_L;
  double inc_A1, inc_l[QMAX + 1];
  int dec_q;        // > 0: order decrease: zn[j] = c[j] zn[dec_q] + zn[j], 2 <= j < dec_q
  double dec_c[QMAX + 1];
  int reset1;       // zn1 = reset_h * f(zn0) (order-1 restart)
  double reset_h;
  int rescale_q;    // > 0: zn[j] = cv[j] zn[j], 1 <= j <= rescale_q
  double rescale_c[QMAX + 1];
  int predict_q;    // > 0: predictor of this order
  // Newton iteration
  int acor_zero;    // the correction starts from zero
  int setup;        // 0: none; 1: Newton matrix from ysav; 2: ysav = y, then the matrix
  double rl1, gamma, gscale;
};

__device__ __forceinline__ void pass_clear(PassD& d)
{
  d.scale1    = 0.0;
  d.restore_q = 0;
  d.save_acor = 0;
  d.inc_q     = 0;
  d.dec_q     = 0;
  d.reset1    = 0;
  d.rescale_q = 0;
  d.predict_q = 0;
  d.acor_zero = 0;
  d.setup     = 0;
}

// ---- per-cell helpers ------------------------------------------------------
template <int M, typename T>
struct Cell
{
  long cell, cc, e;
  bool cv, lv;
};

template <int M, typename T>
__device__ __forceinline__ Cell<M, T> cell_at(const CplArgs& A, const Grp<M, T>& g, long ch)
{
  constexpr int GPW = 64 / pow2ceil(M);
  Cell<M, T> c;
  c.cell = ch * GPW + g.row;
  c.cv   = c.cell < A.ncells;
  c.cc   = c.cv ? c.cell : A.ncells - 1;
  c.lv   = c.cv && g.valid;
  c.e    = c.cc * M + (g.valid ? g.i : 0);
  return c;
}

template <int M, typename T>
__device__ __forceinline__ void stage_rates(const CplArgs& A, const Grp<M, T>& g, long cc)
{
  constexpr int G = pow2ceil(M);
  gsync();
  for (int r = g.i; r < A.R; r += G)
  {
    g.rates[r]       = gptr(A.kf)[cc * A.R + r];
    g.rates[A.R + r] = gptr(A.kb)[cc * A.R + r];
  }
  gsync();
}

// the saved factorisation of a cell, as the register LU left it
// The row permutation is kept as each lane's physical row (groups under 32
// lanes, rows stay where the pivoting left them) or as the inverse
// permutation the solve gathers with (32 and 64 lanes, rows moved to lane
// order after the factorisation); the solve reads it from the group's LDS.
template <int M, typename T>
__device__ __forceinline__ void lu_store(const CplArgs& A, const Grp<M, T>& g,
                                         const Sys<M, T>& S, const Cell<M, T>& c)
{
  constexpr int G = pow2ceil(M);
  int p           = S.pos;
  if constexpr (G >= 32) { p = g.inv[g.valid ? g.i : 0]; }
  if (!c.lv) { return; }
  auto lu = gptr((T*)A.lu) + c.cc * (long)M * M + g.i;
#pragma unroll
  for (int j = 0; j < M; j++) { lu[(long)j * M] = S.lu[j]; }
  gptr(A.pos)[c.e] = p;
  if constexpr (sizeof(T) == 4) { gptr(A.rinv)[c.e] = S.rinv; }
}

template <int M, typename T>
__device__ __forceinline__ void lu_load(const CplArgs& A, const Grp<M, T>& g, Sys<M, T>& S,
                                        const Cell<M, T>& c)
{
  constexpr int G = pow2ceil(M);
  const auto lu   = gptr((const T*)A.lu) + c.cc * (long)M * M + (g.valid ? g.i : 0);
#pragma unroll
  for (int j = 0; j < M; j++) { S.lu[j] = g.valid ? lu[(long)j * M] : (T)0; }
  const int p = g.valid ? gptr(A.pos)[c.e] : g.i;
  if constexpr (sizeof(T) == 4) { S.rinv = g.valid ? gptr(A.rinv)[c.e] : 1.0f; }
  gsync();
  if constexpr (G >= 32)
  {
    S.pos = g.i;
    if (g.valid) { g.inv[g.i] = p; }
  }
  else
  {
    S.pos = p;
    if (g.valid) { g.inv[p] = g.i; }
  }
  gsync();
}

__device__ __forceinline__ void add_term(const CplArgs& A, double (&acc)[3], int k, long e,
                                         bool live, double x, double w, unsigned gen)
{
  const double p = x * w;
  const double t = live ? p * p : 0.0;
  if (A.sequential)
  {
    if (live) { gptr(A.seq)[(buf_of(gen) * 3 + k) * A.N + e] = t; }
  }
  else { acc[k] = acc[k] + t; }
}

// reactions per lane whose rate constants a pass keeps in registers (larger
// networks stage every rate constant of the cell through LDS)
constexpr int RPL = 4;

// The right-hand side with each lane's reactions' rate constants in
// registers (reaction r = g.i + t G in kr[t], br[t]): the same operations in
// the same order as g_rhs, which reads them from the group's LDS slot
template <int M, typename T>
__device__ __forceinline__ double rhs_regs(const CplArgs& A, const Grp<M, T>& g, double yi,
                                           const double (&kr)[RPL], const double (&br)[RPL])
{
  constexpr int G = pow2ceil(M);
  gsync();
  if (g.valid) { g.red[g.i] = yi; }
  gsync();
#pragma unroll
  for (int t = 0; t < RPL; t++)
  {
    const int r = g.i + t * G;
    if (r < A.R)
    {
      const int4 q    = *reinterpret_cast<const int4*>(g.idx + 4 * r);
      const double wf = (kr[t] * g.red[q.x]) * g.red[q.y];
      const double wb = (br[t] * g.red[q.z]) * g.red[q.w];
      g.wbuf[r]       = wf - wb;
    }
  }
  gsync();
  double f = 0.0;
  if (g.valid)
  {
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
  return f;
}

// The Newton matrix rows (I - gamma J)_i of each group, with the groups of a
// wavefront taking turns in the shared rows area: the additions of
// g_newton_matrix in the same order, so the rows are the same bits
template <int M, typename T>
__device__ __forceinline__ void newton_matrix_turns(const Grp<M, T>& g, double yj, double gamma,
                                                    T (&a)[M])
{
  constexpr int GPW = 64 / pow2ceil(M);
  gsync();
  if (g.valid) { g.red[g.i] = yj; }
  gsync();
  const double c = -gamma;
  for (int turn = 0; turn < GPW; turn++)
  {
    T* row = g.jrow;
    if (g.row == turn && g.valid)
    {
#pragma unroll
      for (int j = 0; j < M; j++) { row[j] = (T)0; }
      const int e1 = g.soff[g.i + 1];
      for (int e = g.soff[g.i]; e < e1; e++)
      {
        const int code = g.sent[e];
        const int r = code >> 2, t = code & 3;
        int ix[4]    = {g.idx[4 * r], g.idx[4 * r + 1], g.idx[4 * r + 2], g.idx[4 * r + 3]};
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
    if (g.row == turn)
    {
#pragma unroll
      for (int j = 0; j < M; j++)
      {
        double v = g.valid ? (double)row[j] * c : 0.0;
        if (j == g.i) { v = v + 1.0; }
        a[j] = (T)v;
      }
    }
    gsync();
  }
}

// ---- pass: history updates, then one Newton iteration ----------------------
// terms: 0 = ||delta||, 1 = ||acor||; maximum: a failed factorisation (1)
template <int M, typename T>
__device__ __forceinline__ void pass_newton(const CplArgs& A, const Grp<M, T>& g, Sys<M, T>& S, const PassD& d,
                            long ch0, long ch1, double (&acc)[3], double& mx, unsigned gen)
{
  const long N        = A.N;
  const bool zn_ops   = (d.scale1 != 0.0) || d.restore_q || d.save_acor || d.inc_q || d.dec_q ||
                      d.reset1 || d.rescale_q || d.predict_q;
  constexpr int G  = pow2ceil(M);
  const bool regs  = A.R <= RPL * G;
  for (long ch = ch0; ch < ch1; ch++)
  {
    const Cell<M, T> c = cell_at(A, g, ch);
    // every global operand of the cell is requested before the first wait:
    // the lane's rate constants, the history columns (zn0 and zn1 only in
    // the iterations after the first), the correction, the weights and, when
    // no setup is due, the saved factors
    double kr[RPL], br[RPL];
#pragma unroll
    for (int t = 0; t < RPL; t++)
    {
      const int r = g.i + t * G;
      kr[t]       = (regs && r < A.R) ? gptr(A.kf)[c.cc * A.R + r] : 0.0;
      br[t]       = (regs && r < A.R) ? gptr(A.kb)[c.cc * A.R + r] : 0.0;
    }
    double z[QMAX + 1];
#pragma unroll
    for (int j = 0; j <= QMAX; j++) { z[j] = (c.lv && (j <= 1 || zn_ops)) ? gptr(A.zn)[(long)j * N + c.e] : 0.0; }
    double acor = (d.acor_zero && !d.save_acor) ? 0.0 : (c.lv ? gptr(A.acor)[c.e] : 0.0);
    const double w = c.lv ? gptr(A.ewt)[c.e] : 0.0;
    if (!d.setup) { lu_load(A, g, S, c); }
    // the Newton matrix and the order-1 restart's RHS read any reaction's
    // rates: the cell's rates go to the group's LDS slot
    if (!regs) { stage_rates(A, g, c.cc); }
    else if (d.setup || d.reset1)
    {
      gsync();
#pragma unroll
      for (int t = 0; t < RPL; t++)
      {
        const int r = g.i + t * G;
        if (r < A.R)
        {
          g.rates[r]       = kr[t];
          g.rates[A.R + r] = br[t];
        }
      }
      gsync();
    }
    if (zn_ops)
    {
      if (d.scale1 != 0.0) { z[1] = d.scale1 * z[1]; }
      if (d.restore_q)
      {
#pragma unroll
        for (int k = 1; k <= QMAX; k++)
        {
#pragma unroll
          for (int j = QMAX; j >= 1; j--)
          {
            if (k <= d.restore_q && j >= k && j <= d.restore_q) { z[j - 1] = z[j - 1] - z[j]; }
          }
        }
      }
      if (d.save_acor) { z[QMAX] = 1.0 * acor; }
      if (d.inc_q)
      {
        double src = z[0];
#pragma unroll
        for (int j = 1; j <= QMAX; j++) { src = (j == d.inc_ia) ? z[j] : src; }
        const double zl = d.inc_A1 * src;
#pragma unroll
        for (int j = 0; j <= QMAX; j++) { z[j] = (j == d.inc_L) ? zl : z[j]; }
#pragma unroll
        for (int j = 2; j <= QMAX; j++)
        {
          if (j <= d.inc_q) { z[j] = (d.inc_l[j] * zl) + z[j]; }
        }
      }
      if (d.dec_q)
      {
        double zq = z[0];
#pragma unroll
        for (int j = 1; j <= QMAX; j++) { zq = (j == d.dec_q) ? z[j] : zq; }
#pragma unroll
        for (int j = 2; j < QMAX; j++)
        {
          if (j < d.dec_q) { z[j] = (d.dec_c[j] * zq) + z[j]; }
        }
      }
      if (d.reset1)
      {
        const double f0 = g_rhs(g, z[0]);
        z[1]            = d.reset_h * f0;
      }
      if (d.rescale_q)
      {
#pragma unroll
        for (int j = 1; j <= QMAX; j++)
        {
          if (j <= d.rescale_q) { z[j] = d.rescale_c[j] * z[j]; }
        }
      }
      if (d.predict_q)
      {
#pragma unroll
        for (int k = 1; k <= QMAX; k++)
        {
#pragma unroll
          for (int j = QMAX; j >= 1; j--)
          {
            if (k <= d.predict_q && j >= k && j <= d.predict_q) { z[j - 1] = z[j - 1] + z[j]; }
          }
        }
      }
      if (c.lv)
      {
#pragma unroll
        for (int j = 0; j <= QMAX; j++) { gptr(A.zn)[(long)j * N + c.e] = z[j]; }
      }
    }
    if (d.acor_zero) { acor = 0.0; }
    const double y = lsum(1.0, z[0], 1.0, acor);
    if (d.setup)
    {
      double ys;
      if (d.setup == 2)
      {
        ys = y;
        if (c.lv) { gptr(A.ysav)[c.e] = ys; }
      }
      else { ys = c.lv ? gptr(A.ysav)[c.e] : 0.0; }
      if constexpr (cpl_shared_rows<M, T>()) { newton_matrix_turns(g, ys, d.gamma, S.lu); }
      else { g_newton_matrix(g, ys, d.gamma, S.lu); }
      const int fail = g_getrf(g, S);
      if (c.cv && fail) { mx = 1.0; }
      lu_store(A, g, S, c);
    }
    const double f = regs ? rhs_regs(A, g, y, kr, br) : g_rhs(g, y);
    const double res = lsum(d.rl1, z[1], 1.0, acor);
    double delta     = lsum(-d.gamma, f, 1.0, res);
    delta            = -1.0 * delta;
    delta            = g_getrs(g, S, delta);
    if (d.gscale != 1.0) { delta = d.gscale * delta; }
    acor = lsum(1.0, acor, 1.0, delta);
    if (c.lv) { gptr(A.acor)[c.e] = acor; }
    add_term(A, acc, 0, c.e, c.lv, delta, w, gen);
    add_term(A, acc, 1, c.e, c.lv, acor, w, gen);
  }
}

// ---- pass: accepted step -----------------------------------------------------
// zn[j] += l[j] acor (j <= q), the saved correction, the order-selection norms
// (terms 0: zn[q] for etaqm1, 1: acor - cquot zn[QMAX] for etaqp1, with the
// step's weights), then the next step's weights and term 2: ||zn0||
struct DoneD
{
  int q;
  double l[QMAX + 1];
  int save;    // zn[QMAX] = 1.0 * acor
  int qm1;     // norm of zn[q]
  int qp1;     // norm of acor - cquot zn[QMAX]
  double cquot;
};

template <int M, typename T>
__device__ __forceinline__ void pass_done(const CplArgs& A, const Grp<M, T>& g, const DoneD& d, long ch0, long ch1,
                          double (&acc)[3], unsigned gen)
{
  const long N = A.N;
  for (long ch = ch0; ch < ch1; ch++)
  {
    const Cell<M, T> c = cell_at(A, g, ch);
    if (!c.lv)
    {
      // every lane takes part in nothing collective here; dead lanes add no terms
      continue;
    }
    // columns 0..q, and zn[QMAX] where the saved correction is read or written
    const bool qmax_used = d.qp1 || d.save;
    double z[QMAX + 1];
#pragma unroll
    for (int j = 0; j <= QMAX; j++)
    {
      z[j] = (j <= d.q || (j == QMAX && qmax_used)) ? gptr(A.zn)[(long)j * N + c.e] : 0.0;
    }
    const double acor = gptr(A.acor)[c.e];
    const double w    = gptr(A.ewt)[c.e];
#pragma unroll
    for (int j = 0; j <= QMAX; j++)
    {
      if (j <= d.q) { z[j] = (d.l[j] * acor) + z[j]; }
    }
    if (d.save) { z[QMAX] = 1.0 * acor; }
#pragma unroll
    for (int j = 0; j <= QMAX; j++)
    {
      if (j <= d.q || (j == QMAX && d.save)) { gptr(A.zn)[(long)j * N + c.e] = z[j]; }
    }
    if (d.qm1)
    {
      double zq = z[0];
#pragma unroll
      for (int j = 1; j <= QMAX; j++) { zq = (j == d.q) ? z[j] : zq; }
      add_term(A, acc, 0, c.e, true, zq, w, gen);
    }
    if (d.qp1)
    {
      const double t = lsum(-d.cquot, z[QMAX], 1.0, acor);
      add_term(A, acc, 1, c.e, true, t, w, gen);
    }
    // the next step's error weights (cvEwtSet, scalar tolerances) and the
    // norm of zn0 in them (CVode's tolerance-scale check)
    const double wn = 1.0 / ((A.rtol * fabs(z[0])) + A.atol);
    gptr(A.ewt)[c.e]      = wn;
    add_term(A, acc, 2, c.e, true, z[0], wn, gen);
  }
}

// ---- the integration ---------------------------------------------------------
// At most 256 registers (two wavefronts per SIMD).
// The arguments are read from device memory (a pointer, not a by-value
// struct whose address the shared per-cell code takes, which would put a
// private copy of it in scratch memory)
template <int M, typename T>
__global__ void __launch_bounds__(64 * cpl_waves<M>(), 2) k_coupled_bdf(const CplArgs* __restrict__ Ap)
{
  const CplArgs& A = *Ap;
  constexpr int G   = pow2ceil(M);
  constexpr int GPW = 64 / G;
  constexpr int CW  = cpl_waves<M>();
  constexpr int CT  = 64 * CW;
  __shared__ double s_red[CW][64];
  __shared__ T s_prow[CW][64];
  __shared__ int s_inv[CW][64];
  __shared__ GroupScalars s_sc[CW * GPW];
  __shared__ double s_w[CW][NRED];
  __shared__ int s_dead;
  // the pass descriptors as the passes read them (every thread computes the
  // same descriptor; thread 0 publishes it)
  __shared__ PassD s_pd;
  __shared__ DoneD s_dd;
  extern __shared__ __align__(16) double s_dyn[];
  int* s_idx     = (int*)s_dyn;
  int* s_sent    = s_idx + 4 * A.R;
  int* s_soff    = s_sent + 4 * A.R;
  double* s_grp  = (double*)(s_soff + ((M + 2) & ~1));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t = threadIdx.x; t < 4 * A.R; t += CT)
  {
    s_idx[t]  = gptr(A.idx)[t];
    s_sent[t] = gptr(A.sent)[t];
  }
  for (int t = threadIdx.x; t <= M; t += CT) { s_soff[t] = gptr(A.soff)[t]; }
  __syncthreads();

  const int grp = lane / G;
  Grp<M, T> g;
  g.i     = lane % G;
  g.base  = grp * G;
  g.valid = g.i < M;
  g.red   = s_red[wave] + g.base;
  g.prow  = s_prow[wave] + g.base;
  g.inv   = s_inv[wave] + g.base;
  g.sc    = &s_sc[wave * GPW + grp];
  g.mlp   = nullptr;
  g.wx    = nullptr;
  g.row   = grp;
  g.learn = false;
  g.a     = &A.ba;
  g.idx   = s_idx;
  g.soff  = s_soff;
  g.sent  = s_sent;
  double* wave_area = s_grp + (long)wave * cpl_wave_doubles<M, T>(A.R);
  double* area      = wave_area + (long)grp * cpl_group_doubles<M, T>(A.R);
  g.rates = area;
  g.kf    = area;
  g.kb    = area + A.R;
  g.wbuf  = area + 2 * A.R;
  if constexpr (cpl_shared_rows<M, T>())
  {
    g.jrow = (T*)(wave_area + (long)GPW * cpl_group_doubles<M, T>(A.R)) +
             (long)(g.valid ? g.i : 0) * (M + 1);
  }
  else { g.jrow = (T*)(area + 2 * A.R) + (long)(g.valid ? g.i : 0) * (M + 1); }
  g.cell  = 0;
  Sys<M, T> S;
  S.pos  = g.i;
  S.rinv = 1.0f;

  // this wavefront's cells: chunks [ch0, ch1) of GPW consecutive cells
  const long NWV = (long)A.nwg * CW;
  const long gw  = (long)blockIdx.x * CW + wave;
  const long ch0 = gw * A.nchunks / NWV;
  const long ch1 = (gw + 1) * A.nchunks / NWV;
  const long N   = A.N;
  const double Nd = (double)N;
  unsigned gen   = 0;
  int flag       = CV_SUCCESS;

  CS s;
  s.tn = 0.0;
  s.q = 1;
  s.L = 2;
  s.qwait = 2;
  s.etamax = ETA_MAX_FS;
  s.qu = 0;
  s.hu = 0.0;
  s.nst = s.nfe = s.ncfn = s.netf = s.nni = s.nnf = s.nsetups = s.nje = 0;
  s.nstlp = s.nstlj = s.nscon = 0;
  s.h0u = 0.0;
  s.next_h = 0.0;
  s.next_q = 0;
  s.saved_tq5 = 0.0;
  s.indx_acor = 0;
  s.nls_jcur = false;
  s.jcur = false;
  s.crate = 1.0;
  s.delp = 0.0;
  s.gammap = 0.0;
  s.gamrat = 1.0;
  s.eta = 1.0;
  s.hprime = 0.0;
  s.qprime = 1;
  for (int i = 0; i <= QMAX + 1; i++) { s.tau[i] = 0.0; }
  for (int i = 0; i <= QMAX; i++) { s.l[i] = 0.0; }
  for (int i = 0; i < 6; i++) { s.tq[i] = 0.0; }

  // ---- first call: zn0 = y0, weights, f(t0, y0) into zn1, the norms cvHin
  // and the tolerance check start from
  double acc[3] = {0.0, 0.0, 0.0};
  double mx     = 0.0;
  for (long ch = ch0; ch < ch1; ch++)
  {
    const Cell<M, T> c = cell_at(A, g, ch);
    stage_rates(A, g, c.cc);
    const double y  = c.lv ? gptr(A.y0)[c.e] : 0.0;
    const double w  = 1.0 / ((A.rtol * fabs(y)) + A.atol);
    const double f0 = g_rhs(g, y);
    if (c.lv)
    {
      gptr(A.zn)[c.e]     = y;
      gptr(A.zn)[N + c.e] = f0;
      gptr(A.ewt)[c.e]    = w;
      // cvUpperBoundH0: |f0| / (HUB |y0| + 1/w)
      const double inv = 1.0 / w;
      const double den = (HUB_FACTOR * fabs(y)) + (1.0 * inv);
      const double r   = fabs(f0) / den;
      mx               = (r > mx) ? r : mx;
    }
    add_term(A, acc, 2, c.e, c.lv, y, w, gen);
  }
  Red red   = reduce<CW>(A, acc, mx, gen, s_w, &s_dead);
  bool dead = red.dead;
  s.nfe     = 1;
  double nrm0 = sqrt0(red.s[2] / Nd);

  // ---- cvHin
  {
    const double tout  = A.tout;
    const double tdiff = tout - s.tn;
    const int sign     = (tdiff > 0.0) ? 1 : -1;
    const double tdist = fabs(tdiff);
    const double tround = UROUND * fmax(fabs(s.tn), fabs(tout));
    if (tdiff == 0.0 || tdist < 2.0 * tround) { flag = CV_TOO_CLOSE; }
    if (flag == CV_SUCCESS && !dead)
    {
      const double hlb = HLB_FACTOR * tround;
      const double hub_inv = red.mx;
      double hub = HUB_FACTOR * tdist;
      if (hub * hub_inv > 1.0) { hub = 1.0 / hub_inv; }
      double hg = sqrt0(hlb * hub);
      if (hub < hlb) { s.h = (sign == -1) ? -hg : hg; }
      else
      {
        double hnew = hg, hs = hg;
        for (int count1 = 1; count1 <= MAX_ITERS; count1++)
        {
          const double hgs = hg * sign;
          // cvYddNorm: y = hgs zn1 + zn0, f(t + hgs, y), (f - zn1) / hgs
          double a2[3] = {0.0, 0.0, 0.0};
          const double ra = 1.0 / hgs, rb = -1.0 / hgs;
          for (long ch = ch0; ch < ch1; ch++)
          {
            const Cell<M, T> c = cell_at(A, g, ch);
            stage_rates(A, g, c.cc);
            const double z0 = c.lv ? gptr(A.zn)[c.e] : 0.0;
            const double z1 = c.lv ? gptr(A.zn)[N + c.e] : 0.0;
            const double y  = lsum(hgs, z1, 1.0, z0);
            const double f  = g_rhs(g, y);
            const double t  = lsum(ra, f, rb, z1);
            add_term(A, a2, 0, c.e, c.lv, t, c.lv ? gptr(A.ewt)[c.e] : 0.0, gen);
          }
          Red r2 = reduce<CW>(A, a2, 0.0, gen, s_w, &s_dead);
          s.nfe++;
          dead = r2.dead;
          if (dead) { break; }
          const double yddnrm = sqrt0(r2.s[0] / Nd);
          hs   = hg;
          hnew = (yddnrm * hub * hub > 2.0) ? sqrt0(2.0 / yddnrm) : sqrt0(hg * hub);
          if (count1 == MAX_ITERS) { break; }
          const double hrat = hnew / hg;
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
        s.h = h0;
      }
    }
  }
  s.hscale = s.h;
  s.h0u    = s.h;
  s.hprime = s.h;

  PassD pd;
  pass_clear(pd);
  pd.scale1 = s.h;
  long nstloc  = 0;
  bool out_dky = false; // end: interpolate at tout (else: return zn0)

  while (flag == CV_SUCCESS && !dead)
  {
    s.next_h = s.h;
    s.next_q = s.q;
    if (A.mxstep > 0 && nstloc >= A.mxstep)
    {
      flag = CV_TOO_MUCH_WORK;
      break;
    }
    if (UROUND * nrm0 > 1.0)
    {
      flag = CV_TOO_MUCH_ACC;
      break;
    }

    // ---- cvStep
    if ((s.nst > 0) && (s.hprime != s.h))
    {
      // cvAdjustParams: order change, then the rescale
      if (s.qprime != s.q)
      {
        const int deltaq = s.qprime - s.q;
        if (!((s.q == 2) && (deltaq != 1)))
        {
          if (deltaq == 1)
          {
            // cvIncreaseBDF
            double l[QMAX + 1];
            for (int i = 0; i <= QMAX; i++) { l[i] = 0.0; }
            double alpha1, prod, xiold, alpha0, hsum, xi;
            l[2] = alpha1 = prod = xiold = 1.0;
            alpha0 = -1.0;
            hsum   = s.hscale;
            if (s.q > 1)
            {
              for (int j = 1; j < s.q; j++)
              {
                hsum += s.tau[j + 1];
                xi = hsum / s.hscale;
                prod *= xi;
                alpha0 -= 1.0 / (j + 1);
                alpha1 += 1.0 / xi;
                for (int i = j + 2; i >= 2; i--) { l[i] = l[i] * xiold + l[i - 1]; }
                xiold = xi;
              }
            }
            pd.inc_A1 = (-alpha0 - alpha1) / prod;
            pd.inc_ia = s.indx_acor;
            pd.inc_L  = s.L;
            pd.inc_q  = s.q; // zn[j] += l[j] zn[L] for 2 <= j <= q
            for (int i = 0; i <= QMAX; i++) { pd.inc_l[i] = l[i]; }
            if (s.q <= 1) { pd.inc_q = 1; } // no ScaleAddMulti; zn[L] still set
          }
          else if (deltaq == -1)
          {
            // cvDecreaseBDF
            double l[QMAX + 1];
            for (int i = 0; i <= QMAX; i++) { l[i] = 0.0; }
            l[2]        = 1.0;
            double hsum = 0.0;
            for (int j = 1; j <= s.q - 2; j++)
            {
              hsum += s.tau[j];
              const double xi = hsum / s.hscale;
              for (int i = j + 2; i >= 2; i--) { l[i] = l[i] * xi + l[i - 1]; }
            }
            for (int j = 0; j <= QMAX; j++) { pd.dec_c[j] = (j >= 2 && j < s.q) ? -l[j] : 0.0; }
            pd.dec_q = (s.q > 2) ? s.q : 0;
          }
        }
        s.q     = s.qprime;
        s.L     = s.q + 1;
        s.qwait = s.L;
      }
      // cvRescale
      {
        double cv = s.eta;
        for (int j = 1; j <= s.q; j++)
        {
          pd.rescale_c[j] = cv;
          cv              = s.eta * cv;
        }
        pd.rescale_q = s.q;
        s.h          = s.hscale * s.eta;
        s.next_h     = s.h;
        s.hscale     = s.h;
        s.nscon      = 0;
      }
    }

    const double saved_t = s.tn;
    int ncf = 0, nef = 0;
    int nflag = FIRST_CALL;
    double dsm = 0.0;
    for (;;)
    {
      // cvPredict, cvSet
      s.tn += s.h;
      pd.predict_q = s.q;
      {
        double l[QMAX + 1];
        double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
        l[0] = l[1] = xi_inv = xistar_inv = 1.0;
        for (int i = 2; i <= QMAX; i++) { l[i] = 0.0; }
        alpha0 = alpha0_hat = -1.0;
        hsum                = s.h;
        if (s.q > 1)
        {
          for (int j = 2; j < s.q; j++)
          {
            hsum += s.tau[j - 1];
            xi_inv = s.h / hsum;
            alpha0 -= 1.0 / j;
            for (int i = j; i >= 1; i--) { l[i] += l[i - 1] * xi_inv; }
          }
          alpha0 -= 1.0 / s.q;
          xistar_inv = -l[1] - alpha0;
          hsum += s.tau[s.q - 1];
          xi_inv     = s.h / hsum;
          alpha0_hat = -l[1] - xi_inv;
          for (int i = s.q; i >= 1; i--) { l[i] += l[i - 1] * xistar_inv; }
        }
        // cvSetTqBDF
        double A1 = 1.0 - alpha0_hat + alpha0;
        double A2 = 1.0 + s.q * A1;
        s.tq[2]   = fabs(A1 / (alpha0 * A2));
        s.tq[5]   = fabs(A2 * xistar_inv / (l[s.q] * xi_inv));
        if (s.qwait == 1)
        {
          if (s.q > 1)
          {
            const double Cc    = xistar_inv / l[s.q];
            const double A3    = alpha0 + 1.0 / s.q;
            const double A4    = alpha0_hat + xi_inv;
            const double Cpinv = (1.0 - A4 + A3) / A3;
            s.tq[1]            = fabs(Cc * Cpinv);
          }
          else { s.tq[1] = 1.0; }
          hsum += s.tau[s.q];
          xi_inv              = s.h / hsum;
          const double A5     = alpha0 - (1.0 / (s.q + 1));
          const double A6     = alpha0_hat - xi_inv;
          const double Cppinv = (1.0 - A6 + A5) / A2;
          s.tq[3]             = fabs(Cppinv / (xi_inv * (s.q + 2) * A5));
        }
        s.tq[4] = CORTES / s.tq[2];
        for (int i = 0; i <= QMAX; i++) { s.l[i] = (i <= s.q) ? l[i] : 0.0; }
        s.rl1   = 1.0 / s.l[1];
        s.gamma = s.h * s.rl1;
        if (s.nst == 0) { s.gammap = s.gamma; }
        s.gamrat = (s.nst > 0) ? s.gamma / s.gammap : 1.0;
      }

      // ---- cvNls: setup decision, then the Newton iteration
      const int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES
                                                                               : CV_FAIL_OTHER;
      bool callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (s.nst == 0) ||
                       (s.nst >= s.nstlp + MSBP) || (fabs(s.gamrat - 1.0) > DGMAX_LSETUP);
      long niters = 0, nconvfails = 0;
      bool jbad = false;
      int retval = CV_SUCCESS;
      int conv   = convfail;
      // sunnonlinsol_newton.c's two loops as one loop of passes (a single
      // pass site keeps the per-cell code, its unrolled LU and solve, in one
      // copy): `start` marks the pass that (re)starts the solve from a zero
      // correction -- history updates, residual, the setup when due, solve;
      // the others are the iterations that follow a non-converged test
      bool start  = true;
      int curiter = 0;
      for (;;)
      {
        bool did_setup = false;
        if (start)
        {
          curiter      = 0;
          pd.acor_zero = 1;
          if (callSetup)
          {
            // cvNlsLSetup / cvLsSetup
            if (jbad) { conv = CV_FAIL_BAD_J; }
            const double dgamma = fabs((s.gamma / s.gammap) - 1.0);
            const bool ls_jbad  = (s.nst == 0) || (s.nst >= s.nstlj + MSBJ) ||
                                 ((conv == CV_FAIL_BAD_J) && (dgamma < DGMAX_JBAD)) ||
                                 (conv == CV_FAIL_OTHER);
            pd.setup = ls_jbad ? 2 : 1;
            s.jcur   = ls_jbad;
            if (ls_jbad)
            {
              s.nje++;
              s.nstlj = s.nst;
            }
            s.nsetups++;
            s.nls_jcur = s.jcur;
            s.gamrat   = 1.0;
            s.gammap   = s.gamma;
            s.crate    = 1.0;
            s.nstlp    = s.nst;
            did_setup  = true;
          }
        }
        pd.rl1       = s.rl1;
        pd.gamma     = s.gamma;
        pd.gscale    = (s.gamrat != 1.0) ? 2.0 / (1.0 + s.gamrat) : 1.0;
        double a0[3] = {0.0, 0.0, 0.0};
        double m0    = 0.0;
        __syncthreads();
        if (threadIdx.x == 0) { s_pd = pd; }
        __syncthreads();
        pass_newton(A, g, S, s_pd, ch0, ch1, a0, m0, gen);
        s.nfe++;
        pass_clear(pd);
        start = false;
        Red r = reduce<CW>(A, a0, m0, gen, s_w, &s_dead);
        dead  = r.dead;
        if (dead) { break; }
        // a failed setup ends the solve (no retry)
        if (did_setup && r.mx > 0.0)
        {
          retval = NLS_CONV_RECVR;
          break;
        }
        niters++;
        // cvNlsConvTest
        const double del = sqrt0(r.s[0] / Nd);
        if (curiter > 0)
        {
          const double rr = del / s.delp;
          s.crate         = (CRDOWN * s.crate > rr) ? CRDOWN * s.crate : rr;
        }
        const double dcon = del * ((1.0 < s.crate) ? 1.0 : s.crate) / s.tq[4];
        int ct;
        if (dcon <= 1.0)
        {
          s.acnrm = (curiter == 0) ? del : sqrt0(r.s[1] / Nd);
          ct      = CV_SUCCESS;
        }
        else if ((curiter >= 1) && (del > RDIV * s.delp)) { ct = NLS_CONV_RECVR; }
        else
        {
          s.delp = del;
          ct     = NLS_CONTINUE;
        }
        curiter++;
        if (ct == CV_SUCCESS)
        {
          s.nls_jcur = false;
          retval     = CV_SUCCESS;
          break;
        }
        if (ct == NLS_CONTINUE && curiter < NLS_MAXCOR) { continue; }
        retval = NLS_CONV_RECVR;
        // a convergence failure with stale Jacobian data re-runs the setup
        // and retries from zero
        if (!s.nls_jcur)
        {
          nconvfails++;
          callSetup = true;
          jbad      = true;
          start     = true;
          continue;
        }
        break;
      }
      if (dead) { break; }
      if (retval != CV_SUCCESS) { nconvfails++; }
      s.nni += niters;
      s.nnf += nconvfails;
      if (retval == CV_SUCCESS) { s.jcur = false; }

      // ---- cvHandleNFlag
      if (retval != CV_SUCCESS)
      {
        s.ncfn++;
        s.tn         = saved_t;
        pd.restore_q = s.q;
        ncf++;
        s.etamax = 1.0;
        if (fabs(s.h) <= 0.0 || ncf == MXNCF)
        {
          flag = CV_CONV_FAILURE;
          break;
        }
        s.eta = ETA_CF; // max(eta_cf, hmin / |h|) with hmin = 0
        nflag = PREV_CONV_FAIL;
        {
          double cv = s.eta;
          for (int j = 1; j <= s.q; j++)
          {
            pd.rescale_c[j] = cv;
            cv              = s.eta * cv;
          }
          pd.rescale_q = s.q;
          s.h          = s.hscale * s.eta;
          s.next_h     = s.h;
          s.hscale     = s.h;
          s.nscon      = 0;
        }
        continue;
      }

      // ---- cvDoErrorTest
      dsm = s.acnrm * s.tq[2];
      if (dsm <= 1.0) { break; }
      nef++;
      s.netf++;
      nflag        = PREV_ERR_FAIL;
      s.tn         = saved_t;
      pd.restore_q = s.q;
      if (fabs(s.h) <= 0.0 || nef == MXNEF)
      {
        flag = CV_ERR_FAILURE;
        break;
      }
      s.etamax = 1.0;
      if (nef <= MXNEF1)
      {
        s.eta = 1.0 / (sunamd_pow(BIAS2 * dsm, 1.0 / s.L) + ADDON);
        s.eta = fmax(ETA_MIN_EF, fmax(s.eta, 0.0));
        if (nef >= SMALL_NEF) { s.eta = fmin(s.eta, ETA_MAX_EF); }
      }
      else if (s.q > 1)
      {
        s.eta = fmax(ETA_MIN_EF, 0.0);
        // cvAdjustOrder(-1): cvDecreaseBDF with the current q
        if (!((s.q == 2)))
        {
          double l[QMAX + 1];
          for (int i = 0; i <= QMAX; i++) { l[i] = 0.0; }
          l[2]        = 1.0;
          double hsum = 0.0;
          for (int j = 1; j <= s.q - 2; j++)
          {
            hsum += s.tau[j];
            const double xi = hsum / s.hscale;
            for (int i = j + 2; i >= 2; i--) { l[i] = l[i] * xi + l[i - 1]; }
          }
          for (int j = 0; j <= QMAX; j++) { pd.dec_c[j] = (j >= 2 && j < s.q) ? -l[j] : 0.0; }
          pd.dec_q = (s.q > 2) ? s.q : 0;
        }
        s.L     = s.q;
        s.q     = s.q - 1;
        s.qwait = s.L;
      }
      else
      {
        s.eta = fmax(ETA_MIN_EF, 0.0);
        s.h *= s.eta;
        s.next_h   = s.h;
        s.hscale   = s.h;
        s.qwait    = LONG_WAIT;
        s.nscon    = 0;
        pd.reset1  = 1;
        pd.reset_h = s.h;
        s.nfe++;
        continue;
      }
      {
        double cv = s.eta;
        for (int j = 1; j <= s.q; j++)
        {
          pd.rescale_c[j] = cv;
          cv              = s.eta * cv;
        }
        pd.rescale_q = s.q;
        s.h          = s.hscale * s.eta;
        s.next_h     = s.h;
        s.hscale     = s.h;
        s.nscon      = 0;
      }
    }
    if (flag != CV_SUCCESS || dead) { break; }

    // ---- cvCompleteStep
    s.nst++;
    s.nscon++;
    s.hu = s.h;
    s.qu = s.q;
    for (int i = s.q; i >= 2; i--) { s.tau[i] = s.tau[i - 1]; }
    if ((s.q == 1) && (s.nst > 1)) { s.tau[2] = s.tau[1]; }
    s.tau[1] = s.h;
    s.qwait--;
    DoneD dd;
    dd.q = s.q;
    for (int i = 0; i <= QMAX; i++) { dd.l[i] = s.l[i]; }
    dd.save = 0;
    if ((s.qwait == 1) && (s.q != QMAX))
    {
      dd.save     = 1;
      s.saved_tq5 = s.tq[5];
      s.indx_acor = QMAX;
    }
    // cvPrepareNextStep needs two more norms when the order may change
    const bool choose = (s.etamax != 1.0) && (s.qwait == 0);
    dd.qm1   = choose && (s.q > 1);
    dd.qp1   = choose && (s.q != QMAX) && (s.saved_tq5 != 0.0);
    dd.cquot = 0.0;
    if (dd.qp1)
    {
      double pw = 1.0;
      const double hr = s.h / s.tau[2];
      for (int i = 1; i <= s.L; i++) { pw *= hr; } // SUNRpowerI
      dd.cquot = (s.tq[5] / s.saved_tq5) * pw;
    }
    double a3[3] = {0.0, 0.0, 0.0};
    __syncthreads();
    if (threadIdx.x == 0) { s_dd = dd; }
    __syncthreads();
    pass_done(A, g, s_dd, ch0, ch1, a3, gen);
    Red rd = reduce<CW>(A, a3, 0.0, gen, s_w, &s_dead);
    dead   = rd.dead;
    if (dead) { break; }
    nrm0 = sqrt0(rd.s[2] / Nd);

    // ---- cvPrepareNextStep
    if (s.etamax == 1.0)
    {
      s.qwait  = (s.qwait > 2) ? s.qwait : 2;
      s.qprime = s.q;
      s.hprime = s.h;
      s.eta    = 1.0;
    }
    else
    {
      s.etaq = 1.0 / (sunamd_pow(BIAS2 * dsm, 1.0 / s.L) + ADDON);
      if (s.qwait != 0)
      {
        s.eta    = s.etaq;
        s.qprime = s.q;
      }
      else
      {
        s.qwait  = 2;
        s.etaqm1 = 0.0;
        if (s.q > 1)
        {
          const double ddn = sqrt0(rd.s[0] / Nd) * s.tq[1];
          s.etaqm1         = 1.0 / (sunamd_pow(BIAS1 * ddn, 1.0 / s.q) + ADDON);
        }
        s.etaqp1 = 0.0;
        if (s.q != QMAX && s.saved_tq5 != 0.0)
        {
          const double dup = sqrt0(rd.s[1] / Nd) * s.tq[3];
          s.etaqp1         = 1.0 / (sunamd_pow(BIAS3 * dup, 1.0 / (s.L + 1)) + ADDON);
        }
        // cvChooseEta
        const double etam = fmax(s.etaqm1, fmax(s.etaq, s.etaqp1));
        if ((etam > ETA_MIN_FX) && (etam < ETA_MAX_FX))
        {
          s.eta    = 1.0;
          s.qprime = s.q;
        }
        else if (etam == s.etaq)
        {
          s.eta    = s.etaq;
          s.qprime = s.q;
        }
        else if (etam == s.etaqm1)
        {
          s.eta    = s.etaqm1;
          s.qprime = s.q - 1;
        }
        else
        {
          s.eta        = s.etaqp1;
          s.qprime     = s.q + 1;
          pd.save_acor = 1; // Delta_n saved in zn[qmax] for the order increase
        }
      }
      // cvSetEta
      if ((s.eta > ETA_MIN_FX) && (s.eta < ETA_MAX_FX))
      {
        s.eta    = 1.0;
        s.hprime = s.h;
      }
      else
      {
        if (s.eta >= ETA_MAX_FX)
        {
          s.eta = fmin(s.eta, s.etamax);
          s.eta /= fmax(1.0, 0.0);
        }
        else
        {
          s.eta = fmax(s.eta, ETA_MIN);
          s.eta = fmax(s.eta, 0.0);
        }
        s.hprime = s.h * s.eta;
        if (s.qprime < s.q) { s.nscon = 0; }
      }
    }
    s.etamax = (s.nst <= SMALL_NST) ? ETA_MAX_ES : ETA_MAX_GS;
    nstloc++;

    if ((s.tn - A.tout) * s.h >= 0.0)
    {
      s.next_q = s.qprime;
      s.next_h = s.hprime;
      out_dky  = true;
      break;
    }
  }

  // ---- output: y(tout) from the history (CVodeGetDky, k = 0), or zn0 after
  // any pending undo of the predictor
  if (!dead)
  {
    double c[QMAX + 1];
    int nvec = 0;
    if (out_dky)
    {
      const double sdk = (A.tout - s.tn) / s.h;
      for (int j = s.q; j >= 0; j--)
      {
        double cj = 1.0;
        for (int i = 0; i < j; i++) { cj *= sdk; }
        c[nvec++] = cj;
      }
    }
    for (long ch = ch0; ch < ch1; ch++)
    {
      const Cell<M, T> cl = cell_at(A, g, ch);
      if (!cl.lv) { continue; }
      double z[QMAX + 1];
#pragma unroll
      for (int j = 0; j <= QMAX; j++) { z[j] = gptr(A.zn)[(long)j * N + cl.e]; }
      if (!out_dky && pd.restore_q)
      {
#pragma unroll
        for (int k = 1; k <= QMAX; k++)
        {
#pragma unroll
          for (int j = QMAX; j >= 1; j--)
          {
            if (k <= pd.restore_q && j >= k && j <= pd.restore_q) { z[j - 1] = z[j - 1] - z[j]; }
          }
        }
      }
      double yv;
      if (out_dky)
      {
        // N_VLinearCombination over zn[q], ..., zn[0]: c0 X0, then + ck Xk
        double zq = z[0];
#pragma unroll
        for (int j = 1; j <= QMAX; j++) { zq = (j == s.q) ? z[j] : zq; }
        yv = c[0] * zq;
        for (int k = 1; k < nvec; k++)
        {
          const int j = s.q - k;
          double zj   = z[0];
#pragma unroll
          for (int t = 1; t <= QMAX; t++) { zj = (t == j) ? z[t] : zj; }
          yv = yv + c[k] * zj;
        }
      }
      else { yv = z[0]; }
      gptr(A.yout)[cl.e] = yv;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    auto o       = gptr(A.out);
    o[O_NST]     = (double)s.nst;
    o[O_NFE]     = (double)s.nfe;
    o[O_NSETUPS] = (double)s.nsetups;
    o[O_NETF]    = (double)s.netf;
    o[O_NNI]     = (double)s.nni;
    o[O_NNF]     = (double)s.nnf;
    o[O_NCFN]    = (double)s.ncfn;
    o[O_NJE]     = (double)s.nje;
    o[O_QU]      = (double)s.qu;
    o[O_NEXTQ]   = (double)s.next_q;
    o[O_H0U]     = s.h0u;
    o[O_HU]      = s.hu;
    o[O_NEXTH]   = s.next_h;
    o[O_TN]      = s.tn;
    o[O_FLAG]    = dead ? -100.0 : (double)flag;
    o[O_BARRIERS] = (double)gen;
  }
}

} // namespace sunamd_cpl

// -----------------------------------------------------------------------------
// host side
// -----------------------------------------------------------------------------
using namespace sunamd_cpl;

namespace
{

struct CplState
{
  int m, R;
  long ncells;
  const double* kf_src; // the problem the buffers were built for
  int* soff;
  int* sent;
  double *zn, *ewt, *acor, *ysav, *part, *seq, *out;
  float* rinv;
  unsigned* bar;
  int* abort;
  CplArgs* args; // the kernel's argument block
  long part_cap, seq_cap;
};

// launch shape of the last run (workgroups, threads per workgroup, LDS bytes)
int g_last_shape[3] = {0, 0, 0};

int hip_err(hipError_t e, const char* what)
{
  fprintf(stderr, "[sundials_amd coupled_device] %s failed: %s\n", what, hipGetErrorString(e));
  fflush(stderr);
  return -100;
}

void free_state(CplState* st)
{
  if (!st) { return; }
  void* ps[] = {st->soff, st->sent, st->zn, st->ewt, st->acor, st->ysav,
                st->part, st->seq, st->out, st->rinv, st->bar, st->abort, st->args};
  for (void* p : ps)
  {
    if (p) { (void)hipFree(p); }
  }
  free(st);
}

template <int M, typename T>
int run_sized(const CoupledDeviceJob& job, CplState* st, SUNAMD_SolveStats* out)
{
  constexpr int G   = pow2ceil(M);
  constexpr int GPW = 64 / G;
  constexpr int CW  = cpl_waves<M>();
  constexpr int CT  = 64 * CW;
  const int R       = job.R;
  const size_t shm  = sizeof(int) * (size_t)(8 * R + ((M + 2) & ~1)) +
                     sizeof(double) * (size_t)CW * cpl_wave_doubles<M, T>(R);
  if (shm > 65536) { return -100; }
  auto kern = k_coupled_bdf<M, T>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)shm);
  if (e != hipSuccess) { return hip_err(e, "shared memory attribute"); }
  int dev = 0, ncu = 0, per_cu = 0;
  e = hipGetDevice(&dev);
  if (e == hipSuccess) { e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev); }
  if (e == hipSuccess) { e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, CT, shm); }
  if (e != hipSuccess) { return hip_err(e, "occupancy query"); }
  if (per_cu < 1) { return -100; }
  const long nchunks = (job.ncells + GPW - 1) / GPW;
  long nwg           = (long)per_cu * ncu;
  const long need    = (nchunks + CW - 1) / CW;
  if (nwg > need) { nwg = need; }
  if (nwg < 1) { nwg = 1; }
  if (nwg > st->part_cap) { return -100; }

  CplArgs A;
  memset(&A, 0, sizeof(A));
  A.ba.R       = R;
  A.ba.m       = M;
  A.m          = M;
  A.R          = R;
  A.ncells     = job.ncells;
  A.N          = job.ncells * M;
  A.nchunks    = nchunks;
  A.nwg        = (int)nwg;
  A.idx        = job.idx;
  A.soff       = st->soff;
  A.sent       = st->sent;
  A.kf         = job.kf;
  A.kb         = job.kb;
  A.y0         = job.y0;
  A.zn         = st->zn;
  A.ewt        = st->ewt;
  A.acor       = st->acor;
  A.ysav       = st->ysav;
  A.lu         = job.lu;
  A.pos        = job.piv;
  A.rinv       = st->rinv;
  A.part       = st->part;
  A.seq        = st->seq;
  A.bar        = st->bar;
  A.abort      = st->abort;
  A.yout       = job.yout;
  A.out        = st->out;
  A.rtol       = job.rtol;
  A.atol       = job.atol;
  A.tout       = job.tout;
  A.mxstep     = job.mxstep;
  A.sequential = job.sequential;

  g_last_shape[0] = (int)nwg;
  g_last_shape[1] = CT;
  g_last_shape[2] = (int)shm;
  e = hipMemsetAsync(st->bar, 0, sizeof(unsigned), job.stream);
  if (e == hipSuccess) { e = hipMemsetAsync(st->abort, 0, sizeof(int), job.stream); }
  if (e == hipSuccess)
  {
    e = hipMemcpyAsync(st->args, &A, sizeof(CplArgs), hipMemcpyHostToDevice, job.stream);
  }
  const CplArgs* ap = st->args;
  void* args[]      = {(void*)&ap};
  if (e == hipSuccess)
  {
    e = hipLaunchCooperativeKernel((const void*)kern, dim3((unsigned)nwg), dim3(CT), args,
                                   (unsigned)shm, job.stream);
  }
  double o[O_COUNT];
  if (e == hipSuccess)
  {
    e = hipMemcpyAsync(o, st->out, sizeof(o), hipMemcpyDeviceToHost, job.stream);
  }
  if (e == hipSuccess) { e = hipStreamSynchronize(job.stream); }
  if (e != hipSuccess) { return hip_err(e, "device-stepped integration"); }
  if (out)
  {
    out->nst     = (long)o[O_NST];
    out->nfe     = (long)o[O_NFE];
    out->nsetups = (long)o[O_NSETUPS];
    out->netf    = (long)o[O_NETF];
    out->nni     = (long)o[O_NNI];
    out->nnf     = (long)o[O_NNF];
    out->ncfn    = (long)o[O_NCFN];
    out->nje     = (long)o[O_NJE];
    out->qlast   = (int)o[O_QU];
    out->qcur    = (int)o[O_NEXTQ];
    out->hinused = o[O_H0U];
    out->hlast   = o[O_HU];
    out->hcur    = o[O_NEXTH];
    out->tcur    = o[O_TN];
    out->launches = 1;
  }
  return (int)o[O_FLAG];
}

} // namespace

#define CPL_SIZES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(10) X(12) X(16) X(24) X(32) X(48) X(64)

int sunamd_coupled_device_run(const CoupledDeviceJob& job, void** state, SUNAMD_SolveStats* out)
{
  const int m = job.m;
  bool sized  = false;
#define CPL_OK(MM) sized = sized || (m == MM);
  CPL_SIZES(CPL_OK)
#undef CPL_OK
  if (!sized || job.ncells < 1 || job.R < 1) { return -100; }
  const long N  = job.ncells * (long)m;
  CplState* st  = (CplState*)*state;
  if (st && (st->m != m || st->R != job.R || st->ncells != job.ncells || st->kf_src != job.kf))
  {
    free_state(st);
    st     = nullptr;
    *state = nullptr;
  }
  hipError_t e = hipSuccess;
  if (!st)
  {
    st = (CplState*)calloc(1, sizeof(CplState));
    if (!st) { return -100; }
    st->m      = m;
    st->R      = job.R;
    st->ncells = job.ncells;
    st->kf_src = job.kf;
    e          = hipMalloc((void**)&st->zn, sizeof(double) * (QMAX + 1) * (size_t)N);
    if (e == hipSuccess) { e = hipMalloc((void**)&st->ewt, sizeof(double) * (size_t)N); }
    if (e == hipSuccess) { e = hipMalloc((void**)&st->acor, sizeof(double) * (size_t)N); }
    if (e == hipSuccess) { e = hipMalloc((void**)&st->ysav, sizeof(double) * (size_t)N); }
    if (e == hipSuccess) { e = hipMalloc((void**)&st->rinv, sizeof(float) * (size_t)N); }
    st->part_cap = 1L << 16;
    if (e == hipSuccess)
    {
      e = hipMalloc((void**)&st->part, sizeof(double) * 2 * NRED * st->part_cap);
    }
    if (e == hipSuccess) { e = hipMalloc((void**)&st->out, sizeof(double) * O_COUNT); }
    if (e == hipSuccess) { e = hipMalloc((void**)&st->bar, sizeof(unsigned)); }
    if (e == hipSuccess) { e = hipMalloc((void**)&st->abort, sizeof(int)); }
    if (e == hipSuccess) { e = hipMalloc((void**)&st->args, sizeof(CplArgs)); }
    if (e == hipSuccess) { e = hipMemset(st->acor, 0, sizeof(double) * (size_t)N); }
    if (e == hipSuccess) { e = hipMemset(st->zn, 0, sizeof(double) * (QMAX + 1) * (size_t)N); }
    // species s's RHS / Jacobian-row terms in reaction order, roles a b c d
    int* hidx = (int*)malloc(sizeof(int) * 4 * (size_t)job.R);
    int* off  = (int*)calloc((size_t)m + 1, sizeof(int));
    int* ent  = (int*)malloc(sizeof(int) * 4 * (size_t)job.R);
    int* fill = (int*)malloc(sizeof(int) * (size_t)m);
    if (!hidx || !off || !ent || !fill) { e = hipErrorOutOfMemory; }
    if (e == hipSuccess)
    {
      e = hipMemcpy(hidx, job.idx, sizeof(int) * 4 * (size_t)job.R, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess)
    {
      for (int t = 0; t < 4 * job.R; t++)
      {
        if (hidx[t] < 0 || hidx[t] >= m) { e = hipErrorInvalidValue; }
      }
    }
    if (e == hipSuccess)
    {
      for (int r = 0; r < job.R; r++)
      {
        for (int t = 0; t < 4; t++) { off[hidx[4 * r + t] + 1]++; }
      }
      for (int q = 0; q < m; q++) { off[q + 1] += off[q]; }
      memcpy(fill, off, sizeof(int) * (size_t)m);
      for (int r = 0; r < job.R; r++)
      {
        for (int t = 0; t < 4; t++) { ent[fill[hidx[4 * r + t]]++] = 4 * r + t; }
      }
      e = hipMalloc((void**)&st->soff, sizeof(int) * ((size_t)m + 1));
      if (e == hipSuccess) { e = hipMalloc((void**)&st->sent, sizeof(int) * 4 * (size_t)job.R); }
      if (e == hipSuccess)
      {
        e = hipMemcpy(st->soff, off, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice);
      }
      if (e == hipSuccess)
      {
        e = hipMemcpy(st->sent, ent, sizeof(int) * 4 * (size_t)job.R, hipMemcpyHostToDevice);
      }
    }
    free(hidx);
    free(off);
    free(ent);
    free(fill);
    if (e == hipSuccess) { e = hipStreamSynchronize(job.stream); }
    if (e != hipSuccess)
    {
      free_state(st);
      return hip_err(e, "device-stepped integrator set-up");
    }
    *state = st;
  }
  if (job.sequential && st->seq_cap < 6 * N)
  {
    if (st->seq) { (void)hipFree(st->seq); }
    st->seq     = nullptr;
    st->seq_cap = 0;
    e           = hipMalloc((void**)&st->seq, sizeof(double) * 6 * (size_t)N);
    if (e != hipSuccess) { return hip_err(e, "sequential-norm terms"); }
    st->seq_cap = 6 * N;
  }
  const bool fp32 = job.lu_precision == SUNAMD_LU_FP32;
  switch (m)
  {
#define CPL_CASE(MM)                                                                       \
  case MM: return fp32 ? run_sized<MM, float>(job, st, out) : run_sized<MM, double>(job, st, out);
    CPL_SIZES(CPL_CASE)
#undef CPL_CASE
  default: break;
  }
  return -100;
}

void sunamd_coupled_device_free(void* state) { free_state((CplState*)state); }

extern "C" int SUNAMD_CoupledDeviceLastShape(int* workgroups, int* threads, int* lds_bytes)
{
  if (workgroups) { *workgroups = g_last_shape[0]; }
  if (threads) { *threads = g_last_shape[1]; }
  if (lds_bytes) { *lds_bytes = g_last_shape[2]; }
  return g_last_shape[0] > 0 ? 0 : -1;
}
