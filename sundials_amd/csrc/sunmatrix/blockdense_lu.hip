// -----------------------------------------------------------------------------
// Block-diagonal dense SUNMatrix + batched LU solver for MI355X (gfx950).
//
// Reference behaviour restated:
//   matrix ops        src/sunmatrix/dense/sunmatrix_dense.c:229-334
//   LU factor/solve   src/sundials/sundials_dense.c:85-186 (GETRF / GETRS)
//   solver wrapper    src/sunlinsol/dense/sunlinsol_dense.c:128-183
//   batched precedent src/sunlinsol/magmadense/sunlinsol_magmadense.cpp:314-440
//
// Kernel design (one m x m block per group of G = pow2ceil(m) lanes, 64/G
// blocks per 64-wide wavefront, one wavefront per workgroup):
//   * lane r holds row r of its block in registers (m values); every column
//     load/store is one coalesced 8*m-byte segment per block;
//   * partial pivoting without moving data: each lane carries the physical
//     row position it currently occupies; the pivot search is a butterfly
//     arg-max over the group keyed on (|a_rk|, -position), which reproduces
//     the reference's "first maximum, strict >" choice in the reference's
//     row order, and a "swap" just exchanges two position tags;
//   * the pivot row is broadcast through LDS; each remaining row scales its
//     multiplier with the reciprocal of the pivot and subtracts
//     pivot_row[j] * multiplier column by column, skipping zero pivot-row
//     entries -- the reference's exact per-element operation sequence, so in
//     FP64 the factors are bit-identical;
//   * rows are written back to their final physical position, so the stored
//     factors and the pivot vector are those of the reference;
//   * GETRS gathers b through the pivot sequence, then runs the unit-lower
//     and upper sweeps with shuffle broadcasts of the solved entry.
// The FP32 variant factors in single precision into a separate 4*m^2-byte
// factor array and solves in FP64 arithmetic from those factors.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "../core/lane_ops.h"
#include "../core/profiler_internal.h"
#include "sunamd/nvector_mi355.h"
#include "sunamd/sunamd_profiler.h"
#include "sunamd/sunmatrix_mi355block.h"

#define EW_BLOCK 256

static void hip_fail(SUNContext ctx, hipError_t e, const char* what, int line)
{
  (void)ctx;
  fprintf(stderr, "[sundials_amd blockdense] %s failed (line %d): %s\n", what, line,
          hipGetErrorString(e));
  fflush(stderr);
}
#define HIPCHK(ctx, call)                                           \
  do {                                                              \
    hipError_t e_ = (call);                                         \
    if (e_ != hipSuccess) { hip_fail((ctx), e_, #call, __LINE__); } \
  } while (0)
#define LAUNCHCHK(ctx) HIPCHK(ctx, hipGetLastError())

// ----------------------------------------------------------------------------
// matrix content
// ----------------------------------------------------------------------------
struct BlockMat
{
  sunindextype nblocks, m;
  double* dev;
  double* host;
  hipStream_t stream;
};
#define BM(A) ((BlockMat*)((A)->content))

// ----------------------------------------------------------------------------
// elementwise matrix kernels
// ----------------------------------------------------------------------------
__global__ void k_fill(double* __restrict__ a, long n, double v)
{
  for (long i = (long)blockIdx.x * EW_BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * EW_BLOCK)
  {
    a[i] = v;
  }
}

// A = c*A + I (per element: a *= c; a += 1 on the diagonal)
__global__ void k_scaleaddi(double* __restrict__ a, long n, long m, double c)
{
  const long mm = m * m;
  for (long i = (long)blockIdx.x * EW_BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * EW_BLOCK)
  {
    long e   = i % mm;
    double v = a[i] * c;
    if (e % (m + 1) == 0) { v = v + 1.0; }
    a[i] = v;
  }
}

// A = c*A + B
__global__ void k_scaleadd(double* __restrict__ a, const double* __restrict__ b, long n, double c)
{
  for (long i = (long)blockIdx.x * EW_BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * EW_BLOCK)
  {
    a[i] = c * a[i] + b[i];
  }
}

// y = A x per block; one thread per row, column order as the reference
__global__ void k_matvec(const double* __restrict__ a, const double* __restrict__ x,
                         double* __restrict__ y, long nblocks, long m)
{
  const long nrows = nblocks * m;
  for (long g = (long)blockIdx.x * EW_BLOCK + threadIdx.x; g < nrows; g += (long)gridDim.x * EW_BLOCK)
  {
    long blk          = g / m;
    long i            = g - blk * m;
    const double* ab  = a + blk * m * m;
    const double* xb  = x + blk * m;
    double acc        = 0.0;
    for (long j = 0; j < m; j++) { acc = acc + ab[j * m + i] * xb[j]; }
    y[g] = acc;
  }
}

static int ew_grid(long n)
{
  long g = (n + EW_BLOCK - 1) / EW_BLOCK;
  if (g > 65536) { g = 65536; }
  if (g < 1) { g = 1; }
  return (int)g;
}

// ----------------------------------------------------------------------------
// batched LU kernels
// ----------------------------------------------------------------------------
__host__ __device__ constexpr int pow2ceil(int m)
{
  return m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : m <= 8 ? 8 : m <= 16 ? 16 : m <= 32 ? 32 : 64;
}

template <typename T>
__device__ __forceinline__ T to_t(double v)
{
  return (T)v;
}

// How k_getrf forms the matrix it factors from its source blocks A:
//   SRC_PLAIN  the blocks themselves;
//   SRC_SCALED c*A + I (SUNMatScaleAddI's per-element a*c, then +1 on the
//              diagonal), A read-only -- the Newton matrix from a saved
//              Jacobian without the copy and the scaling pass;
//   SRC_SAVE   as SRC_SCALED, and A's blocks are also written to `save` first
//              (a fresh Jacobian kept for later setups).
// With SRC_SCALED / SRC_SAVE and FP32 factors, `scaled` receives c*A + I in
// FP64 -- the matrix the separate copy / scale / factor calls leave in the
// solver's matrix; with FP64 factors the factors themselves go there.
enum GetrfSource
{
  SRC_PLAIN  = 0,
  SRC_SCALED = 1,
  SRC_SAVE   = 2
};

// wavefronts per workgroup of the register-row factorisation (each wavefront
// factors its own blocks; workgroups of several wavefronts put more blocks in
// flight per CU than one-wavefront workgroups)
#define LU_WAVES 4

// Factor nblocks m x m blocks in place (T = double) or into an FP32 factor
// array (T = float).  info receives min over failing blocks of
// (block*m + k + 1) for the first zero pivot k.  Lane r of a block's group of
// G = pow2ceil(m) lanes holds row r; the row swaps exchange position tags
// only, and the rows are written to their final positions at the end.
template <int M, typename T, int SRC = SRC_PLAIN>
__global__ void __launch_bounds__(64 * LU_WAVES) k_getrf(const double* A, T* F /* may alias A */,
                                                         int* __restrict__ piv, long nblocks,
                                                         unsigned long long* info, double c = 0.0,
                                                         double* save = nullptr,
                                                         double* scaled = nullptr)
{
  using namespace sunamd_lane;
  constexpr int G   = pow2ceil(M);
  constexpr int BPW = 64 / G;
  __shared__ T prow[LU_WAVES * BPW][M];

  const int lane   = threadIdx.x & 63;
  const int slot   = (threadIdx.x >> 6) * BPW + lane / G;
  const int r      = lane % G;
  const long blk   = (long)blockIdx.x * (LU_WAVES * BPW) + slot;
  const bool valid = (blk < nblocks) && (r < M);

  // every lane loads from inside block 0 or its own block (rows past M read
  // row M-1), so the loads are unconditional and in flight together; idle
  // lanes' values are replaced afterwards (a load under the validity test
  // compiles to a branch per element, which waits for its load wherever a
  // conversion or the scaling follows it).  FP32 factors, and the scaled
  // forms at m > 32, take the FP64 source in chunks of 16 columns (a
  // scheduling barrier between chunks), so the loads in flight do not cost the
  // kernel its occupancy.
  const long blk_s = (blk < nblocks) ? blk : 0;
  const int r_s    = (r < M) ? r : M - 1;
  const double* Ab = A + blk_s * (long)M * M;
  const long o0    = blk * (long)M * M + r;
  constexpr int CH = ((sizeof(T) == 4 && M > 16) || (SRC != SRC_PLAIN && M > 32)) ? 16 : M;
  T a[M];
#pragma unroll
  for (int j0 = 0; j0 < M; j0 += CH)
  {
    double raw[CH];
#pragma unroll
    for (int q = 0; q < CH; q++)
    {
      if (j0 + q < M) { raw[q] = Ab[(j0 + q) * M + r_s]; }
    }
#pragma unroll
    for (int q = 0; q < CH; q++)
    {
      const int j = j0 + q;
      if (j >= M) { break; }
      if (SRC == SRC_PLAIN) { a[j] = valid ? to_t<T>(raw[q]) : (T)0; }
      else
      {
        double v = valid ? raw[q] : 0.0;
        if (SRC == SRC_SAVE && valid) { save[o0 + (long)j * M] = v; }
        v = v * c;
        if (j == r) { v = v + 1.0; }
        if (!std::is_same<T, double>::value && valid) { scaled[o0 + (long)j * M] = v; }
        a[j] = valid ? to_t<T>(v) : (T)0;
      }
      // the single-precision value is formed here, not where it is first
      // used (the compiler would otherwise keep the double until then)
      if constexpr (sizeof(T) == 4) { asm volatile("" : "+v"(a[j])); }
    }
    if constexpr (CH < M) { __builtin_amdgcn_sched_barrier(0); }
  }

  int pos  = r; // physical row currently held by this lane
  int myp  = 0; // pivot row chosen at step k == r
  int fail = 0;

#pragma unroll
  for (int k = 0; k < M; k++)
  {
    // arg-max of |a_rk| over rows at positions >= k; ties -> lowest position
    T v    = (valid && pos >= k) ? (T)fabs((double)a[k]) : (T)-1;
    int bp = valid ? pos : INT_MAX;
    g_argmax<G>(v, bp);
    if (r == k) { myp = bp; }
    if (v == (T)0 && fail == 0) { fail = k + 1; }

    // exchange the position tags of rows k and bp
    if (pos == k) { pos = bp; }
    else if (pos == bp) { pos = k; }

    const bool up = valid && pos > k && fail == 0;
    if constexpr (G == 64)
    {
      // one block per wavefront: the pivot row's lane is wavefront-uniform,
      // its entries are read straight from that lane's registers
      const unsigned long long holder = __ballot(valid && pos == k);
      const int pl = __builtin_amdgcn_readfirstlane(__ffsll((long long)holder) - 1) & 63;
      const T pk   = readlane_d(a[k], pl);
      T l          = (T)0;
      if (up)
      {
        const T mult = (T)1 / pk;
        l            = a[k] * mult;
        a[k]         = l;
      }
#pragma unroll
      for (int j = k + 1; j < M; j++)
      {
        const T pj = readlane_d(a[j], pl);
        if (up && pj != (T)0) { a[j] = a[j] - pj * l; }
      }
    }
    else
    {
      // the pivot row (now at position k) publishes columns k..M-1
      wave_sync();
      if (valid && pos == k)
      {
#pragma unroll
        for (int j = k; j < M; j++) { prow[slot][j] = a[j]; }
      }
      wave_sync();
      if (up)
      {
        const T mult = (T)1 / prow[slot][k];
        const T l    = a[k] * mult;
        a[k]         = l;
#pragma unroll
        for (int j = k + 1; j < M; j++)
        {
          const T pj = prow[slot][j];
          if (pj != (T)0) { a[j] = a[j] - pj * l; }
        }
      }
    }
  }

  if (valid)
  {
    T* Fb = F + blk * (long)M * M;
#pragma unroll
    for (int j = 0; j < M; j++) { Fb[j * M + pos] = a[j]; }
    piv[blk * M + r] = myp;
    if (r == 0 && fail != 0)
    {
      atomicMin(info, (unsigned long long)(blk * M + fail));
    }
  }
}

// x = LU^{-1} P b per block (x may alias b)
template <int M, typename T>
__global__ void __launch_bounds__(64) k_getrs(const T* __restrict__ F, const int* __restrict__ piv,
                                              const double* b, double* x, long nblocks,
                                              double scale = 1.0)
{
  constexpr int G   = pow2ceil(M);
  constexpr int BPW = 64 / G;
  const int lane    = threadIdx.x;
  const int grp     = lane / G;
  const int r       = lane % G;
  const long blk    = (long)blockIdx.x * BPW + grp;
  const bool valid  = (blk < nblocks) && (r < M);
  const int base    = grp * G;

  // every lane loads from inside block 0 or its own block (rows past M read
  // row M-1), so the loads are unconditional and all in flight at once; the
  // values of idle lanes are then replaced by zeros (a load under the
  // validity test would be a branch per element, and for FP32 factors the
  // conversion made each one wait for its load)
  const long blk_s = (blk < nblocks) ? blk : 0;
  const int r_s    = (r < M) ? r : M - 1;
  const T* Fb      = F + blk_s * (long)M * M;
  T raw[M];
#pragma unroll
  for (int j = 0; j < M; j++) { raw[j] = Fb[j * M + r_s]; }
  const int* p = piv + blk_s * M;
  int pv[M];
#pragma unroll
  for (int k = 0; k < M; k++) { pv[k] = p[k]; }
  double lu[M];
#pragma unroll
  for (int j = 0; j < M; j++) { lu[j] = valid ? (double)raw[j] : 0.0; }

  // source position of b after the sequential swaps b[k] <-> b[p[k]]
  int idx = r_s;
#pragma unroll
  for (int k = M - 1; k >= 0; k--)
  {
    const int pk = pv[k];
    if (idx == k) { idx = pk; }
    else if (idx == pk) { idx = k; }
  }
  // (pivots of an unfactored block are not trusted to stay in range)
  const double b_s = b[blk_s * M + (((unsigned)idx < (unsigned)M) ? idx : 0)];
  double bi        = valid ? b_s : 0.0;

  // L y = P b (unit lower)
#pragma unroll
  for (int k = 0; k < M - 1; k++)
  {
    double bk = __shfl(bi, base + k, 64);
    if (r > k) { bi = bi - lu[k] * bk; }
  }
  // U x = y
#pragma unroll
  for (int k = M - 1; k > 0; k--)
  {
    if (r == k) { bi = bi / lu[k]; }
    double bk = __shfl(bi, base + k, 64);
    if (r < k) { bi = bi - lu[k] * bk; }
  }
  if (r == 0) { bi = bi / lu[0]; }
  // scale * x is N_VScale's per-element product (exact for scale = 1)
  if (valid) { x[blk * M + r] = scale * bi; }
}

// generic-size fallback (any m <= 64): block staged in LDS, same arithmetic
template <typename T>
__global__ void __launch_bounds__(64) k_getrf_lds(const double* A, T* F /* may alias A */,
                                                  int* __restrict__ piv, long nblocks, int M,
                                                  unsigned long long* info)
{
  extern __shared__ unsigned char smem_raw[];
  T* s         = reinterpret_cast<T*>(smem_raw); // M*M, column-major, physical rows
  const long blk = blockIdx.x;
  const int r    = threadIdx.x;
  if (blk >= nblocks) { return; }
  const double* Ab = A + blk * (long)M * M;
  for (int e = r; e < M * M; e += 64) { s[e] = to_t<T>(Ab[e]); }
  __syncthreads();
  __shared__ int sp[64];
  __shared__ int sfail;
  if (r == 0)
  {
    sfail = 0;
    for (int k = 0; k < M; k++)
    {
      int l = k;
      for (int i = k + 1; i < M; i++)
      {
        if (fabs((double)s[k * M + i]) > fabs((double)s[k * M + l])) { l = i; }
      }
      sp[k] = l;
      if (s[k * M + l] == (T)0)
      {
        sfail = k + 1;
        break;
      }
      if (l != k)
      {
        for (int j = 0; j < M; j++)
        {
          T t          = s[j * M + l];
          s[j * M + l] = s[j * M + k];
          s[j * M + k] = t;
        }
      }
      T mult = (T)1 / s[k * M + k];
      for (int i = k + 1; i < M; i++) { s[k * M + i] *= mult; }
      for (int j = k + 1; j < M; j++)
      {
        T akj = s[j * M + k];
        if (akj != (T)0)
        {
          for (int i = k + 1; i < M; i++) { s[j * M + i] = s[j * M + i] - akj * s[k * M + i]; }
        }
      }
    }
  }
  __syncthreads();
  T* Fb = F + blk * (long)M * M;
  for (int e = r; e < M * M; e += 64) { Fb[e] = s[e]; }
  if (r < M) { piv[blk * M + r] = sp[r]; }
  if (r == 0 && sfail) { atomicMin(info, (unsigned long long)(blk * M + sfail)); }
}

template <typename T>
__global__ void __launch_bounds__(64) k_getrs_lds(const T* __restrict__ F, const int* __restrict__ piv,
                                                  const double* b, double* x, long nblocks, int M)
{
  __shared__ double sb[64];
  const long blk = blockIdx.x;
  const int r    = threadIdx.x;
  if (blk >= nblocks) { return; }
  if (r < M) { sb[r] = b[blk * M + r]; }
  __syncthreads();
  if (r == 0)
  {
    const T* Fb  = F + blk * (long)M * M;
    const int* p = piv + blk * M;
    for (int k = 0; k < M; k++)
    {
      int pk = p[k];
      if (pk != k)
      {
        double t = sb[k];
        sb[k]    = sb[pk];
        sb[pk]   = t;
      }
    }
    for (int k = 0; k < M - 1; k++)
    {
      for (int i = k + 1; i < M; i++) { sb[i] = sb[i] - (double)Fb[k * M + i] * sb[k]; }
    }
    for (int k = M - 1; k > 0; k--)
    {
      sb[k] = sb[k] / (double)Fb[k * M + k];
      for (int i = 0; i < k; i++) { sb[i] = sb[i] - (double)Fb[k * M + i] * sb[k]; }
    }
    sb[0] = sb[0] / (double)Fb[0];
  }
  __syncthreads();
  if (r < M) { x[blk * M + r] = sb[r]; }
}

__global__ void k_info_reset(unsigned long long* info) { *info = ULLONG_MAX; }

// ----------------------------------------------------------------------------
// FP32 blocked LU with an MFMA trailing update (blocks of 32 and 64)
// ----------------------------------------------------------------------------
// One block per wavefront, staged in LDS row-major (stride M + 1 floats).
// Right-looking with panels of 16 columns: the panel is factored column by
// column (partial pivoting with physical row swaps across the whole block,
// the reference's choice of the first largest |a_ik|), the panel's pivot
// rows get their trailing columns by forward substitution with the panel's
// unit lower triangle (U12), and the trailing block is updated
// A22 -= L21 U12 with v_mfma_f32_16x16x4f32 tiles: lane l supplies
// L21[row l%16][k + l/16] and U12[k + l/16][column l%16] and holds rows
// 4(l/16) .. 4(l/16)+3 of column l%16 of each 16 x 16 tile.  Factors and
// pivots are stored as the register kernel stores them, so the FP32 solve is
// shared.
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int M>
__global__ void __launch_bounds__(64) k_getrf_mfma(const double* __restrict__ A,
                                                   float* __restrict__ F, int* __restrict__ piv,
                                                   long nblocks, unsigned long long* info)
{
  constexpr int NB = 16, LD = M + 1;
  __shared__ float S[M * LD];
  const int t    = threadIdx.x;
  const long blk = blockIdx.x;
  if (blk >= nblocks) { return; } // whole workgroup
  const double* Ab = A + blk * (long)M * M;
  for (int e = t; e < M * M; e += 64) { S[(e % M) * LD + e / M] = (float)Ab[e]; }
  __syncthreads();
  int fail = 0;
  int* pb  = piv + blk * M;
  for (int k0 = 0; k0 < M; k0 += NB)
  {
    // panel: columns k0 .. k0+NB-1, lane t = row t
    for (int k = k0; k < k0 + NB; k++)
    {
      const bool cand = (t < M) && (t >= k);
      float v         = cand ? fabsf(S[t * LD + k]) : -1.0f;
      int p           = cand ? t : INT_MAX;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1)
      {
        const float ov = __shfl_xor(v, off, 64);
        const int op   = __shfl_xor(p, off, 64);
        if (ov > v || (ov == v && op < p))
        {
          v = ov;
          p = op;
        }
      }
      if (p < k || p >= M) { p = k; } // no comparable candidate (NaN column): no swap
      if (t == 0) { pb[k] = p; }
      if (!(v > 0.0f) && fail == 0) { fail = k + 1; }
      __syncthreads();
      if (p != k && t < M)
      {
        const float a = S[k * LD + t], b = S[p * LD + t];
        S[k * LD + t] = b;
        S[p * LD + t] = a;
      }
      __syncthreads();
      if (fail == 0 && t > k && t < M)
      {
        const float l = S[t * LD + k] * (1.0f / S[k * LD + k]);
        S[t * LD + k] = l;
        for (int j = k + 1; j < k0 + NB; j++) { S[t * LD + j] = fmaf(-l, S[k * LD + j], S[t * LD + j]); }
      }
      __syncthreads();
    }
    if (k0 + NB >= M) { break; }
    // U12: the pivot rows' trailing columns (lane t = column k0 + NB + t)
    const int j = k0 + NB + t;
    if (fail == 0 && j < M)
    {
      for (int k = k0; k < k0 + NB; k++)
      {
        const float ukj = S[k * LD + j];
        for (int i = k + 1; i < k0 + NB; i++) { S[i * LD + j] = fmaf(-S[i * LD + k], ukj, S[i * LD + j]); }
      }
    }
    __syncthreads();
    // A22 -= L21 U12, one 16 x 16 tile at a time
    if (fail == 0)
    {
      const int r = t & 15, q = t >> 4;
      const int n2 = (M - k0 - NB) / 16;
      for (int ti = 0; ti < n2; ti++)
      {
        for (int tj = 0; tj < n2; tj++)
        {
          const int i0 = k0 + NB + 16 * ti, j0 = k0 + NB + 16 * tj;
          f32x4_t acc  = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int kk = 0; kk < NB; kk += 4)
          {
            const float a = S[(i0 + r) * LD + k0 + kk + q];
            const float b = S[(k0 + kk + q) * LD + j0 + r];
            acc           = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
          }
#pragma unroll
          for (int e = 0; e < 4; e++)
          {
            float* d = &S[(i0 + 4 * q + e) * LD + j0 + r];
            *d       = *d - acc[e];
          }
        }
      }
    }
    __syncthreads();
  }
  float* Fb = F + blk * (long)M * M;
  for (int e = t; e < M * M; e += 64) { Fb[e] = S[(e % M) * LD + e / M]; }
  if (t == 0 && fail != 0) { atomicMin(info, (unsigned long long)(blk * M + fail)); }
}

// ----------------------------------------------------------------------------
// launch dispatch over block sizes with register-resident kernels
// ----------------------------------------------------------------------------
#define LU_SIZES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(12) X(16) X(24) X(32) X(48) X(64)

template <typename T>
static bool launch_getrf(int m, const double* A, T* F, int* piv, long nblocks,
                         unsigned long long* info, hipStream_t s)
{
  switch (m)
  {
#define GETRF_CASE(MM)                                                                     \
  case MM:                                                                                 \
  {                                                                                        \
    constexpr int BPG = LU_WAVES * (64 / pow2ceil(MM));                                    \
    long g            = (nblocks + BPG - 1) / BPG;                                         \
    k_getrf<MM, T><<<(unsigned)g, 64 * LU_WAVES, 0, s>>>(A, F, piv, nblocks, info);        \
    return true;                                                                           \
  }
    LU_SIZES(GETRF_CASE)
#undef GETRF_CASE
  default: break;
  }
  if (m > 64) { return false; }
  k_getrf_lds<T><<<(unsigned)nblocks, 64, sizeof(T) * m * m, s>>>(A, F, piv, nblocks, m, info);
  return true;
}

// The register-resident sizes factor c*src + I in one pass (SRC_SCALED, or
// SRC_SAVE with src also copied to save); false for other sizes, which take
// the separate copy / scale / factor calls.
template <typename T>
static bool launch_getrf_scaled(int m, bool save_src, const double* src, double c, double* save,
                                double* scaled, T* F, int* piv, long nblocks,
                                unsigned long long* info, hipStream_t s)
{
  switch (m)
  {
#define GETRF_SC_CASE(MM)                                                                  \
  case MM:                                                                                 \
  {                                                                                        \
    constexpr int BPG = LU_WAVES * (64 / pow2ceil(MM));                                    \
    long g            = (nblocks + BPG - 1) / BPG;                                         \
    if (save_src)                                                                          \
    {                                                                                      \
      k_getrf<MM, T, SRC_SAVE><<<(unsigned)g, 64 * LU_WAVES, 0, s>>>(                      \
        src, F, piv, nblocks, info, c, save, scaled);                                      \
    }                                                                                      \
    else                                                                                   \
    {                                                                                      \
      k_getrf<MM, T, SRC_SCALED><<<(unsigned)g, 64 * LU_WAVES, 0, s>>>(                    \
        src, F, piv, nblocks, info, c, save, scaled);                                      \
    }                                                                                      \
    return true;                                                                           \
  }
    LU_SIZES(GETRF_SC_CASE)
#undef GETRF_SC_CASE
  default: break;
  }
  return false;
}

template <typename T>
static bool launch_getrs(int m, const T* F, const int* piv, const double* b, double* x,
                         long nblocks, hipStream_t s, double scale = 1.0)
{
  switch (m)
  {
#define GETRS_CASE(MM)                                                                     \
  case MM:                                                                                 \
  {                                                                                        \
    constexpr int BPW = 64 / pow2ceil(MM);                                                 \
    long g            = (nblocks + BPW - 1) / BPW;                                         \
    k_getrs<MM, T><<<(unsigned)g, 64, 0, s>>>(F, piv, b, x, nblocks, scale);               \
    return true;                                                                           \
  }
    LU_SIZES(GETRS_CASE)
#undef GETRS_CASE
  default: break;
  }
  if (m > 64 || scale != 1.0) { return false; }
  k_getrs_lds<T><<<(unsigned)nblocks, 64, 0, s>>>(F, piv, b, x, nblocks, m);
  return true;
}

// ----------------------------------------------------------------------------
// SUNMatrix
// ----------------------------------------------------------------------------
static SUNMatrix_ID bm_getid(SUNMatrix A)
{
  (void)A;
  return SUNMATRIX_CUSTOM;
}

static void bm_destroy(SUNMatrix A)
{
  if (!A) { return; }
  if (A->content)
  {
    BlockMat* c = BM(A);
    if (c->dev)
    {
      (void)hipStreamSynchronize(c->stream);
      (void)hipFree(c->dev);
    }
    free(c->host);
    free(c);
    A->content = NULL;
  }
  SUNMatFreeEmpty(A);
}

static SUNMatrix bm_clone(SUNMatrix A)
{
  return SUNMatrix_Mi355BlockDense(BM(A)->nblocks, BM(A)->m, (void*)BM(A)->stream, A->sunctx);
}

static long bm_len(SUNMatrix A) { return (long)(BM(A)->nblocks * BM(A)->m * BM(A)->m); }

static SUNErrCode bm_zero(SUNMatrix A)
{
  long n = bm_len(A);
  k_fill<<<ew_grid(n), EW_BLOCK, 0, BM(A)->stream>>>(BM(A)->dev, n, 0.0);
  LAUNCHCHK(A->sunctx);
  return SUN_SUCCESS;
}

static bool compatible(SUNMatrix A, SUNMatrix B)
{
  return B && B->ops->getid == bm_getid && BM(A)->nblocks == BM(B)->nblocks &&
         BM(A)->m == BM(B)->m;
}

static SUNErrCode bm_copy(SUNMatrix A, SUNMatrix B)
{
  if (!compatible(A, B)) { return SUN_ERR_ARG_DIMSMISMATCH; }
  HIPCHK(A->sunctx, hipMemcpyAsync(BM(B)->dev, BM(A)->dev, sizeof(double) * bm_len(A),
                                   hipMemcpyDeviceToDevice, BM(A)->stream));
  return SUN_SUCCESS;
}

static SUNErrCode bm_scaleaddi(sunrealtype c, SUNMatrix A)
{
  long n = bm_len(A);
  k_scaleaddi<<<ew_grid(n), EW_BLOCK, 0, BM(A)->stream>>>(BM(A)->dev, n, (long)BM(A)->m, c);
  LAUNCHCHK(A->sunctx);
  return SUN_SUCCESS;
}

static SUNErrCode bm_scaleadd(sunrealtype c, SUNMatrix A, SUNMatrix B)
{
  if (!compatible(A, B)) { return SUN_ERR_ARG_DIMSMISMATCH; }
  long n = bm_len(A);
  k_scaleadd<<<ew_grid(n), EW_BLOCK, 0, BM(A)->stream>>>(BM(A)->dev, BM(B)->dev, n, c);
  LAUNCHCHK(A->sunctx);
  return SUN_SUCCESS;
}

static SUNErrCode bm_matvec(SUNMatrix A, N_Vector x, N_Vector y)
{
  if (N_VGetLength(x) != BM(A)->nblocks * BM(A)->m ||
      N_VGetLength(y) != BM(A)->nblocks * BM(A)->m)
  {
    return SUN_ERR_ARG_DIMSMISMATCH;
  }
  double* xd = N_VGetDeviceArrayPointer(x);
  double* yd = N_VGetDeviceArrayPointer(y);
  if (!xd || !yd || xd == yd) { return SUN_ERR_ARG_INCOMPATIBLE; }
  long rows = (long)(BM(A)->nblocks * BM(A)->m);
  k_matvec<<<ew_grid(rows), EW_BLOCK, 0, BM(A)->stream>>>(BM(A)->dev, xd, yd,
                                                         (long)BM(A)->nblocks, (long)BM(A)->m);
  LAUNCHCHK(A->sunctx);
  return SUN_SUCCESS;
}

static SUNErrCode bm_space(SUNMatrix A, long int* lenrw, long int* leniw)
{
  *lenrw = bm_len(A);
  *leniw = 2;
  return SUN_SUCCESS;
}

extern "C" SUNMatrix SUNMatrix_Mi355BlockDense(sunindextype nblocks, sunindextype m,
                                               void* stream, SUNContext sunctx)
{
  if (nblocks <= 0 || m <= 0 || !sunctx) { return NULL; }
  SUNMatrix A = SUNMatNewEmpty(sunctx);
  if (!A) { return NULL; }
  A->ops->getid     = bm_getid;
  A->ops->clone     = bm_clone;
  A->ops->destroy   = bm_destroy;
  A->ops->zero      = bm_zero;
  A->ops->copy      = bm_copy;
  A->ops->scaleadd  = bm_scaleadd;
  A->ops->scaleaddi = bm_scaleaddi;
  A->ops->matvec    = bm_matvec;
  A->ops->space     = bm_space;
  BlockMat* c       = (BlockMat*)calloc(1, sizeof(BlockMat));
  if (!c)
  {
    SUNMatFreeEmpty(A);
    return NULL;
  }
  A->content = c;
  c->nblocks = nblocks;
  c->m       = m;
  c->stream  = (hipStream_t)stream;
  hipError_t e = hipMalloc((void**)&c->dev, sizeof(double) * (size_t)(nblocks * m * m));
  if (e != hipSuccess)
  {
    hip_fail(sunctx, e, "hipMalloc(block matrix)", __LINE__);
    c->dev = NULL;
    bm_destroy(A);
    return NULL;
  }
  return A;
}

extern "C" sunindextype SUNMatrix_Mi355BlockDense_NumBlocks(SUNMatrix A) { return BM(A)->nblocks; }
extern "C" sunindextype SUNMatrix_Mi355BlockDense_BlockSize(SUNMatrix A) { return BM(A)->m; }
extern "C" sunrealtype* SUNMatrix_Mi355BlockDense_DeviceData(SUNMatrix A) { return BM(A)->dev; }

extern "C" sunrealtype* SUNMatrix_Mi355BlockDense_HostData(SUNMatrix A)
{
  if (!BM(A)->host)
  {
    BM(A)->host = (double*)calloc((size_t)bm_len(A), sizeof(double));
  }
  return BM(A)->host;
}

extern "C" SUNErrCode SUNMatrix_Mi355BlockDense_CopyToDevice(SUNMatrix A)
{
  double* h = SUNMatrix_Mi355BlockDense_HostData(A);
  if (!h) { return SUN_ERR_MALLOC_FAIL; }
  hipError_t e = hipMemcpyAsync(BM(A)->dev, h, sizeof(double) * bm_len(A),
                                hipMemcpyHostToDevice, BM(A)->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(BM(A)->stream); }
  if (e != hipSuccess)
  {
    hip_fail(A->sunctx, e, "matrix copy to device", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  return SUN_SUCCESS;
}

extern "C" SUNErrCode SUNMatrix_Mi355BlockDense_CopyFromDevice(SUNMatrix A)
{
  double* h = SUNMatrix_Mi355BlockDense_HostData(A);
  if (!h) { return SUN_ERR_MALLOC_FAIL; }
  hipError_t e = hipMemcpyAsync(h, BM(A)->dev, sizeof(double) * bm_len(A),
                                hipMemcpyDeviceToHost, BM(A)->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(BM(A)->stream); }
  if (e != hipSuccess)
  {
    hip_fail(A->sunctx, e, "matrix copy from device", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  return SUN_SUCCESS;
}

// ----------------------------------------------------------------------------
// SUNLinearSolver
// ----------------------------------------------------------------------------
struct BlockLU
{
  sunindextype nblocks, m;
  int precision;
  int mfma;                  // FP32 factors by the MFMA blocked kernel (m = 32, 64)
  int* piv;                  // device, nblocks*m
  float* f32;                // device FP32 factors (FP32 mode)
  unsigned long long* dinfo; // device
  unsigned long long* hinfo; // pinned
  sunindextype last_flag;
  hipStream_t stream;
};
#define LU(S) ((BlockLU*)((S)->content))

static SUNLinearSolver_Type lu_gettype(SUNLinearSolver S)
{
  (void)S;
  return SUNLINEARSOLVER_DIRECT;
}
static SUNLinearSolver_ID lu_getid(SUNLinearSolver S)
{
  (void)S;
  return SUNLINEARSOLVER_CUSTOM;
}
static SUNErrCode lu_initialize(SUNLinearSolver S)
{
  LU(S)->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}

static int lu_finish(SUNLinearSolver S);

static int lu_setup(SUNLinearSolver S, SUNMatrix A)
{
  BlockLU* c = LU(S);
  if (A->ops->getid != bm_getid || BM(A)->nblocks != c->nblocks || BM(A)->m != c->m)
  {
    return SUN_ERR_ARG_INCOMPATIBLE;
  }
  k_info_reset<<<1, 1, 0, c->stream>>>(c->dinfo);
  bool ok;
  sunamd_prof_begin(SUNAMD_K_GETRF, c->stream);
  if (c->precision == SUNAMD_LU_FP32 && c->mfma)
  {
    ok = true;
    if (c->m == 32)
    {
      k_getrf_mfma<32><<<(unsigned)c->nblocks, 64, 0, c->stream>>>(BM(A)->dev, c->f32, c->piv,
                                                                   (long)c->nblocks, c->dinfo);
    }
    else
    {
      k_getrf_mfma<64><<<(unsigned)c->nblocks, 64, 0, c->stream>>>(BM(A)->dev, c->f32, c->piv,
                                                                   (long)c->nblocks, c->dinfo);
    }
  }
  else if (c->precision == SUNAMD_LU_FP32)
  {
    ok = launch_getrf<float>((int)c->m, BM(A)->dev, c->f32, c->piv, (long)c->nblocks, c->dinfo,
                             c->stream);
  }
  else
  {
    ok = launch_getrf<double>((int)c->m, BM(A)->dev, BM(A)->dev, c->piv, (long)c->nblocks,
                              c->dinfo, c->stream);
  }
  sunamd_prof_end(SUNAMD_K_GETRF, c->stream);
  if (!ok) { return SUN_ERR_ARG_OUTOFRANGE; }
  return lu_finish(S);
}

// the end of a setup: fetch the factorisation's zero-pivot report
static int lu_finish(SUNLinearSolver S)
{
  BlockLU* c   = LU(S);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
  {
    e = hipMemcpyAsync(c->hinfo, c->dinfo, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       c->stream);
  }
  if (e == hipSuccess) { e = hipStreamSynchronize(c->stream); }
  if (e != hipSuccess)
  {
    hip_fail(S->sunctx, e, "block LU setup", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  if (*c->hinfo != ULLONG_MAX)
  {
    c->last_flag = (sunindextype)(*c->hinfo);
    return SUNLS_LUFACT_FAIL;
  }
  c->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}

static int lu_solve(SUNLinearSolver S, SUNMatrix A, N_Vector x, N_Vector b, sunrealtype tol)
{
  (void)tol;
  BlockLU* c = LU(S);
  // the solve runs on the matrix's stream: vectors on another stream would
  // race with it
  if (!SUNAMD_NVectorIsMi355(x) || !SUNAMD_NVectorIsMi355(b) ||
      (hipStream_t)N_VGetStream_Mi355(x) != c->stream ||
      (hipStream_t)N_VGetStream_Mi355(b) != c->stream)
  {
    c->last_flag = SUN_ERR_ARG_INCOMPATIBLE;
    return SUN_ERR_ARG_INCOMPATIBLE;
  }
  double* xd = N_VGetDeviceArrayPointer(x);
  double* bd = N_VGetDeviceArrayPointer(b);
  if (!xd || !bd) { return SUN_ERR_ARG_CORRUPT; }
  bool ok;
  sunamd_prof_begin(SUNAMD_K_GETRS, c->stream);
  if (c->precision == SUNAMD_LU_FP32)
  {
    ok = launch_getrs<float>((int)c->m, c->f32, c->piv, bd, xd, (long)c->nblocks, c->stream);
  }
  else
  {
    ok = launch_getrs<double>((int)c->m, BM(A)->dev, c->piv, bd, xd, (long)c->nblocks,
                              c->stream);
  }
  sunamd_prof_end(SUNAMD_K_GETRS, c->stream);
  if (!ok) { return SUN_ERR_ARG_OUTOFRANGE; }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
  {
    hip_fail(S->sunctx, e, "block LU solve", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  c->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}

static sunindextype lu_lastflag(SUNLinearSolver S) { return LU(S)->last_flag; }

static SUNErrCode lu_space(SUNLinearSolver S, long int* lenrw, long int* leniw)
{
  *lenrw = (LU(S)->precision == SUNAMD_LU_FP32) ? (long)(LU(S)->nblocks * LU(S)->m * LU(S)->m / 2) : 0;
  *leniw = (long)(LU(S)->nblocks * LU(S)->m);
  return SUN_SUCCESS;
}

static SUNErrCode lu_free(SUNLinearSolver S)
{
  if (!S) { return SUN_SUCCESS; }
  if (S->content)
  {
    BlockLU* c = LU(S);
    (void)hipStreamSynchronize(c->stream);
    if (c->piv) { (void)hipFree(c->piv); }
    if (c->f32) { (void)hipFree(c->f32); }
    if (c->dinfo) { (void)hipFree(c->dinfo); }
    if (c->hinfo) { (void)hipHostFree(c->hinfo); }
    free(c);
    S->content = NULL;
  }
  SUNLinSolFreeEmpty(S);
  return SUN_SUCCESS;
}

extern "C" SUNLinearSolver SUNLinSol_Mi355BlockLU(N_Vector y, SUNMatrix A, int precision,
                                                  SUNContext sunctx)
{
  if (!A || A->ops->getid != bm_getid) { return NULL; }
  if (precision != SUNAMD_LU_FP64 && precision != SUNAMD_LU_FP32) { return NULL; }
  if (N_VGetLength(y) != BM(A)->nblocks * BM(A)->m) { return NULL; }
  if (BM(A)->m > 64) { return NULL; }
  // factor and solves are ordered with the vector ops only on one stream
  if (!SUNAMD_NVectorIsMi355(y) || (hipStream_t)N_VGetStream_Mi355(y) != BM(A)->stream)
  {
    return NULL;
  }
  SUNLinearSolver S = SUNLinSolNewEmpty(sunctx);
  if (!S) { return NULL; }
  S->ops->gettype    = lu_gettype;
  S->ops->getid      = lu_getid;
  S->ops->initialize = lu_initialize;
  S->ops->setup      = lu_setup;
  S->ops->solve      = lu_solve;
  S->ops->lastflag   = lu_lastflag;
  S->ops->space      = lu_space;
  S->ops->free       = lu_free;
  BlockLU* c         = (BlockLU*)calloc(1, sizeof(BlockLU));
  if (!c)
  {
    SUNLinSolFreeEmpty(S);
    return NULL;
  }
  S->content   = c;
  c->nblocks   = BM(A)->nblocks;
  c->m         = BM(A)->m;
  c->precision = precision;
  c->stream    = BM(A)->stream;
  hipError_t e = hipMalloc((void**)&c->piv, sizeof(int) * (size_t)(c->nblocks * c->m));
  if (e == hipSuccess && precision == SUNAMD_LU_FP32)
  {
    e = hipMalloc((void**)&c->f32, sizeof(float) * (size_t)(c->nblocks * c->m * c->m));
  }
  if (e == hipSuccess) { e = hipMalloc((void**)&c->dinfo, sizeof(unsigned long long)); }
  if (e == hipSuccess)
  {
    e = hipHostMalloc((void**)&c->hinfo, sizeof(unsigned long long), hipHostMallocDefault);
  }
  if (e != hipSuccess)
  {
    hip_fail(sunctx, e, "block LU allocation", __LINE__);
    lu_free(S);
    return NULL;
  }
  return S;
}

extern "C" SUNErrCode SUNLinSol_Mi355BlockLUSetMFMA(SUNLinearSolver S, int onoff)
{
  if (!S || !S->content) { return SUN_ERR_ARG_CORRUPT; }
  BlockLU* c = LU(S);
  if (onoff && (c->precision != SUNAMD_LU_FP32 || (c->m != 32 && c->m != 64)))
  {
    return SUN_ERR_ARG_INCOMPATIBLE;
  }
  if (onoff && c->nblocks > 0x7fffffffL) { return SUN_ERR_ARG_OUTOFRANGE; }
  c->mfma = onoff ? 1 : 0;
  return SUN_SUCCESS;
}

extern "C" int* SUNLinSol_Mi355BlockLU_DevicePivots(SUNLinearSolver S) { return LU(S)->piv; }

// One pass instead of three: SUNMatCopy + SUNMatScaleAddI(c) + SUNLinSolSetup
// (cvode_ls.c:1311-1395 issues them one after the other).  Declines with
// SUN_ERR_ARG_INCOMPATIBLE, before touching anything, where the separate
// calls must run instead.
extern "C" int SUNLinSol_Mi355BlockLUSetupScaled(SUNLinearSolver S, SUNMatrix A, SUNMatrix J,
                                                 sunrealtype c, int save_to_J)
{
  if (!S || !S->content || !A || !J || A == J || S->ops->setup != lu_setup ||
      A->ops->getid != bm_getid || J->ops->getid != bm_getid)
  {
    return SUN_ERR_ARG_INCOMPATIBLE;
  }
  BlockLU* lc = LU(S);
  if (!compatible(A, J) || BM(A)->nblocks != lc->nblocks || BM(A)->m != lc->m || lc->mfma ||
      BM(A)->stream != lc->stream || BM(J)->stream != lc->stream)
  {
    return SUN_ERR_ARG_INCOMPATIBLE;
  }
  const int m = (int)lc->m;
  if (m > 10 && m != 12 && m != 16 && m != 24 && m != 32 && m != 48 && m != 64)
  {
    return SUN_ERR_ARG_INCOMPATIBLE;
  }
  const double* src = save_to_J ? BM(A)->dev : BM(J)->dev;
  double* save      = save_to_J ? BM(J)->dev : nullptr;
  k_info_reset<<<1, 1, 0, lc->stream>>>(lc->dinfo);
  sunamd_prof_begin(SUNAMD_K_GETRF, lc->stream);
  if (lc->precision == SUNAMD_LU_FP32)
  {
    launch_getrf_scaled<float>(m, save_to_J != 0, src, c, save, BM(A)->dev, lc->f32, lc->piv,
                               (long)lc->nblocks, lc->dinfo, lc->stream);
  }
  else
  {
    launch_getrf_scaled<double>(m, save_to_J != 0, src, c, save, nullptr, BM(A)->dev, lc->piv,
                                (long)lc->nblocks, lc->dinfo, lc->stream);
  }
  sunamd_prof_end(SUNAMD_K_GETRF, lc->stream);
  return lu_finish(S);
}

// One pass instead of four: cvode_ls.c:1744-1784 (cvLsSolve) zeroes x, solves
// A x = b, copies x into b and scales b by 2/(1 + gamrat); here b is solved in
// place and each entry leaves the kernel as scale * x, the product N_VScale
// forms (scale = 1 leaves x unchanged).  Declines with
// SUN_ERR_ARG_INCOMPATIBLE, before touching anything, where the separate calls
// must run instead.
extern "C" int SUNLinSol_Mi355BlockLUSolveScaled(SUNLinearSolver S, SUNMatrix A, N_Vector b,
                                                 sunrealtype scale)
{
  if (!S || !S->content || !A || !b || S->ops->solve != lu_solve || A->ops->getid != bm_getid)
  {
    return SUN_ERR_ARG_INCOMPATIBLE;
  }
  BlockLU* c = LU(S);
  const int m = (int)c->m;
  if (!SUNAMD_NVectorIsMi355(b) || (hipStream_t)N_VGetStream_Mi355(b) != c->stream ||
      BM(A)->nblocks != c->nblocks || BM(A)->m != c->m || BM(A)->stream != c->stream ||
      (m > 10 && m != 12 && m != 16 && m != 24 && m != 32 && m != 48 && m != 64))
  {
    return SUN_ERR_ARG_INCOMPATIBLE;
  }
  double* bd = N_VGetDeviceArrayPointer(b);
  if (!bd) { return SUN_ERR_ARG_INCOMPATIBLE; }
  bool ok;
  sunamd_prof_begin(SUNAMD_K_GETRS, c->stream);
  if (c->precision == SUNAMD_LU_FP32)
  {
    ok = launch_getrs<float>(m, c->f32, c->piv, bd, bd, (long)c->nblocks, c->stream, scale);
  }
  else
  {
    ok = launch_getrs<double>(m, BM(A)->dev, c->piv, bd, bd, (long)c->nblocks, c->stream,
                              scale);
  }
  sunamd_prof_end(SUNAMD_K_GETRS, c->stream);
  if (!ok) { return SUN_ERR_ARG_INCOMPATIBLE; }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
  {
    hip_fail(S->sunctx, e, "k_getrs", __LINE__);
    c->last_flag = SUN_ERR_EXT_FAIL;
    return SUN_ERR_EXT_FAIL;
  }
  c->last_flag = SUN_SUCCESS;
  return SUN_SUCCESS;
}

// ----------------------------------------------------------------------------
// factorisation benchmark: nblocks blocks of m x m (m = 32 or 64), entries
// U(-1, 1)/m plus the identity, generated on the device; average ms per
// factorisation over reps launches after a warm-up (HIP events).  variant 0:
// FP64 register kernel (into a separate factor array), 1: FP32 register
// kernel, 2: FP32 MFMA blocked kernel.
// ----------------------------------------------------------------------------
__global__ void k_bench_fill(double* __restrict__ a, long nblocks, int m)
{
  const long n = nblocks * m * m;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E5FULL;
    z                    = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z                    = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const double u   = ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const long inb   = i % ((long)m * m);
    const bool diag  = (inb / m) == (inb % m);
    a[i]             = (2.0 * u - 1.0) / m + (diag ? 1.0 : 0.0);
  }
}

extern "C" int SUNAMD_BenchBlockLU(long nblocks, int m, int variant, int reps, double* ms_out)
{
  if (nblocks < 1 || (m != 32 && m != 64) || variant < 0 || variant > 2 || reps < 1 ||
      nblocks > 0x7fffffffL)
  {
    return -1;
  }
  double *A = nullptr, *F64 = nullptr;
  float* F32              = nullptr;
  int* piv                = nullptr;
  unsigned long long* inf = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const size_t nel = (size_t)nblocks * m * m;
  hipError_t e     = hipMalloc((void**)&A, sizeof(double) * nel);
  if (e == hipSuccess && variant == 0) { e = hipMalloc((void**)&F64, sizeof(double) * nel); }
  if (e == hipSuccess && variant > 0) { e = hipMalloc((void**)&F32, sizeof(float) * nel); }
  if (e == hipSuccess) { e = hipMalloc((void**)&piv, sizeof(int) * (size_t)nblocks * m); }
  if (e == hipSuccess) { e = hipMalloc((void**)&inf, sizeof(unsigned long long)); }
  if (e == hipSuccess) { e = hipEventCreate(&e0); }
  if (e == hipSuccess) { e = hipEventCreate(&e1); }
  if (e == hipSuccess)
  {
    k_bench_fill<<<4096, 256>>>(A, nblocks, m);
    auto run = [&]() {
      if (variant == 2)
      {
        if (m == 32) { k_getrf_mfma<32><<<(unsigned)nblocks, 64>>>(A, F32, piv, nblocks, inf); }
        else { k_getrf_mfma<64><<<(unsigned)nblocks, 64>>>(A, F32, piv, nblocks, inf); }
      }
      else if (variant == 1) { (void)launch_getrf<float>(m, A, F32, piv, nblocks, inf, 0); }
      else { (void)launch_getrf<double>(m, A, F64, piv, nblocks, inf, 0); }
    };
    run();
    e = hipDeviceSynchronize();
    if (e == hipSuccess) { e = hipEventRecord(e0, 0); }
    for (int r = 0; r < reps && e == hipSuccess; r++) { run(); }
    if (e == hipSuccess) { e = hipEventRecord(e1, 0); }
    if (e == hipSuccess) { e = hipEventSynchronize(e1); }
    float ms = 0.0f;
    if (e == hipSuccess) { e = hipEventElapsedTime(&ms, e0, e1); }
    if (e == hipSuccess) { e = hipGetLastError(); }
    *ms_out = (double)ms / reps;
  }
  if (e0) { (void)hipEventDestroy(e0); }
  if (e1) { (void)hipEventDestroy(e1); }
  (void)hipFree(A);
  (void)hipFree(F64);
  (void)hipFree(F32);
  (void)hipFree(piv);
  (void)hipFree(inf);
  if (e != hipSuccess)
  {
    fprintf(stderr, "[sundials_amd blockdense] LU benchmark: %s\n", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

extern "C" SUNErrCode SUNLinSol_Mi355BlockLU_GetPivots(SUNLinearSolver S, int* host_out)
{
  BlockLU* c   = LU(S);
  hipError_t e = hipMemcpyAsync(host_out, c->piv, sizeof(int) * (size_t)(c->nblocks * c->m),
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) { e = hipStreamSynchronize(c->stream); }
  if (e != hipSuccess)
  {
    hip_fail(S->sunctx, e, "pivot copy", __LINE__);
    return SUN_ERR_EXT_FAIL;
  }
  return SUN_SUCCESS;
}
