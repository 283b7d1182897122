// -----------------------------------------------------------------------------
// MLP for the learned Newton initial iterate: forward, backward and Adam on
// the MI355X.  The batched products run on MFMA bf16 tiles
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) staged through LDS; bias,
// SiLU, loss and the optimiser are elementwise kernels.  Weights are kept in FP32 (master copy
// and Adam moments) and re-rounded to BF16 after every update.
// -----------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sunamd/learned_predictor.h"

namespace
{

constexpr int IN  = SUNAMD_MLP_IN_PAD;
constexpr int HID = SUNAMD_MLP_HID;
constexpr int OUT = SUNAMD_MLP_OUT;
// parameter offsets (floats)
constexpr long OW1 = 0;
constexpr long OB1 = OW1 + (long)HID * IN;
constexpr long OW2 = OB1 + HID;
constexpr long OB2 = OW2 + (long)HID * HID;
constexpr long OW3 = OB2 + HID;
constexpr long OB3 = OW3 + (long)OUT * HID;
constexpr long NP  = OB3 + OUT;
static_assert(NP == SUNAMD_MLP_NPARAM, "parameter layout");
// device weight block used by the integrator: bf16 W1 W2 W3, then FP32 biases
constexpr long NW_BF16 = (long)HID * IN + (long)HID * HID + (long)OUT * HID;

typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t to_bf16(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) { return (uint16_t)((u >> 16) | 0x40); }
  u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
  return (uint16_t)(u >> 16);
}

__host__ __device__ inline float from_bf16(uint16_t h)
{
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }

__device__ inline float dsilu(float z)
{
  float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// C[i][j] = sum_k A(i,k) B(k,j), A(i,k) = A[i*lai + k*lak], B(k,j) = B[k*lbk + j*lbj],
// bf16 operands, FP32 result; M and N multiples of 16, any K.
//
// A workgroup of four wavefronts owns a 64 x 64 block of C and walks its K
// range (all of K, or one of `splits` consecutive 256-deep ranges) in
// 32-deep slices: the slice of A (64 x 32) and of B (32 x 64, stored
// transposed, k contiguous) is staged in LDS with bounds-checked strided
// loads, then each wavefront runs the 2 x 2 MFMA tiles of its 32 x 32
// quadrant with gfx950's v_mfma_f32_16x16x32_bf16: lane l reads the 8
// consecutive k of its quarter l/16 for row l%16 of its A and B fragments
// (16 bytes each) and holds rows 4(l/16)..4(l/16)+3 of column l%16 of the
// tile.  Each A and B element is loaded from memory once per workgroup
// instead of once per 16 x 16 tile.  With splits > 1 every range writes its
// partial block to part[split][M][N] and k_splitsum adds the ranges in
// order (deterministic).
constexpr int GT  = 64;      // block tile (M and N)
constexpr int GK  = 32;      // K slice
constexpr int GKP = GK + 8;  // LDS row stride (bf16): 16-byte aligned, rows in distinct banks
constexpr int GSPLIT_K = 256; // K range per split

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const uint16_t* __restrict__ A,
                                              long lai, long lak, const uint16_t* __restrict__ B,
                                              long lbk, long lbj, float* __restrict__ C, long ldc,
                                              int splits, float* __restrict__ part)
{
  __shared__ __align__(16) uint16_t As[GT * GKP];
  __shared__ __align__(16) uint16_t Bs[GT * GKP];
  const int tiles_n = (N + GT - 1) / GT;
  const int bi = blockIdx.x / tiles_n, bj = blockIdx.x % tiles_n;
  const int split = blockIdx.y;
  const int k_lo  = (splits > 1) ? split * GSPLIT_K : 0;
  const int k_hi  = (splits > 1) ? min(K, k_lo + GSPLIT_K) : K;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int r = lane & 15, q = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
  {
#pragma unroll
    for (int b = 0; b < 2; b++) { acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f}; }
  }
  for (int k0 = k_lo; k0 < k_hi; k0 += GK)
  {
    // stage: 64 x 32 of A and of B^T, 8 elements per thread each
#pragma unroll
    for (int e = 0; e < 8; e++)
    {
      const int idx = t + 256 * e;      // 0 .. 2047
      const int row = idx >> 5, kk = idx & 31;
      const long gi = (long)bi * GT + row, gj = (long)bj * GT + row;
      const long k  = k0 + kk;
      As[row * GKP + kk] = (gi < M && k < k_hi) ? A[gi * lai + k * lak] : (uint16_t)0;
      Bs[row * GKP + kk] = (gj < N && k < k_hi) ? B[k * lbk + gj * lbj] : (uint16_t)0;
    }
    __syncthreads();
    bf16x8_t af[2], bf[2];
#pragma unroll
    for (int a = 0; a < 2; a++)
    {
      af[a] = *reinterpret_cast<const bf16x8_t*>(&As[(wm + 16 * a + r) * GKP + 8 * q]);
      bf[a] = *reinterpret_cast<const bf16x8_t*>(&Bs[(wn + 16 * a + r) * GKP + 8 * q]);
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
    {
#pragma unroll
      for (int b = 0; b < 2; b++)
      {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bf[b], acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* out   = (splits > 1) ? part + (long)split * M * N : C;
  const long ld = (splits > 1) ? (long)N : ldc;
#pragma unroll
  for (int a = 0; a < 2; a++)
  {
#pragma unroll
    for (int b = 0; b < 2; b++)
    {
      const long col = (long)bj * GT + wn + 16 * b + r;
#pragma unroll
      for (int i = 0; i < 4; i++)
      {
        const long row = (long)bi * GT + wm + 16 * a + 4 * q + i;
        if (row < M && col < N) { out[row * ld + col] = acc[a][b][i]; }
      }
    }
  }
}

// C = sum over the splits of part[split] (in split order)
__global__ void k_splitsum(const float* __restrict__ part, int splits, int M, int N,
                           float* __restrict__ C, long ldc)
{
  const long n = (long)M * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    float s = part[i];
    for (int p = 1; p < splits; p++) { s += part[(long)p * n + i]; }
    C[(i / N) * ldc + i % N] = s;
  }
}

float* g_gemm_part     = nullptr; // split-K partials of the calling network (set per call)
long g_gemm_part_elems = 0;

void gemm(hipStream_t s, int M, int N, int K, const uint16_t* A, long lai, long lak,
          const uint16_t* B, long lbk, long lbj, float* C, long ldc)
{
  const int tiles = ((M + GT - 1) / GT) * ((N + GT - 1) / GT);
  int splits      = (K > 2 * GSPLIT_K && tiles < 64) ? (K + GSPLIT_K - 1) / GSPLIT_K : 1;
  if ((long)splits * M * N > g_gemm_part_elems) { splits = 1; }
  k_gemm<<<dim3((unsigned)tiles, (unsigned)splits), 256, 0, s>>>(M, N, K, A, lai, lak, B, lbk, lbj,
                                                                  C, ldc, splits, g_gemm_part);
  if (splits > 1)
  {
    const long n = (long)M * N;
    k_splitsum<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g_gemm_part, splits, M, N, C, ldc);
  }
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    y[i] = to_bf16(x[i]);
  }
}

// Z[b][j] += bias[j]; A = silu(Z) in bf16 (hidden layers) or Y = Z (output)
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ bias, int B, int N,
                           uint16_t* __restrict__ Abf, int act)
{
  const long n = (long)B * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    float z = Z[i] + bias[i % N];
    Z[i]    = z;
    if (act) { Abf[i] = to_bf16(silu(z)); }
  }
}

// dY = 2 mask (Y - T) / msum (FP32 and bf16), loss accumulated in *acc[0]
__global__ void k_loss(const float* __restrict__ Y, const float* __restrict__ T,
                       const float* __restrict__ Mk, long n, float inv_msum, float* __restrict__ dY,
                       uint16_t* __restrict__ dYbf, float* __restrict__ acc)
{
  float part = 0.0f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    float d = (Y[i] - T[i]) * Mk[i];
    part += d * d;
    float g = 2.0f * d * inv_msum;
    dY[i]   = g;
    dYbf[i] = to_bf16(g);
  }
  for (int off = 32; off >= 1; off >>= 1) { part += __shfl_xor(part, off, 64); }
  if ((threadIdx.x & 63) == 0) { atomicAdd(acc, part * inv_msum); }
}

// dZ = dA * silu'(Z) (FP32 and bf16)
__global__ void k_dsilu(const float* __restrict__ dA, const float* __restrict__ Z, long n,
                        float* __restrict__ dZ, uint16_t* __restrict__ dZbf)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    float g = dA[i] * dsilu(Z[i]);
    dZ[i]   = g;
    dZbf[i] = to_bf16(g);
  }
}

// db[j] = sum_b dZ[b][j]: one workgroup per column, 256 rows at a time,
// fixed-order tree over the threads (deterministic)
__global__ void __launch_bounds__(256) k_colsum(const float* __restrict__ dZ, int B, int N,
                                                float* __restrict__ db)
{
  __shared__ float part[256];
  const int j = blockIdx.x;
  float s     = 0.0f;
  for (int b = threadIdx.x; b < B; b += 256) { s += dZ[(long)b * N + j]; }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1)
  {
    if ((int)threadIdx.x < w) { part[threadIdx.x] += part[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { db[j] = part[0]; }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float c1,
                       float c2, float eps)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i]     = mi;
    v[i]     = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ void k_fill(float* __restrict__ x, long n, float val)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
  {
    x[i] = val;
  }
}

// FP32 master -> device weight block (bf16 weights, FP32 biases)
__global__ void k_pack(const float* __restrict__ p, uint16_t* __restrict__ w, float* __restrict__ bias)
{
  const long nw1 = (long)HID * IN, nw2 = (long)HID * HID, nw3 = (long)OUT * HID;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < NW_BF16 + HID + HID + OUT;
       i += (long)gridDim.x * blockDim.x)
  {
    if (i < nw1) { w[i] = to_bf16(p[OW1 + i]); }
    else if (i < nw1 + nw2) { w[i] = to_bf16(p[OW2 + (i - nw1)]); }
    else if (i < NW_BF16) { w[i] = to_bf16(p[OW3 + (i - nw1 - nw2)]); }
    else
    {
      long j = i - NW_BF16; // b1 b2 b3
      bias[j] = (j < HID) ? p[OB1 + j] : (j < 2 * HID) ? p[OB2 + (j - HID)] : p[OB3 + (j - 2 * HID)];
    }
  }
}

unsigned blocks_for(long n) { return (unsigned)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048); }

} // namespace

struct SUNAMD_MLP_
{
  hipStream_t stream;
  int max_batch;
  float* p;    // FP32 master parameters
  float* g;    // gradient
  float *m, *v; // Adam moments
  long t;      // Adam step count
  void* wblock; // bf16 weights + FP32 biases
  uint16_t* w1;
  uint16_t* w2;
  uint16_t* w3;
  float* bias; // b1 b2 b3
  // batch work buffers
  uint16_t *xb, *a1b, *a2b, *dyb, *dz2b, *dz1b;
  float *z1, *z2, *y, *dy, *da2, *dz2, *da1, *dz1, *acc;
  float *hx, *ht, *hm; // staging for the host-array entry points (device)
  float* gpart;        // split-K partials of the weight-gradient products
  long gpart_elems;
};

static int mfail(hipError_t e, const char* what)
{
  fprintf(stderr, "[sundials_amd mlp] %s: %s\n", what, hipGetErrorString(e));
  return -1;
}

#define MCHK(call, what)                                 \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) { return mfail(e_, what); }    \
  } while (0)

static int repack(SUNAMD_MLP* M)
{
  k_pack<<<blocks_for(NW_BF16 + 320), 256, 0, M->stream>>>(M->p, M->w1, M->bias);
  MCHK(hipGetLastError(), "pack");
  return 0;
}

extern "C" SUNAMD_MLP* SUNAMD_MLPCreate(uint64_t seed, int max_batch)
{
  if (max_batch < 16 || max_batch % 16) { return NULL; }
  SUNAMD_MLP* M = (SUNAMD_MLP*)calloc(1, sizeof(SUNAMD_MLP));
  if (!M) { return NULL; }
  M->max_batch = max_batch;
  const long B = max_batch;
  bool ok      = hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) == hipSuccess;
  auto alloc   = [&](void** ptr, size_t bytes) {
    if (ok && hipMalloc(ptr, bytes) != hipSuccess) { ok = false; }
  };
  alloc((void**)&M->p, sizeof(float) * NP);
  alloc((void**)&M->g, sizeof(float) * NP);
  alloc((void**)&M->m, sizeof(float) * NP);
  alloc((void**)&M->v, sizeof(float) * NP);
  alloc(&M->wblock, sizeof(uint16_t) * NW_BF16 + sizeof(float) * (2 * HID + OUT));
  alloc((void**)&M->xb, sizeof(uint16_t) * B * IN);
  alloc((void**)&M->a1b, sizeof(uint16_t) * B * HID);
  alloc((void**)&M->a2b, sizeof(uint16_t) * B * HID);
  alloc((void**)&M->dyb, sizeof(uint16_t) * B * OUT);
  alloc((void**)&M->dz2b, sizeof(uint16_t) * B * HID);
  alloc((void**)&M->dz1b, sizeof(uint16_t) * B * HID);
  alloc((void**)&M->z1, sizeof(float) * B * HID);
  alloc((void**)&M->z2, sizeof(float) * B * HID);
  alloc((void**)&M->y, sizeof(float) * B * OUT);
  alloc((void**)&M->dy, sizeof(float) * B * OUT);
  alloc((void**)&M->da2, sizeof(float) * B * HID);
  alloc((void**)&M->dz2, sizeof(float) * B * HID);
  alloc((void**)&M->da1, sizeof(float) * B * HID);
  alloc((void**)&M->dz1, sizeof(float) * B * HID);
  alloc((void**)&M->acc, sizeof(float) * 4);
  alloc((void**)&M->hx, sizeof(float) * B * IN);
  alloc((void**)&M->ht, sizeof(float) * B * OUT);
  alloc((void**)&M->hm, sizeof(float) * B * OUT);
  // the longest K is the batch: at most ceil(B / 256) ranges of the largest
  // weight block (HID x IN)
  M->gpart_elems = ((B + 255) / 256) * (long)HID * IN;
  alloc((void**)&M->gpart, sizeof(float) * M->gpart_elems);
  if (!ok)
  {
    SUNAMD_MLPFree(M);
    return NULL;
  }
  M->w1   = (uint16_t*)M->wblock;
  M->w2   = M->w1 + (long)HID * IN;
  M->w3   = M->w2 + (long)HID * HID;
  M->bias = (float*)(M->w1 + NW_BF16);
  // seeded init on the host (SplitMix64 -> uniform)
  std::vector<float> h(NP, 0.0f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E5Full;
  auto uni    = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto init = [&](long off, int nout, int nin, int nin_used, double scale) {
    double a = scale * std::sqrt(6.0 / nin_used);
    for (int o = 0; o < nout; o++)
    {
      for (int i = 0; i < nin; i++)
      {
        h[off + (long)o * nin + i] = (i < nin_used) ? (float)((2.0 * uni() - 1.0) * a) : 0.0f;
      }
    }
  };
  init(OW1, HID, IN, SUNAMD_MLP_IN, 1.0);
  init(OW2, HID, HID, HID, 1.0);
  init(OW3, OUT, HID, HID, 0.1);
  if (SUNAMD_MLPSetParams(M, h.data()) != 0)
  {
    SUNAMD_MLPFree(M);
    return NULL;
  }
  return M;
}

extern "C" void SUNAMD_MLPFree(SUNAMD_MLP* M)
{
  if (!M) { return; }
  if (M->stream) { (void)hipStreamSynchronize(M->stream); }
  void* ptrs[] = {M->p,   M->g,   M->m,   M->v,   M->wblock, M->xb,  M->a1b, M->a2b,
                  M->dyb, M->dz2b, M->dz1b, M->z1, M->z2,  M->y,   M->dy,  M->da2,
                  M->dz2, M->da1, M->dz1, M->acc, M->hx,     M->ht,  M->hm, M->gpart};
  for (void* q : ptrs)
  {
    if (q) { (void)hipFree(q); }
  }
  if (M->stream) { (void)hipStreamDestroy(M->stream); }
  free(M);
}

extern "C" int SUNAMD_MLPGetParams(SUNAMD_MLP* M, float* params)
{
  MCHK(hipStreamSynchronize(M->stream), "sync");
  MCHK(hipMemcpy(params, M->p, sizeof(float) * NP, hipMemcpyDeviceToHost), "get params");
  return 0;
}

extern "C" int SUNAMD_MLPSetParams(SUNAMD_MLP* M, const float* params)
{
  MCHK(hipStreamSynchronize(M->stream), "sync");
  MCHK(hipMemcpy(M->p, params, sizeof(float) * NP, hipMemcpyHostToDevice), "set params");
  k_fill<<<blocks_for(NP), 256, 0, M->stream>>>(M->m, NP, 0.0f);
  k_fill<<<blocks_for(NP), 256, 0, M->stream>>>(M->v, NP, 0.0f);
  M->t = 0;
  if (repack(M) != 0) { return -1; }
  MCHK(hipStreamSynchronize(M->stream), "set params sync");
  return 0;
}

extern "C" const void* SUNAMD_MLPDeviceWeights(SUNAMD_MLP* M) { return M->wblock; }

// forward on device rows x[B x IN] (FP32): fills z1/a1b/z2/a2b/y
static void forward_dev(SUNAMD_MLP* M, const float* x, int B)
{
  hipStream_t s = M->stream;
  g_gemm_part       = M->gpart;
  g_gemm_part_elems = M->gpart_elems;
  k_to_bf16<<<blocks_for((long)B * IN), 256, 0, s>>>(x, M->xb, (long)B * IN);
  gemm(s, B, HID, IN, M->xb, IN, 1, M->w1, 1, IN, M->z1, HID);
  k_bias_act<<<blocks_for((long)B * HID), 256, 0, s>>>(M->z1, M->bias, B, HID, M->a1b, 1);
  gemm(s, B, HID, HID, M->a1b, HID, 1, M->w2, 1, HID, M->z2, HID);
  k_bias_act<<<blocks_for((long)B * HID), 256, 0, s>>>(M->z2, M->bias + HID, B, HID, M->a2b, 1);
  gemm(s, B, OUT, HID, M->a2b, HID, 1, M->w3, 1, HID, M->y, OUT);
  k_bias_act<<<blocks_for((long)B * OUT), 256, 0, s>>>(M->y, M->bias + 2 * HID, B, OUT, nullptr, 0);
}

// loss + gradient into M->g for device rows (x, t, mask), mask sum msum
static void backward_dev(SUNAMD_MLP* M, const float* x, const float* t, const float* mk, int B,
                         float msum)
{
  hipStream_t s = M->stream;
  forward_dev(M, x, B);
  k_fill<<<1, 64, 0, s>>>(M->acc, 4, 0.0f);
  const float inv = msum > 0.0f ? 1.0f / msum : 0.0f;
  k_loss<<<blocks_for((long)B * OUT), 256, 0, s>>>(M->y, t, mk, (long)B * OUT, inv, M->dy, M->dyb,
                                                   M->acc);
  // layer 3: dW3 = dY^T A2, db3 = colsum dY, dA2 = dY W3
  gemm(s, OUT, HID, B, M->dyb, 1, OUT, M->a2b, HID, 1, M->g + OW3, HID);
  k_colsum<<<OUT, 256, 0, s>>>(M->dy, B, OUT, M->g + OB3);
  gemm(s, B, HID, OUT, M->dyb, OUT, 1, M->w3, HID, 1, M->da2, HID);
  k_dsilu<<<blocks_for((long)B * HID), 256, 0, s>>>(M->da2, M->z2, (long)B * HID, M->dz2, M->dz2b);
  // layer 2
  gemm(s, HID, HID, B, M->dz2b, 1, HID, M->a1b, HID, 1, M->g + OW2, HID);
  k_colsum<<<HID, 256, 0, s>>>(M->dz2, B, HID, M->g + OB2);
  gemm(s, B, HID, HID, M->dz2b, HID, 1, M->w2, HID, 1, M->da1, HID);
  k_dsilu<<<blocks_for((long)B * HID), 256, 0, s>>>(M->da1, M->z1, (long)B * HID, M->dz1, M->dz1b);
  // layer 1
  gemm(s, HID, IN, B, M->dz1b, 1, HID, M->xb, IN, 1, M->g + OW1, IN);
  k_colsum<<<HID, 256, 0, s>>>(M->dz1, B, HID, M->g + OB1);
}

extern "C" int SUNAMD_MLPForward(SUNAMD_MLP* M, const float* x, int B, float* y)
{
  if (B <= 0 || B > M->max_batch || B % 16) { return -1; }
  MCHK(hipMemcpyAsync(M->hx, x, sizeof(float) * B * IN, hipMemcpyHostToDevice, M->stream), "x");
  forward_dev(M, M->hx, B);
  MCHK(hipGetLastError(), "forward");
  MCHK(hipMemcpyAsync(y, M->y, sizeof(float) * B * OUT, hipMemcpyDeviceToHost, M->stream), "y");
  MCHK(hipStreamSynchronize(M->stream), "forward sync");
  return 0;
}

extern "C" int SUNAMD_MLPLossGrad(SUNAMD_MLP* M, const float* x, const float* t, const float* mask,
                                  int B, float* loss, float* grad)
{
  if (B <= 0 || B > M->max_batch || B % 16) { return -1; }
  float msum = 0.0f;
  for (long i = 0; i < (long)B * OUT; i++) { msum += mask[i]; }
  MCHK(hipMemcpyAsync(M->hx, x, sizeof(float) * B * IN, hipMemcpyHostToDevice, M->stream), "x");
  MCHK(hipMemcpyAsync(M->ht, t, sizeof(float) * B * OUT, hipMemcpyHostToDevice, M->stream), "t");
  MCHK(hipMemcpyAsync(M->hm, mask, sizeof(float) * B * OUT, hipMemcpyHostToDevice, M->stream),
       "mask");
  backward_dev(M, M->hx, M->ht, M->hm, B, msum);
  MCHK(hipGetLastError(), "backward");
  float acc[4];
  MCHK(hipMemcpyAsync(acc, M->acc, sizeof(acc), hipMemcpyDeviceToHost, M->stream), "loss");
  MCHK(hipMemcpyAsync(grad, M->g, sizeof(float) * NP, hipMemcpyDeviceToHost, M->stream), "grad");
  MCHK(hipStreamSynchronize(M->stream), "grad sync");
  *loss = acc[0];
  return 0;
}

// sum of a mask slice (device), one workgroup
__global__ void k_masksum(const float* __restrict__ mk, long n, float* __restrict__ out)
{
  __shared__ float part[256];
  float s = 0.0f;
  for (long i = threadIdx.x; i < n; i += 256) { s += mk[i]; }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1)
  {
    if ((int)threadIdx.x < w) { part[threadIdx.x] += part[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[1] = part[0]; }
}

extern "C" int SUNAMD_MLPTrain(SUNAMD_MLP* M, const float* x_dev, const float* t_dev,
                               const float* mask_dev, long n, int batch, int steps, float lr,
                               float* loss)
{
  if (batch <= 0 || batch > M->max_batch || batch % 16 || n < batch) { return -1; }
  hipStream_t s = M->stream;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double tail = 0.0;
  int ntail   = 0;
  for (int k = 0; k < steps; k++)
  {
    long start = ((long)k * batch) % n;
    if (start + batch > n) { start = n - batch; }
    const float* x  = x_dev + start * IN;
    const float* t  = t_dev + start * OUT;
    const float* mk = mask_dev + start * OUT;
    k_masksum<<<1, 256, 0, s>>>(mk, (long)batch * OUT, M->acc);
    float hacc[4];
    MCHK(hipMemcpyAsync(hacc, M->acc, sizeof(hacc), hipMemcpyDeviceToHost, s), "mask sum");
    MCHK(hipStreamSynchronize(s), "mask sum sync");
    backward_dev(M, x, t, mk, batch, hacc[1]);
    M->t++;
    const float c1 = 1.0f - powf(b1, (float)M->t), c2 = 1.0f - powf(b2, (float)M->t);
    k_adam<<<blocks_for(NP), 256, 0, s>>>(M->p, M->g, M->m, M->v, NP, lr, b1, b2, c1, c2, eps);
    if (repack(M) != 0) { return -1; }
    MCHK(hipGetLastError(), "train step");
    if (k >= steps - 10)
    {
      MCHK(hipMemcpyAsync(hacc, M->acc, sizeof(hacc), hipMemcpyDeviceToHost, s), "loss");
      MCHK(hipStreamSynchronize(s), "loss sync");
      tail += hacc[0];
      ntail++;
    }
  }
  MCHK(hipStreamSynchronize(s), "train sync");
  if (loss) { *loss = ntail ? (float)(tail / ntail) : 0.0f; }
  return 0;
}
