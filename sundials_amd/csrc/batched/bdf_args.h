// -----------------------------------------------------------------------------
// Interface between the batched BDF host code and its per-size kernels.
// -----------------------------------------------------------------------------
#ifndef SUNAMD_BDF_ARGS_H
#define SUNAMD_BDF_ARGS_H

#include <hip/hip_runtime.h>

namespace sunamd_bdf
{

__host__ __device__ constexpr int pow2ceil(int m)
{
  return m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : m <= 8 ? 8 : m <= 16 ? 16 : m <= 32 ? 32 : 64;
}

// one size bucket as the kernel sees it
struct BucketArgs
{
  int m, R;
  long ncells;
  const int* idx;     // 4R
  const int* soff;    // m+1: species s's terms are sent[soff[s] .. soff[s+1])
  const int* sent;    // 4R terms, 4*reaction + role (0,1 reactant; 2,3 product)
  const double* kf;   // ncells*R, cell-major
  const double* kb;
  const double* y0;   // ncells*m
  double* yout;       // ncells*m
  long* stats;        // ncells*SUNAMD_BS_COUNT
  unsigned long long* next;
  double rtol, atol, tout;
  long mxstep;
  // learned initial Newton iterate (mlp_w == nullptr: off)
  const unsigned short* mlp_w; // bf16 W1[128x144] W2[128x128] W3[64x128]
  const float* mlp_b;          // b1[128] b2[128] b3[64]
  // training-pair log (log_x == nullptr: off): rows of 144 inputs, 64
  // targets and 64 mask entries, one row every log_stride-th step
  float* log_x;
  float* log_t;
  float* log_m;
  unsigned long long* log_next;
  long log_cap;
  int log_stride;
  // learned-guess counters of the bucket: [steps tried, guesses kept]
  unsigned long long* lstats;
  // non-null: a flag the reaction table's validation set on the device; when
  // it is nonzero every system is returned unintegrated with flag -22
  const int* bad;
  // instrumented build only: clock ticks per phase, summed over the groups
  unsigned long long* phase;
};

// LDS floats per lane group staging one training-log row (the inputs)
constexpr int MLP_LOG_FLOATS = 144;
// the wavefront's network area: 16 rows of bf16 inputs (stride MLP_XS: 144
// used, zero to 160 for the 32-deep MFMA slices) and two hidden layers
// (stride MLP_HS); the strides are padded so that the 16 rows' 16-byte
// operand reads fall in different LDS banks
constexpr int MLP_XS         = 168;
constexpr int MLP_HS         = 136;
constexpr int MLP_WAVE_BYTES = 2 * (16 * MLP_XS + 2 * 16 * MLP_HS);

// doubles of a lane group's LDS rows area: the m Jacobian rows (stride m+1,
// entries of row_bytes: 8 in the FP64 form, 4 in the FP32 form) while the
// Newton matrix is built, the R reaction rates (doubles) during an RHS
__host__ __device__ constexpr long group_rows_stride(int m, int R, int row_bytes)
{
  return ((long)m * (m + 1) * row_bytes + 7) / 8 > (long)R ? ((long)m * (m + 1) * row_bytes + 7) / 8
                                                          : (long)R;
}

// dynamic LDS of one workgroup (GPW lane groups): staged rate constants,
// network tables, the groups' rows areas, the wavefront's network area with
// the learned start and the log-row staging with the training log
inline size_t bucket_lds_bytes(int m, int R, int gpw, bool learned, bool log, int row_bytes)
{
  size_t b = sizeof(double) * (size_t)gpw * 2 * R;
  b += sizeof(int) * (size_t)(8 * R + ((m + 2) & ~1));
  b += sizeof(double) * (size_t)gpw * group_rows_stride(m, R, row_bytes);
  if (learned) { b += MLP_WAVE_BYTES; }
  if (log) { b += sizeof(float) * (size_t)gpw * MLP_LOG_FLOATS; }
  return b;
}

// launches the persistent kernel for systems of M species (T: LU precision)
// compiled for at least `waves` wavefronts per SIMD (1, 2 or 4); defined in
// bdf_device.h, instantiated in the bdf_inst_*.hip units
template <int M, typename T>
void launch(const BucketArgs& A, int grid, hipStream_t stream, int waves);

} // namespace sunamd_bdf

#endif
