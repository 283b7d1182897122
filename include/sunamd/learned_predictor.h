/* -----------------------------------------------------------------------------
 * learned_predictor.h -- a small MLP that proposes the initial Newton
 * correction of a BDF step, trained on the MI355X by this library's own
 * forward / backward kernels.  No counterpart in the reference: the
 * reference starts every Newton solve from a zero correction (CVODE
 * N_VConst(ZERO, acor) before SUNNonlinSolSolve, cvode.c:3095-3109); this is
 * the alternative initial iterate the batched integrator can use instead,
 * with the zero correction kept as the fallback.
 *
 * Network: 129 inputs (padded to 144 with zeros) -> 128 -> 128 -> 64, SiLU
 * on the hidden layers, linear output.  Weights are held in FP32 (master
 * copy, Adam state) and used in BF16; every product accumulates in FP32
 * (MFMA bf16 tiles for the batched training GEMMs).
 *
 * Inputs of one system (m species, m <= 64), all scaled by 1/10:
 *   x[0..63]    log10 max(|y_pred_i|, 1e-30) for i < m, 0 beyond
 *   x[64]       log10 gamma (gamma = h / l_1, the Newton-matrix scale)
 *   x[65..128]  log10(1 + sum of k_f + k_b over the reactions touching
 *               species i) for i < m, 0 beyond
 * Output y[i] (i < m): the converged correction acor_i in units of the
 * error weight, i.e. acor_i = y[i] / ewt_i.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_LEARNED_PREDICTOR_H
#define SUNAMD_LEARNED_PREDICTOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SUNAMD_MLP_IN     129
#define SUNAMD_MLP_IN_PAD 144
#define SUNAMD_MLP_HID    128
#define SUNAMD_MLP_OUT    64

/* number of FP32 parameters: W1[128x144] b1[128] W2[128x128] b2[128]
   W3[64x128] b3[64], in that order, row-major (W[out][in]) */
#define SUNAMD_MLP_NPARAM                                                          \
  (SUNAMD_MLP_HID * SUNAMD_MLP_IN_PAD + SUNAMD_MLP_HID + SUNAMD_MLP_HID * SUNAMD_MLP_HID + \
   SUNAMD_MLP_HID + SUNAMD_MLP_OUT * SUNAMD_MLP_HID + SUNAMD_MLP_OUT)

/* bf16 entries of the device weight block (W1, W2, W3); the FP32 biases
   b1, b2, b3 follow them */
#define SUNAMD_MLP_WEIGHTS_BF16                                                   \
  (SUNAMD_MLP_HID * SUNAMD_MLP_IN_PAD + SUNAMD_MLP_HID * SUNAMD_MLP_HID +             \
   SUNAMD_MLP_OUT * SUNAMD_MLP_HID)

typedef struct SUNAMD_MLP_ SUNAMD_MLP;

/* weights drawn from a seeded uniform(+-sqrt(6/fan_in)) (x 0.1 on the
   output layer), biases zero; max_batch bounds the training batch size */
SUNAMD_MLP* SUNAMD_MLPCreate(uint64_t seed, int max_batch);
void SUNAMD_MLPFree(SUNAMD_MLP* M);

/* FP32 parameters in the layout above (host arrays of SUNAMD_MLP_NPARAM);
   SetParams also refreshes the BF16 copies and clears the Adam state */
int SUNAMD_MLPGetParams(SUNAMD_MLP* M, float* params);
int SUNAMD_MLPSetParams(SUNAMD_MLP* M, const float* params);

/* device pointer of the BF16 weights + FP32 biases used by the integrator
   kernels (layout: W1 W2 W3 as bf16, then b1 b2 b3 as float) */
const void* SUNAMD_MLPDeviceWeights(SUNAMD_MLP* M);

/* forward pass on a batch (host arrays): x[B x 144] -> y[B x 64] */
int SUNAMD_MLPForward(SUNAMD_MLP* M, const float* x, int B, float* y);

/* loss and gradient on a batch (host arrays): masked mean squared error
   sum_{b,i} mask[b,i] (y[b,i] - t[b,i])^2 / sum mask; grad has
   SUNAMD_MLP_NPARAM entries (the W1 columns beyond 129 get zero gradient) */
int SUNAMD_MLPLossGrad(SUNAMD_MLP* M, const float* x, const float* t, const float* mask, int B,
                       float* loss, float* grad);

/* the batched integrator's own in-kernel evaluation of the network (one
   wavefront per 16 rows, bf16 weights as 16 x 16 x 32 MFMA operands, FP32
   accumulation) on host arrays x[rows x 144] -> y[rows x 64]; weights =
   SUNAMD_MLPDeviceWeights.  For checking it against SUNAMD_MLPForward and a
   CPU reference. */
int SUNAMD_MLPForwardWave(const void* weights, const float* x, int rows, float* y);
/* training on device-resident data: x[n x 144], t[n x 64], mask[n x 64]
   (device pointers); runs `steps` Adam steps on consecutive batches of
   `batch` rows (wrapping around), learning rate lr; *loss receives the mean
   loss of the last 10 steps */
int SUNAMD_MLPTrain(SUNAMD_MLP* M, const float* x_dev, const float* t_dev,
                    const float* mask_dev, long n, int batch, int steps, float lr, float* loss);

/* ---- training-pair log file (versioned binary, no reference counterpart) --
   Header: magic "SUNAMDTP", version 1, header size 64, rows, 144 inputs and
   64 targets per row, and the logging runs' seed, rtol, atol and stride; then
   per row the system's species count m (uint32), the 144 float32 inputs and
   the 64 float32 targets (little-endian).  The species mask of a row is its
   first m entries.  Return codes: 0 success, -1 bad argument, -2 cannot open,
   -3 I/O error or short file, -4 a mask that is not a prefix of ones (write),
   -5 not a log file, -6 other version or row shape, -7 rows out of range,
   -8 corrupt row. */
typedef struct
{
  int version;
  long rows;
  uint64_t seed;
  double rtol, atol;
  int stride;
} SUNAMD_TrainingLogInfo;

/* host arrays x[rows x 144], t[rows x 64], mask[rows x 64] (the layout the
   batched integrator logs into); info supplies seed / rtol / atol / stride
   (NULL: zeros) */
int SUNAMD_TrainingLogWrite(const char* path, const float* x, const float* t,
                            const float* mask, long rows, const SUNAMD_TrainingLogInfo* info);
int SUNAMD_TrainingLogInfoRead(const char* path, SUNAMD_TrainingLogInfo* info);
/* rows [row0, row0 + nrows) into host arrays of the write layout */
int SUNAMD_TrainingLogRead(const char* path, long row0, long nrows, float* x, float* t,
                           float* mask);

#ifdef __cplusplus
}
#endif

#endif
