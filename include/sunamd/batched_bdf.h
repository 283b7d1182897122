/* -----------------------------------------------------------------------------
 * batched_bdf.h -- thousands of independent stiff systems of uneven size,
 * each integrated by its own BDF / modified-Newton / dense-LU integrator with
 * its own step size and order, all inside one persistent launch per size
 * bucket on the MI355X.
 *
 * Each system follows the reference CVODE algorithm exactly as the host
 * restatement (cvode_bdf.c) does for a serial N_Vector of length m with a
 * dense m x m LU (cvode.c cvStep / cvNls / error test / eta selection,
 * cvode_nls.c, cvode_ls.c, sunnonlinsol_newton.c, sundials_dense.c):
 *   - one group of pow2ceil(m) lanes per system, lane = species;
 *   - WRMS norms summed in species order (the serial vector's order);
 *   - the Newton matrix I - gamma*J factored by the register-resident
 *     pivoting LU of the block solver, bit-identical in FP64;
 *   - the saved Jacobian is kept as the state it was evaluated at and
 *     rebuilt when reused (the same values as a stored copy);
 *   - x^y in the step-size ratios is the correctly rounded sunamd_pow used by
 *     the host integrator.
 * In FP64 mode every system's step count, iteration counts and final state
 * equal the host integrator's on the same system, bit for bit.  FP32 mode is
 * the fast product form: the Newton matrix is factored and solved in single
 * precision, WRMS norms are summed as a butterfly over the group's lanes and
 * x^y in the step ratios uses the hardware log2 / exp2 (relative error
 * ~1e-7); residuals, Newton updates and the norms themselves stay FP64.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_BATCHED_BDF_H
#define SUNAMD_BATCHED_BDF_H

#include "massaction.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SUNAMD_BATCH_MAXB 8

/* systems grouped in size buckets; bucket b holds nb[b] systems of m[b]
   species, stored as the cells of its own mass-action network */
typedef struct
{
  int nbuckets;
  SUNAMD_MassAction* bucket[SUNAMD_BATCH_MAXB];
  int64_t nsys;
  int* sys_bucket;     /* bucket of system s */
  int64_t* sys_cell;   /* cell index of system s inside its bucket */
  int64_t* sys_yoff;   /* offset of system s in the concatenated state */
  int64_t ytotal;      /* sum of all m */
  uint64_t seed;
} SUNAMD_Batch;

/* nsys systems; m drawn from m_values with the given weights (counter-based,
   stream 5); R = round(rfactor * m); bucket b's network uses seed
   seed + 7919*(b+1); systems of a bucket are its cells 0, 1, ... in order */
SUNAMD_Batch* SUNAMD_BatchCreate(int64_t nsys, uint64_t seed, int nvalues, const int* m_values,
                                 const double* weights, double rfactor, double sigma);
void SUNAMD_BatchFree(SUNAMD_Batch* b);

/* per-system counters, 10 per system, in this order */
#define SUNAMD_BS_NST     0
#define SUNAMD_BS_NFE     1
#define SUNAMD_BS_NSETUPS 2
#define SUNAMD_BS_NETF    3
#define SUNAMD_BS_NNI     4
#define SUNAMD_BS_NNF     5
#define SUNAMD_BS_NCFN    6
#define SUNAMD_BS_NJE     7
#define SUNAMD_BS_QLAST   8
#define SUNAMD_BS_FLAG    9
#define SUNAMD_BS_COUNT   10

typedef struct SUNAMD_BatchedBDF_ SUNAMD_BatchedBDF;

/* upload the batch and allocate results (lu_precision: SUNAMD_LU_FP64/FP32) */
SUNAMD_BatchedBDF* SUNAMD_BatchedBDFCreate(const SUNAMD_Batch* batch, double rtol,
                                           double atol, int lu_precision);
/* largest number of steps per system (CVodeSetMaxNumSteps, cvode_io.c):
   0 restores the default 500, a negative value disables the limit; a system
   that reaches it stops with flag -1 (CV_TOO_MUCH_WORK) */
int SUNAMD_BatchedBDFSetMaxNumSteps(SUNAMD_BatchedBDF* B, long int mxsteps);
/* integrate every system from t = 0 to tout; returns 0 when every system
   succeeded; *wall_s receives the device time of the whole batch */
int SUNAMD_BatchedBDFRun(SUNAMD_BatchedBDF* B, double tout, double* wall_s);
/* results in system order: y (ytotal), counters (nsys*SUNAMD_BS_COUNT) */
int SUNAMD_BatchedBDFGetResults(SUNAMD_BatchedBDF* B, double* y, long int* counters);
/* register budget of each bucket's kernel: waves[k] in {1, 2, 3, 4} is the
   minimum number of wavefronts per SIMD its compilation allows (1: no limit,
   default; 2 / 3 / 4: at most 256 / 168 / 128 registers, the rest spilled;
   3 applies to buckets without the learned start, which otherwise use 2) --
   the occupancy-versus-spilling choice an autotuner makes per size bucket */
int SUNAMD_BatchedBDFSetOccupancy(SUNAMD_BatchedBDF* B, const int* waves);
/* launch size of each bucket: at most per_cu[k] workgroups (one wavefront
   each) per CU, 1..64 (default 16).  Fewer wavefronts than the bucket's
   systems make each lane group take several systems from the bucket's work
   queue, which evens out systems of different step counts inside a
   wavefront; more give the GPU more wavefronts to schedule -- the second
   launch-shape choice an autotuner makes per size bucket */
int SUNAMD_BatchedBDFSetGridCap(SUNAMD_BatchedBDF* B, const int* per_cu);
/* Learned initial Newton iterate (no counterpart in the reference, whose
   Newton solve always starts from a zero correction, cvode.c:3095-3109):
   weights = SUNAMD_MLPDeviceWeights of a trained network (NULL: off),
   bit k of bucket_mask enables it for bucket k.  At the first attempt of
   every step after the first, the network proposes a correction (evaluated
   once per wavefront for all of its lane groups, bf16 weights as
   16 x 16 x 32 MFMA operands, FP32 accumulation); the network's correction
   and the zero correction are both evaluated and the one with the smaller
   weighted residual norm starts the Newton iteration (two RHS
   evaluations). */
int SUNAMD_BatchedBDFSetLearnedPredictor(SUNAMD_BatchedBDF* B, const void* weights,
                                         int bucket_mask);
/* Training-pair log: device arrays x[capacity x 144], t[capacity x 64],
   mask[capacity x 64] receive, for every stride-th step of every system, the
   network inputs at the predictor, the converged correction in error-weight
   units and the species mask (learned_predictor.h); the row counter restarts
   here and keeps counting across runs */
int SUNAMD_BatchedBDFSetTrainingLog(SUNAMD_BatchedBDF* B, float* x, float* t, float* mask,
                                    long capacity, int stride);
long int SUNAMD_BatchedBDFGetTrainingLogCount(SUNAMD_BatchedBDF* B);
/* steps of the last run where the learned guess was tried / kept */
int SUNAMD_BatchedBDFGetLearnedStats(SUNAMD_BatchedBDF* B, long int* tried, long int* kept);
/* the same per bucket (arrays of the bucket count, bucket order); returns
   the number of buckets */
int SUNAMD_BatchedBDFGetLearnedStatsByBucket(SUNAMD_BatchedBDF* B, long int* tried,
                                             long int* kept);
/* kernel time of every bucket in the last run, in bucket order (HIP events
   on the bucket's stream around its launch; the buckets run concurrently);
   returns the number of buckets */
int SUNAMD_BatchedBDFGetKernelTimes(SUNAMD_BatchedBDF* B, double* ms);
void SUNAMD_BatchedBDFFree(SUNAMD_BatchedBDF* B);

/* Phase clock (profiling aid, no counterpart in the reference).  A library
   built with SUNAMD_BDF_PHASES (libsundials_amd_phase.so, the batched
   integrator only, two wavefronts per SIMD) reads the shader clock at every
   phase change of every lane group.  For the last run, summed over bucket
   k's groups, out[k * SUNAMD_BDF_PHASE_VALUES + ...] receives: [p] ticks
   spent in phase p, [SUNAMD_BDF_NPHASES + p] entries into phase p, then the
   ticks and the number of a calibration run of back-to-back phase changes
   (each group's, at its start) -- the cost of one change, which lands in the
   phase entered.  Returns the number of buckets, or -1 from a library built
   without the clock. */
#define SUNAMD_BDF_NPHASES  16
#define SUNAMD_BDF_PHASE_VALUES (2 * SUNAMD_BDF_NPHASES + 2)
#define SUNAMD_PH_ACQUIRE   0  /* taking a system: state set-up, f(t0, y0), initial step */
#define SUNAMD_PH_GUARD     1  /* error weights, step-limit and accuracy guards */
#define SUNAMD_PH_PREDICT   2  /* order / step change, predictor, BDF coefficients */
#define SUNAMD_PH_NEWTON    3  /* the Newton iteration's own arithmetic and control */
#define SUNAMD_PH_RHS       4  /* right-hand side evaluations */
#define SUNAMD_PH_JAC       5  /* Jacobian and Newton matrix I - gamma J */
#define SUNAMD_PH_GETRF     6  /* LU factorisation */
#define SUNAMD_PH_GETRS     7  /* triangular solves */
#define SUNAMD_PH_NORM      8  /* WRMS norms */
#define SUNAMD_PH_ERRTEST   9  /* failure handling and the local error test */
#define SUNAMD_PH_COMPLETE  10 /* history update after an accepted step */
#define SUNAMD_PH_PREPARE   11 /* next step size and order */
#define SUNAMD_PH_POW       12 /* x^y in the step-ratio formulas */
#define SUNAMD_PH_LEARN     13 /* learned start: features and network */
#define SUNAMD_PH_OUTPUT    14 /* tout test, interpolation, results */
#define SUNAMD_PH_IDLE      15 /* group without a system while its wavefront runs */
int SUNAMD_BatchedBDFGetPhaseTicks(SUNAMD_BatchedBDF* B, unsigned long long* out);

/* Allocation-free integration of one size bucket.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  It runs the same kernel as SUNAMD_BatchedBDFRun (two
   wavefronts per SIMD) on ncells systems of m species with the R reactions
   of idx, all caller-owned device buffers laid out as in SUNAMD_MassAction:
   idx[4R], kf[ncells*R], kb[ncells*R] (cell-major), y0[ncells*m] ->
   yout[ncells*m], counters[ncells*SUNAMD_BS_COUNT].  The work queue, the
   validated reaction table and the per-species term lists live in a fixed
   device array of the library (about 33 KB per device), so R is limited to
   SUNAMD_NOALLOC_MAX_REACTIONS; m must be one of 1-8, 10, 12, 16, 24, 32,
   48, 64.  mxsteps as in SUNAMD_BatchedBDFSetMaxNumSteps.  The call is
   asynchronous on `stream` (a hipStream_t of the current device; NULL is
   the default stream) and calls no allocator; the results are complete
   when the stream is.  An index of idx outside [0, m) is caught on the
   device: every system then returns y0 with flag -22 (CV_ILL_INPUT).
   Calls that overlap on one device share the array and race: serialise
   them (one stream, or events).  Returns 0 when queued, -1 for unsupported
   sizes or NULL buffers (nothing queued), -2 on a HIP error. */
#define SUNAMD_NOALLOC_MAX_REACTIONS 1024
int SUNAMD_BatchedBDFSolveBucketNoAlloc(int m, int R, int64_t ncells, const int* idx,
                                        const double* kf, const double* kb, const double* y0,
                                        double rtol, double atol, double tout, long int mxsteps,
                                        int lu_precision, double* yout, long int* counters,
                                        void* stream);

#ifdef __cplusplus
}
#endif

#endif
