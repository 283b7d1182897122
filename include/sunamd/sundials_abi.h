/* -----------------------------------------------------------------------------
 * sundials_abi.h -- the plug-in ABI of the SUNDIALS generic objects, restated.
 *
 * The structs below are laid out field-for-field like the reference's generic
 * objects so that an N_Vector / SUNMatrix / SUNLinearSolver built by this
 * library can be handed to a SUNDIALS integrator unchanged:
 *   struct _generic_N_Vector_Ops      include/sundials/sundials_nvector.h:101-195
 *   struct _generic_N_Vector          include/sundials/sundials_nvector.h:200-205
 *   struct _generic_SUNMatrix_Ops     include/sundials/sundials_matrix.h:89-105
 *   struct _generic_SUNLinearSolver_Ops include/sundials/sundials_linearsolver.h:117-140
 *   struct _generic_SUNNonlinearSolver_Ops include/sundials/sundials_nonlinearsolver.h:103-121
 * Scalar types follow the reference's default build: sunrealtype = double
 * (sundials_types.h:108), sunindextype = int64 (SundialsBuildOptionsPre.cmake:47-52),
 * sunbooleantype = int (sundials_types.h:171), SUNComm = int without MPI
 * (sundials_types.h:253).  Error codes keep the reference's numbering
 * (sundials_errors.h:33-100, sundials_linearsolver.h:222-238,
 * sundials_nonlinearsolver.h:200-204).
 *
 * tests/test_abi.py checks every ops slot's name, return and parameter types
 * and every object member against the reference headers, and the compiled
 * offsets and sizes against the layout those declarations give on LP64.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_SUNDIALS_ABI_H
#define SUNAMD_SUNDIALS_ABI_H

#include <stdint.h>
#include <stdio.h>
#include <float.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar types ------------------------------------------------------- */
typedef double sunrealtype;
typedef int64_t sunindextype;
typedef long int suncountertype;
#define sunbooleantype int
#define SUNFALSE 0
#define SUNTRUE 1
typedef int SUNErrCode;
typedef int SUNComm;
#define SUN_COMM_NULL 0

/* statistics output formats (sundials_types.h:190-194) */
typedef enum
{
  SUN_OUTPUTFORMAT_TABLE,
  SUN_OUTPUTFORMAT_CSV
} SUNOutputFormat;
#define SUN_RCONST(x) x
#define SUN_BIG_REAL DBL_MAX
#define SUN_SMALL_REAL DBL_MIN
#define SUN_UNIT_ROUNDOFF DBL_EPSILON

typedef struct SUNContext_* SUNContext;

/* ---- error codes (same values as the reference enum) -------------------- */
enum
{
  SUN_ERR_MINIMUM = -10000,
  SUN_ERR_ARG_CORRUPT,
  SUN_ERR_ARG_INCOMPATIBLE,
  SUN_ERR_ARG_OUTOFRANGE,
  SUN_ERR_ARG_WRONGTYPE,
  SUN_ERR_ARG_DIMSMISMATCH,
  SUN_ERR_GENERIC,
  SUN_ERR_CORRUPT,
  SUN_ERR_OUTOFRANGE,
  SUN_ERR_FILE_OPEN,
  SUN_ERR_OP_FAIL,
  SUN_ERR_MEM_FAIL,
  SUN_ERR_MALLOC_FAIL,
  SUN_ERR_EXT_FAIL,
  SUN_ERR_DESTROY_FAIL,
  SUN_ERR_NOT_IMPLEMENTED,
  SUN_ERR_USER_FCN_FAIL,
  SUN_ERR_DATANODE_NODENOTFOUND,
  SUN_ERR_PROFILER_MAPFULL,
  SUN_ERR_PROFILER_MAPGET,
  SUN_ERR_PROFILER_MAPINSERT,
  SUN_ERR_PROFILER_MAPKEYNOTFOUND,
  SUN_ERR_PROFILER_MAPSORT,
  SUN_ERR_ADJOINT_STEPPERFAILED,
  SUN_ERR_ADJOINT_STEPPERINVALIDSTOP,
  SUN_ERR_CHECKPOINT_NOT_FOUND,
  SUN_ERR_CHECKPOINT_MISMATCH,
  SUN_ERR_SUNCTX_CORRUPT,
  SUN_ERR_MPI_FAIL,
  SUN_ERR_UNREACHABLE,
  SUN_ERR_UNKNOWN,
  SUN_ERR_MAXIMUM = -1000,
  SUN_SUCCESS     = 0
};

/* linear solver return codes (sundials_linearsolver.h:222-238) */
#define SUNLS_ATIMES_NULL       -804
#define SUNLS_ATIMES_FAIL_UNREC -805
#define SUNLS_PSET_FAIL_UNREC   -806
#define SUNLS_PSOLVE_NULL       -807
#define SUNLS_PSOLVE_FAIL_UNREC -808
#define SUNLS_GS_FAIL           -810
#define SUNLS_QRSOL_FAIL        -811
#define SUNLS_RECOV_FAILURE     800
#define SUNLS_RES_REDUCED       801
#define SUNLS_CONV_FAIL         802
#define SUNLS_ATIMES_FAIL_REC   803
#define SUNLS_PSET_FAIL_REC     804
#define SUNLS_PSOLVE_FAIL_REC   805
#define SUNLS_PACKAGE_FAIL_REC  806
#define SUNLS_QRFACT_FAIL       807
#define SUNLS_LUFACT_FAIL       808

/* nonlinear solver recoverable codes (sundials_nonlinearsolver.h:200-204) */
#define SUN_NLS_CONTINUE   +901
#define SUN_NLS_CONV_RECVR +902

/* ---- N_Vector ------------------------------------------------------------ */
typedef enum
{
  SUNDIALS_NVEC_SERIAL,
  SUNDIALS_NVEC_PARALLEL,
  SUNDIALS_NVEC_OPENMP,
  SUNDIALS_NVEC_PTHREADS,
  SUNDIALS_NVEC_PARHYP,
  SUNDIALS_NVEC_PETSC,
  SUNDIALS_NVEC_CUDA,
  SUNDIALS_NVEC_HIP,
  SUNDIALS_NVEC_SYCL,
  SUNDIALS_NVEC_RAJA,
  SUNDIALS_NVEC_KOKKOS,
  SUNDIALS_NVEC_OPENMPDEV,
  SUNDIALS_NVEC_TRILINOS,
  SUNDIALS_NVEC_MANYVECTOR,
  SUNDIALS_NVEC_MPIMANYVECTOR,
  SUNDIALS_NVEC_MPIPLUSX,
  SUNDIALS_NVEC_CUSTOM
} N_Vector_ID;

typedef struct _generic_N_Vector_Ops* N_Vector_Ops;
typedef struct _generic_N_Vector* N_Vector;
typedef N_Vector* N_Vector_S;

struct _generic_N_Vector_Ops
{
  /* constructors, destructors, and utility operations */
  N_Vector_ID (*nvgetvectorid)(N_Vector);
  N_Vector (*nvclone)(N_Vector);
  N_Vector (*nvcloneempty)(N_Vector);
  void (*nvdestroy)(N_Vector);
  void (*nvspace)(N_Vector, sunindextype*, sunindextype*);
  sunrealtype* (*nvgetarraypointer)(N_Vector);
  sunrealtype* (*nvgetdevicearraypointer)(N_Vector);
  void (*nvsetarraypointer)(sunrealtype*, N_Vector);
  SUNComm (*nvgetcommunicator)(N_Vector);
  sunindextype (*nvgetlength)(N_Vector);
  sunindextype (*nvgetlocallength)(N_Vector);
  /* standard vector operations */
  void (*nvlinearsum)(sunrealtype, N_Vector, sunrealtype, N_Vector, N_Vector);
  void (*nvconst)(sunrealtype, N_Vector);
  void (*nvprod)(N_Vector, N_Vector, N_Vector);
  void (*nvdiv)(N_Vector, N_Vector, N_Vector);
  void (*nvscale)(sunrealtype, N_Vector, N_Vector);
  void (*nvabs)(N_Vector, N_Vector);
  void (*nvinv)(N_Vector, N_Vector);
  void (*nvaddconst)(N_Vector, sunrealtype, N_Vector);
  sunrealtype (*nvdotprod)(N_Vector, N_Vector);
  sunrealtype (*nvmaxnorm)(N_Vector);
  sunrealtype (*nvwrmsnorm)(N_Vector, N_Vector);
  sunrealtype (*nvwrmsnormmask)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvmin)(N_Vector);
  sunrealtype (*nvwl2norm)(N_Vector, N_Vector);
  sunrealtype (*nvl1norm)(N_Vector);
  void (*nvcompare)(sunrealtype, N_Vector, N_Vector);
  sunbooleantype (*nvinvtest)(N_Vector, N_Vector);
  sunbooleantype (*nvconstrmask)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvminquotient)(N_Vector, N_Vector);
  /* fused vector operations */
  SUNErrCode (*nvlinearcombination)(int, sunrealtype*, N_Vector*, N_Vector);
  SUNErrCode (*nvscaleaddmulti)(int, sunrealtype*, N_Vector, N_Vector*, N_Vector*);
  SUNErrCode (*nvdotprodmulti)(int, N_Vector, N_Vector*, sunrealtype*);
  /* vector array operations */
  SUNErrCode (*nvlinearsumvectorarray)(int, sunrealtype, N_Vector*, sunrealtype,
                                       N_Vector*, N_Vector*);
  SUNErrCode (*nvscalevectorarray)(int, sunrealtype*, N_Vector*, N_Vector*);
  SUNErrCode (*nvconstvectorarray)(int, sunrealtype, N_Vector*);
  SUNErrCode (*nvwrmsnormvectorarray)(int, N_Vector*, N_Vector*, sunrealtype*);
  SUNErrCode (*nvwrmsnormmaskvectorarray)(int, N_Vector*, N_Vector*, N_Vector,
                                          sunrealtype*);
  SUNErrCode (*nvscaleaddmultivectorarray)(int, int, sunrealtype*, N_Vector*,
                                           N_Vector**, N_Vector**);
  SUNErrCode (*nvlinearcombinationvectorarray)(int, int, sunrealtype*,
                                               N_Vector**, N_Vector*);
  /* local reduction kernels (no parallel communication) */
  sunrealtype (*nvdotprodlocal)(N_Vector, N_Vector);
  sunrealtype (*nvmaxnormlocal)(N_Vector);
  sunrealtype (*nvminlocal)(N_Vector);
  sunrealtype (*nvl1normlocal)(N_Vector);
  sunbooleantype (*nvinvtestlocal)(N_Vector, N_Vector);
  sunbooleantype (*nvconstrmasklocal)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvminquotientlocal)(N_Vector, N_Vector);
  sunrealtype (*nvwsqrsumlocal)(N_Vector, N_Vector);
  sunrealtype (*nvwsqrsummasklocal)(N_Vector, N_Vector, N_Vector);
  /* single buffer reduction operations */
  SUNErrCode (*nvdotprodmultilocal)(int, N_Vector, N_Vector*, sunrealtype*);
  SUNErrCode (*nvdotprodmultiallreduce)(int, N_Vector, sunrealtype*);
  /* XBraid interface operations */
  SUNErrCode (*nvbufsize)(N_Vector, sunindextype*);
  SUNErrCode (*nvbufpack)(N_Vector, void*);
  SUNErrCode (*nvbufunpack)(N_Vector, void*);
  /* debugging functions */
  void (*nvprint)(N_Vector);
  void (*nvprintfile)(N_Vector, FILE*);
};

struct _generic_N_Vector
{
  void* content;
  N_Vector_Ops ops;
  SUNContext sunctx;
};

/* ---- SUNMatrix ----------------------------------------------------------- */
typedef enum
{
  SUNMATRIX_DENSE,
  SUNMATRIX_MAGMADENSE,
  SUNMATRIX_ONEMKLDENSE,
  SUNMATRIX_BAND,
  SUNMATRIX_SPARSE,
  SUNMATRIX_SLUNRLOC,
  SUNMATRIX_CUSPARSE,
  SUNMATRIX_GINKGO,
  SUNMATRIX_GINKGOBATCH,
  SUNMATRIX_KOKKOSDENSE,
  SUNMATRIX_CUSTOM
} SUNMatrix_ID;

typedef struct _generic_SUNMatrix_Ops* SUNMatrix_Ops;
typedef struct _generic_SUNMatrix* SUNMatrix;

struct _generic_SUNMatrix_Ops
{
  SUNMatrix_ID (*getid)(SUNMatrix);
  SUNMatrix (*clone)(SUNMatrix);
  void (*destroy)(SUNMatrix);
  SUNErrCode (*zero)(SUNMatrix);
  SUNErrCode (*copy)(SUNMatrix, SUNMatrix);
  SUNErrCode (*scaleadd)(sunrealtype, SUNMatrix, SUNMatrix);
  SUNErrCode (*scaleaddi)(sunrealtype, SUNMatrix);
  SUNErrCode (*matvecsetup)(SUNMatrix);
  SUNErrCode (*matvec)(SUNMatrix, N_Vector, N_Vector);
  SUNErrCode (*mathermitiantransposevec)(SUNMatrix, N_Vector, N_Vector);
  SUNErrCode (*space)(SUNMatrix, long int*, long int*);
};

struct _generic_SUNMatrix
{
  void* content;
  SUNMatrix_Ops ops;
  SUNContext sunctx;
};

/* ---- SUNLinearSolver ----------------------------------------------------- */
typedef int (*SUNATimesFn)(void* A_data, N_Vector v, N_Vector z);
typedef int (*SUNPSetupFn)(void* P_data);
typedef int (*SUNPSolveFn)(void* P_data, N_Vector r, N_Vector z,
                           sunrealtype tol, int lr);

enum { SUN_PREC_NONE, SUN_PREC_LEFT, SUN_PREC_RIGHT, SUN_PREC_BOTH };
enum { SUN_MODIFIED_GS = 1, SUN_CLASSICAL_GS = 2 };

typedef enum
{
  SUNLINEARSOLVER_DIRECT,
  SUNLINEARSOLVER_ITERATIVE,
  SUNLINEARSOLVER_MATRIX_ITERATIVE,
  SUNLINEARSOLVER_MATRIX_EMBEDDED
} SUNLinearSolver_Type;

typedef enum
{
  SUNLINEARSOLVER_BAND,
  SUNLINEARSOLVER_DENSE,
  SUNLINEARSOLVER_KLU,
  SUNLINEARSOLVER_LAPACKBAND,
  SUNLINEARSOLVER_LAPACKDENSE,
  SUNLINEARSOLVER_PCG,
  SUNLINEARSOLVER_SPBCGS,
  SUNLINEARSOLVER_SPFGMR,
  SUNLINEARSOLVER_SPGMR,
  SUNLINEARSOLVER_SPTFQMR,
  SUNLINEARSOLVER_SUPERLUDIST,
  SUNLINEARSOLVER_SUPERLUMT,
  SUNLINEARSOLVER_CUSOLVERSP_BATCHQR,
  SUNLINEARSOLVER_MAGMADENSE,
  SUNLINEARSOLVER_ONEMKLDENSE,
  SUNLINEARSOLVER_GINKGO,
  SUNLINEARSOLVER_GINKGOBATCH,
  SUNLINEARSOLVER_KOKKOSDENSE,
  SUNLINEARSOLVER_CUSTOM
} SUNLinearSolver_ID;

typedef struct _generic_SUNLinearSolver_Ops* SUNLinearSolver_Ops;
typedef struct _generic_SUNLinearSolver* SUNLinearSolver;

struct _generic_SUNLinearSolver_Ops
{
  SUNLinearSolver_Type (*gettype)(SUNLinearSolver);
  SUNLinearSolver_ID (*getid)(SUNLinearSolver);
  SUNErrCode (*setatimes)(SUNLinearSolver, void*, SUNATimesFn);
  SUNErrCode (*setpreconditioner)(SUNLinearSolver, void*, SUNPSetupFn, SUNPSolveFn);
  SUNErrCode (*setscalingvectors)(SUNLinearSolver, N_Vector, N_Vector);
  SUNErrCode (*setoptions)(SUNLinearSolver, const char* LSid,
                           const char* file_name, int argc, char* argv[]);
  SUNErrCode (*setzeroguess)(SUNLinearSolver, sunbooleantype);
  SUNErrCode (*initialize)(SUNLinearSolver);
  int (*setup)(SUNLinearSolver, SUNMatrix);
  int (*solve)(SUNLinearSolver, SUNMatrix, N_Vector, N_Vector, sunrealtype);
  int (*numiters)(SUNLinearSolver);
  sunrealtype (*resnorm)(SUNLinearSolver);
  sunindextype (*lastflag)(SUNLinearSolver);
  SUNErrCode (*space)(SUNLinearSolver, long int*, long int*);
  N_Vector (*resid)(SUNLinearSolver);
  SUNErrCode (*free)(SUNLinearSolver);
};

struct _generic_SUNLinearSolver
{
  void* content;
  SUNLinearSolver_Ops ops;
  SUNContext sunctx;
};

/* ---- SUNNonlinearSolver -------------------------------------------------- */
typedef struct _generic_SUNNonlinearSolver_Ops* SUNNonlinearSolver_Ops;
typedef struct _generic_SUNNonlinearSolver* SUNNonlinearSolver;

typedef int (*SUNNonlinSolSysFn)(N_Vector y, N_Vector F, void* mem);
typedef int (*SUNNonlinSolLSetupFn)(sunbooleantype jbad, sunbooleantype* jcur,
                                    void* mem);
typedef int (*SUNNonlinSolLSolveFn)(N_Vector b, void* mem);
typedef int (*SUNNonlinSolConvTestFn)(SUNNonlinearSolver NLS, N_Vector y,
                                      N_Vector del, sunrealtype tol,
                                      N_Vector ewt, void* mem);

typedef enum
{
  SUNNONLINEARSOLVER_ROOTFIND,
  SUNNONLINEARSOLVER_FIXEDPOINT
} SUNNonlinearSolver_Type;

struct _generic_SUNNonlinearSolver_Ops
{
  SUNNonlinearSolver_Type (*gettype)(SUNNonlinearSolver);
  SUNErrCode (*initialize)(SUNNonlinearSolver);
  int (*setup)(SUNNonlinearSolver, N_Vector, void*);
  int (*solve)(SUNNonlinearSolver, N_Vector, N_Vector, N_Vector, sunrealtype,
               sunbooleantype, void*);
  SUNErrCode (*free)(SUNNonlinearSolver);
  SUNErrCode (*setsysfn)(SUNNonlinearSolver, SUNNonlinSolSysFn);
  SUNErrCode (*setlsetupfn)(SUNNonlinearSolver, SUNNonlinSolLSetupFn);
  SUNErrCode (*setlsolvefn)(SUNNonlinearSolver, SUNNonlinSolLSolveFn);
  SUNErrCode (*setctestfn)(SUNNonlinearSolver, SUNNonlinSolConvTestFn, void*);
  SUNErrCode (*setoptions)(SUNNonlinearSolver NLS, const char* NLSid,
                           const char* file_name, int argc, char* argv[]);
  SUNErrCode (*setmaxiters)(SUNNonlinearSolver, int);
  SUNErrCode (*getnumiters)(SUNNonlinearSolver, long int*);
  SUNErrCode (*getcuriter)(SUNNonlinearSolver, int*);
  SUNErrCode (*getnumconvfails)(SUNNonlinearSolver, long int*);
};

struct _generic_SUNNonlinearSolver
{
  void* content;
  SUNNonlinearSolver_Ops ops;
  SUNContext sunctx;
};

#ifdef __cplusplus
}
#endif

#endif /* SUNAMD_SUNDIALS_ABI_H */
