"""ctypes bindings for the in-tree C-ABI libraries.

libsunamd_core.so   generic layer, Newton, BDF integrator, problem generator
libsundials_amd.so  HIP N_Vector, block-diagonal matrix + batched LU, kernels

The HIP library is the product path: `hip()` raises if it has not been built,
there is no fallback implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")

c_real = C.c_double
c_index = C.c_int64
c_void_p = C.c_void_p

_core = None
_hip = None


class MassAction(C.Structure):
    """Mirror of SUNAMD_MassAction (include/sunamd/massaction.h)."""
    _fields_ = [
        ("m", C.c_int),
        ("R", C.c_int),
        ("ncells", C.c_int64),
        ("seed", C.c_uint64),
        ("sigma", C.c_double),
        ("idx", C.POINTER(C.c_int)),
        ("kf0", C.POINTER(C.c_double)),
        ("kb0", C.POINTER(C.c_double)),
        ("kf", C.POINTER(C.c_double)),
        ("kb", C.POINTER(C.c_double)),
        ("y0", C.POINTER(C.c_double)),
    ]


class SolveStats(C.Structure):
    """Mirror of SUNAMD_SolveStats (include/sunamd/massaction_mi355.h)."""
    _fields_ = [
        ("nst", C.c_long), ("nfe", C.c_long), ("nsetups", C.c_long), ("netf", C.c_long),
        ("nni", C.c_long), ("nnf", C.c_long), ("ncfn", C.c_long), ("nje", C.c_long),
        ("qlast", C.c_int), ("qcur", C.c_int),
        ("hinused", C.c_double), ("hlast", C.c_double), ("hcur", C.c_double),
        ("tcur", C.c_double), ("wall_s", C.c_double), ("launches", C.c_long),
        ("flag", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BATCH_MAXB = 8
BS_NAMES = ("nst", "nfe", "nsetups", "netf", "nni", "nnf", "ncfn", "nje", "qlast", "flag")
BS_COUNT = len(BS_NAMES)


class TridiagData(C.Structure):
    """Mirror of SUNAMD_TridiagData (include/sunamd/tridiag_mi355.h)."""
    _fields_ = [("N", C.c_int64), ("d", C.c_void_p), ("s1", C.c_void_p), ("s2", C.c_void_p)]


class Adr(C.Structure):
    """SUNAMD_Adr (include/sunamd/imex_adr.h)."""
    _fields_ = [
        ("n", C.c_int),
        ("ncells", C.c_int64),
        ("cx", C.c_double),
        ("cy", C.c_double),
        ("cz", C.c_double),
        ("d", C.c_double),
        ("A", C.c_double),
        ("B", C.c_double),
        ("kappa", C.POINTER(C.c_double)),
        ("y0", C.POINTER(C.c_double)),
        ("seed", C.c_uint64),
    ]


class AdrOptions(C.Structure):
    """SUNAMD_AdrOptions (include/sunamd/imex_adr.h)."""
    _fields_ = [
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("tf", C.c_double),
        ("maxl", C.c_int),
        ("gstype", C.c_int),
        ("controller", C.c_int),
        ("sequential", C.c_int),
        ("prec_fp32", C.c_int),
        ("mxsteps", C.c_long),
    ]


ADR_STATS = 16


class QRData(C.Structure):
    """struct _SUNQRData (include/sunamd/sundials_iterative.h)."""
    _fields_ = [("vtemp", C.c_void_p), ("vtemp2", C.c_void_p),
                ("temp_array", C.POINTER(C.c_double))]


class KinOptions(C.Structure):
    """SUNAMD_KinOptions (include/sunamd/kin_fp.h)."""
    _fields_ = [
        ("ncells", C.c_long),
        ("variant", C.c_int),
        ("perturb", C.c_int),
        ("tol", C.c_double),
        ("maxiter", C.c_long),
        ("m_aa", C.c_long),
        ("delay_aa", C.c_long),
        ("orth_aa", C.c_int),
        ("damping_fp", C.c_double),
        ("damping_aa", C.c_double),
        ("damping_fn", C.c_int),
        ("depth_fn", C.c_int),
        ("sequential", C.c_int),
        ("fused", C.c_int),
    ]


KIN_STATS = 5


class KinNewtonOptions(C.Structure):
    """SUNAMD_KinNewtonOptions (include/sunamd/kin_newton.h)."""
    _fields_ = [
        ("ncells", C.c_long),
        ("variant", C.c_int),
        ("guess", C.c_int),
        ("perturb", C.c_int),
        ("strategy", C.c_int),
        ("msbset", C.c_long),
        ("fnormtol", C.c_double),
        ("scsteptol", C.c_double),
        ("maxiter", C.c_long),
        ("jac", C.c_int),
        ("sequential", C.c_int),
        ("fused", C.c_int),
    ]


KINN_STATS = 8
KINN_STAT_NAMES = ("flag", "nni", "nfe", "nje", "nfeD", "nbcf", "nbacktrack", "wall")
KIN_NONE, KIN_LINESEARCH, KIN_PICARD, KIN_FP = 0, 1, 2, 3
ADR_STAT_NAMES = ("nst", "attempts", "netf", "ncfn", "nfe", "nfi", "nni", "nnf", "nsetups",
                  "nli", "npe", "nps", "njtimes", "ncfl", "wall", "hlast")


class TrainingLogInfo(C.Structure):
    """Mirror of SUNAMD_TrainingLogInfo (include/sunamd/learned_predictor.h)."""
    _fields_ = [("version", C.c_int), ("rows", C.c_long), ("seed", C.c_uint64),
                ("rtol", C.c_double), ("atol", C.c_double), ("stride", C.c_int)]


class Batch(C.Structure):
    """Mirror of SUNAMD_Batch (include/sunamd/batched_bdf.h)."""
    _fields_ = [
        ("nbuckets", C.c_int),
        ("bucket", C.POINTER(MassAction) * BATCH_MAXB),
        ("nsys", C.c_int64),
        ("sys_bucket", C.POINTER(C.c_int)),
        ("sys_cell", C.POINTER(C.c_int64)),
        ("sys_yoff", C.POINTER(C.c_int64)),
        ("ytotal", C.c_int64),
        ("seed", C.c_uint64),
    ]


def _sig(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def core():
    """Load libsunamd_core.so (plain C)."""
    global _core
    if _core is not None:
        return _core
    path = os.path.join(LIBDIR, "libsunamd_core.so")
    if not os.path.exists(path):
        raise RuntimeError("libsunamd_core.so not built: run __graft_entry__.build()")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    P = C.POINTER
    _sig(lib, "SUNContext_Create", C.c_int, [C.c_int, P(c_void_p)])
    _sig(lib, "SUNContext_Free", C.c_int, [P(c_void_p)])
    _sig(lib, "SUNContext_GetLastError", C.c_int, [c_void_p])
    for n in ("N_VDestroy",):
        _sig(lib, n, None, [c_void_p])
    _sig(lib, "N_VClone", c_void_p, [c_void_p])
    _sig(lib, "N_VGetLength", c_index, [c_void_p])
    _sig(lib, "N_VGetArrayPointer", P(c_real), [c_void_p])
    _sig(lib, "N_VGetDeviceArrayPointer", c_void_p, [c_void_p])
    _sig(lib, "N_VLinearSum", None, [c_real, c_void_p, c_real, c_void_p, c_void_p])
    _sig(lib, "N_VConst", None, [c_real, c_void_p])
    for n in ("N_VProd", "N_VDiv"):
        _sig(lib, n, None, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "N_VScale", None, [c_real, c_void_p, c_void_p])
    for n in ("N_VAbs", "N_VInv"):
        _sig(lib, n, None, [c_void_p, c_void_p])
    _sig(lib, "N_VAddConst", None, [c_void_p, c_real, c_void_p])
    _sig(lib, "N_VDotProd", c_real, [c_void_p, c_void_p])
    _sig(lib, "N_VMaxNorm", c_real, [c_void_p])
    _sig(lib, "N_VWrmsNorm", c_real, [c_void_p, c_void_p])
    _sig(lib, "N_VWrmsNormMask", c_real, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "N_VMin", c_real, [c_void_p])
    _sig(lib, "N_VWL2Norm", c_real, [c_void_p, c_void_p])
    _sig(lib, "N_VL1Norm", c_real, [c_void_p])
    _sig(lib, "N_VCompare", None, [c_real, c_void_p, c_void_p])
    _sig(lib, "N_VInvTest", C.c_int, [c_void_p, c_void_p])
    _sig(lib, "N_VConstrMask", C.c_int, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "N_VMinQuotient", c_real, [c_void_p, c_void_p])
    _sig(lib, "N_VLinearCombination", C.c_int, [C.c_int, P(c_real), P(c_void_p), c_void_p])
    _sig(lib, "N_VScaleAddMulti", C.c_int,
         [C.c_int, P(c_real), c_void_p, P(c_void_p), P(c_void_p)])
    _sig(lib, "N_VDotProdMulti", C.c_int, [C.c_int, c_void_p, P(c_void_p), P(c_real)])
    _sig(lib, "N_VScaleVectorArray", C.c_int,
         [C.c_int, P(c_real), P(c_void_p), P(c_void_p)])
    _sig(lib, "N_VWrmsNormVectorArray", C.c_int,
         [C.c_int, P(c_void_p), P(c_void_p), P(c_real)])
    _sig(lib, "SUNMatDestroy", None, [c_void_p])
    _sig(lib, "SUNMatZero", C.c_int, [c_void_p])
    _sig(lib, "SUNMatCopy", C.c_int, [c_void_p, c_void_p])
    _sig(lib, "SUNMatScaleAddI", C.c_int, [c_real, c_void_p])
    _sig(lib, "SUNMatScaleAdd", C.c_int, [c_real, c_void_p, c_void_p])
    _sig(lib, "SUNMatMatvec", C.c_int, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNLinSolSetup", C.c_int, [c_void_p, c_void_p])
    _sig(lib, "SUNLinSolSolve", C.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_real])
    _sig(lib, "SUNLinSolLastFlag", c_index, [c_void_p])
    _sig(lib, "SUNLinSolFree", C.c_int, [c_void_p])
    _sig(lib, "SUNLinSolInitialize", C.c_int, [c_void_p])
    _sig(lib, "SUNAMD_Uniform", c_real, [C.c_uint64, C.c_uint64, C.c_uint64])
    _sig(lib, "SUNAMD_MassActionCreate", P(MassAction),
         [C.c_int, C.c_int, C.c_int64, C.c_uint64, C.c_double])
    _sig(lib, "SUNAMD_MassActionCreateRange", P(MassAction),
         [C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_uint64, C.c_double])
    _sig(lib, "SUNAMD_MassActionTile", P(MassAction), [P(MassAction), C.c_int64])
    _sig(lib, "SUNAMD_MassActionFree", None, [P(MassAction)])
    _sig(lib, "SUNAMD_MassActionRhsHost", None,
         [P(MassAction), P(c_real), P(c_real), C.c_int64, C.c_int64])
    _sig(lib, "SUNAMD_MassActionJacHost", None,
         [P(MassAction), P(c_real), P(c_real), C.c_int64, C.c_int64])
    _sig(lib, "SUNLinSolSetATimes", C.c_int, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNLinSolSetPreconditioner", C.c_int, [c_void_p, c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNLinSolSetScalingVectors", C.c_int, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNLinSolSetZeroGuess", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "SUNLinSolNumIters", C.c_int, [c_void_p])
    _sig(lib, "SUNLinSolResNorm", c_real, [c_void_p])
    _sig(lib, "SUNLinSolGetType", C.c_int, [c_void_p])
    _sig(lib, "SUNLinSolGetID", C.c_int, [c_void_p])
    _sig(lib, "SUNLinSolSpace", C.c_int, [c_void_p, P(C.c_long), P(C.c_long)])
    _sig(lib, "SUNLinSol_SPGMR", c_void_p, [c_void_p, C.c_int, C.c_int, c_void_p])
    _sig(lib, "SUNLinSol_SPGMRSetPrecType", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "SUNLinSol_SPGMRSetGSType", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "SUNLinSol_SPGMRSetMaxRestarts", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "SUNLinSol_SPGMRSetClassicalGSFn", C.c_int, [c_void_p, c_void_p])
    _sig(lib, "SUNModifiedGS", C.c_int, [P(c_void_p), P(P(c_real)), C.c_int, C.c_int, P(c_real)])
    for name in ("SUNQRAdd_MGS", "SUNQRAdd_ICWY", "SUNQRAdd_CGS2", "SUNQRAdd_DCGS2"):
        _sig(lib, name, C.c_int, [P(c_void_p), P(c_real), c_void_p, C.c_int, C.c_int, P(QRData)])
    _sig(lib, "SUNClassicalGS", C.c_int, [P(c_void_p), P(P(c_real)), C.c_int, C.c_int, P(c_real),
                                          P(c_real), P(c_void_p)])
    _sig(lib, "SUNQRfact", C.c_int, [C.c_int, P(P(c_real)), P(c_real), C.c_int])
    _sig(lib, "SUNQRsol", C.c_int, [C.c_int, P(P(c_real)), P(c_real), P(c_real)])
    _sig(lib, "SUNRCompareTol", C.c_int, [c_real, c_real, c_real])
    _sig(lib, "SUNAMD_Pow", c_real, [c_real, c_real])
    _sig(lib, "SUNAMD_AdrCreate", P(Adr), [C.c_int, C.c_uint64])
    _sig(lib, "SUNAMD_AdrFree", None, [P(Adr)])
    _sig(lib, "SUNAMD_TrainingLogWrite", C.c_int,
         [C.c_char_p, c_void_p, c_void_p, c_void_p, C.c_long, P(TrainingLogInfo)])
    _sig(lib, "SUNAMD_TrainingLogInfoRead", C.c_int, [C.c_char_p, P(TrainingLogInfo)])
    _sig(lib, "SUNAMD_TrainingLogRead", C.c_int,
         [C.c_char_p, C.c_long, C.c_long, c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNAMD_ARKTable", C.c_int, [C.c_int, C.c_int, P(C.c_int)] + [P(c_real)] * 4 +
         [P(C.c_int), P(C.c_int)])
    _sig(lib, "SUNAMD_ARK324Table", C.c_int, [C.c_int, P(c_real), P(c_real), P(c_real),
                                                 P(c_real), P(C.c_int), P(C.c_int)])
    _sig(lib, "SUNAMD_PowExact", c_real, [c_real, c_real])
    _sig(lib, "SUNAMD_PowFastParts", C.c_int, [c_real, c_real] + [P(c_real)] * 4 + [P(C.c_int)])
    _sig(lib, "SUNAMD_BatchCreate", P(Batch),
         [C.c_int64, C.c_uint64, C.c_int, P(C.c_int), P(c_real), c_real, c_real])
    _sig(lib, "SUNAMD_BatchFree", None, [P(Batch)])
    _core = lib
    return lib


def _batched_sigs(lib):
    """ctypes signatures of the batched per-system BDF entry points."""
    P = C.POINTER
    _sig(lib, "SUNAMD_BatchedBDFCreate", c_void_p, [P(Batch), c_real, c_real, C.c_int])
    _sig(lib, "SUNAMD_BatchedBDFSetMaxNumSteps", C.c_int, [c_void_p, C.c_long])
    _sig(lib, "SUNAMD_BatchedBDFRun", C.c_int, [c_void_p, c_real, P(c_real)])
    _sig(lib, "SUNAMD_BatchedBDFGetResults", C.c_int, [c_void_p, P(c_real), P(C.c_long)])
    _sig(lib, "SUNAMD_BatchedBDFGetKernelTimes", C.c_int, [c_void_p, P(c_real)])
    _sig(lib, "SUNAMD_BatchedBDFSetLearnedPredictor", C.c_int, [c_void_p, c_void_p, C.c_int])
    _sig(lib, "SUNAMD_BatchedBDFSetOccupancy", C.c_int, [c_void_p, P(C.c_int)])
    _sig(lib, "SUNAMD_BatchedBDFSetGridCap", C.c_int, [c_void_p, P(C.c_int)])
    _sig(lib, "SUNAMD_BatchedBDFSetTrainingLog", C.c_int,
         [c_void_p, c_void_p, c_void_p, c_void_p, C.c_long, C.c_int])
    _sig(lib, "SUNAMD_BatchedBDFGetTrainingLogCount", C.c_long, [c_void_p])
    _sig(lib, "SUNAMD_BatchedBDFGetLearnedStats", C.c_int, [c_void_p, P(C.c_long), P(C.c_long)])
    _sig(lib, "SUNAMD_BatchedBDFGetLearnedStatsByBucket", C.c_int,
         [c_void_p, P(C.c_long), P(C.c_long)])
    _sig(lib, "SUNAMD_BatchedBDFSolveBucketNoAlloc", C.c_int,
         [C.c_int, C.c_int, C.c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_real, c_real,
          c_real, C.c_long, C.c_int, c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNAMD_BatchedBDFGetPhaseTicks", C.c_int, [c_void_p, P(C.c_ulonglong)])
    _sig(lib, "SUNAMD_BatchedBDFFree", None, [c_void_p])


_hip_phase = None


def hip_phase():
    """Load libsundials_amd_phase.so: the batched integrator alone, built with
    the phase clock (a profiling aid; same entry points as hip())."""
    global _hip_phase
    if _hip_phase is not None:
        return _hip_phase
    core()
    path = os.path.join(LIBDIR, "libsundials_amd_phase.so")
    if not os.path.exists(path):
        raise RuntimeError("libsundials_amd_phase.so not built: run __graft_entry__.build()")
    lib = C.CDLL(path)
    _batched_sigs(lib)
    _hip_phase = lib
    return lib


def hip():
    """Load libsundials_amd.so (the HIP product path); raises if missing."""
    global _hip
    if _hip is not None:
        return _hip
    core()
    path = os.path.join(LIBDIR, "libsundials_amd.so")
    if not os.path.exists(path):
        raise RuntimeError("libsundials_amd.so not built: run __graft_entry__.build()")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    P = C.POINTER
    _sig(lib, "N_VNew_Mi355", c_void_p, [c_index, c_void_p])
    _sig(lib, "N_VNewEmpty_Mi355", c_void_p, [c_index, c_void_p])
    _sig(lib, "N_VMake_Mi355", c_void_p, [c_index, c_void_p, c_void_p, c_void_p])
    for n in ("N_VCopyToDevice_Mi355", "N_VCopyFromDevice_Mi355", "N_VSynchronize_Mi355"):
        _sig(lib, n, C.c_int, [c_void_p])
    _sig(lib, "N_VCopyFromHostArray_Mi355", C.c_int, [c_void_p, P(c_real)])
    _sig(lib, "N_VCopyToHostArray_Mi355", C.c_int, [c_void_p, P(c_real)])
    _sig(lib, "N_VGetStream_Mi355", c_void_p, [c_void_p])
    _sig(lib, "N_VEnableFusedOps_Mi355", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "N_VGetNumLaunches_Mi355", C.c_long, [c_void_p])
    _sig(lib, "SUNMatrix_Mi355BlockDense", c_void_p, [c_index, c_index, c_void_p, c_void_p])
    _sig(lib, "SUNMatrix_Mi355BlockDense_DeviceData", c_void_p, [c_void_p])
    _sig(lib, "SUNMatrix_Mi355BlockDense_HostData", P(c_real), [c_void_p])
    _sig(lib, "SUNMatrix_Mi355BlockDense_CopyToDevice", C.c_int, [c_void_p])
    _sig(lib, "SUNMatrix_Mi355BlockDense_CopyFromDevice", C.c_int, [c_void_p])
    _sig(lib, "SUNLinSol_Mi355BlockLU", c_void_p, [c_void_p, c_void_p, C.c_int, c_void_p])
    _sig(lib, "SUNLinSol_Mi355BlockLU_DevicePivots", c_void_p, [c_void_p])
    _sig(lib, "SUNAMD_MassActionToDevice", c_void_p, [P(MassAction), c_void_p])
    _sig(lib, "SUNAMD_MassActionDevFree", None, [c_void_p])
    _sig(lib, "SUNAMD_MassActionRhsDevArrays", C.c_int, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNAMD_MassActionJacDevArrays", C.c_int, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNAMD_MassActionSolveCoupled_Mi355", C.c_int,
         [P(MassAction), c_real, c_real, c_real, C.c_int, C.c_int, P(c_real), P(SolveStats)])
    _sig(lib, "SUNAMD_CoupledCreate_Mi355", c_void_p,
         [P(MassAction), c_real, c_real, C.c_int, C.c_int])
    _sig(lib, "N_VSetReductionOrder_Mi355", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "N_VNewWithStream_Mi355", c_void_p, [c_index, c_void_p, c_void_p])
    _sig(lib, "SUNAMD_CoupledRun_Mi355", C.c_int, [c_void_p, c_real, P(SolveStats)])
    _sig(lib, "SUNAMD_CoupledGetState_Mi355", C.c_int, [c_void_p, P(c_real)])
    _sig(lib, "SUNAMD_CoupledFree_Mi355", None, [c_void_p])
    _sig(lib, "SUNAMD_CoupledSetFusedKernels_Mi355", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "SUNAMD_NVectorIsMi355", C.c_int, [c_void_p])
    _sig(lib, "SUNLinSol_Mi355BlockLU_GetPivots", C.c_int, [c_void_p, P(C.c_int)])
    _sig(lib, "SUNAMD_ProfilerEnable", None, [C.c_int])
    _sig(lib, "SUNAMD_ProfilerReset", None, [])
    _sig(lib, "SUNAMD_ProfilerGet", C.c_int, [C.c_int, P(C.c_double), P(C.c_long)])
    _sig(lib, "SUNAMD_BenchVectorOps", C.c_int, [c_index, C.c_int, C.c_int, P(c_real)])
    _sig(lib, "SUNAMD_BenchBlockLU", C.c_int, [C.c_long, C.c_int, C.c_int, C.c_int, P(c_real)])
    _sig(lib, "SUNLinSol_Mi355BlockLUSetMFMA", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "SUNAMD_CoupledSetDeviceStepping_Mi355", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "SUNAMD_CoupledDeviceLastShape", C.c_int, [P(C.c_int), P(C.c_int), P(C.c_int)])
    _sig(lib, "SUNLinSol_Mi355BlockLUSetupScaled", C.c_int,
         [c_void_p, c_void_p, c_void_p, c_real, C.c_int])
    _sig(lib, "SUNLinSol_Mi355BlockLUSolveScaled", C.c_int,
         [c_void_p, c_void_p, c_void_p, c_real])
    _sig(lib, "SUNAMD_ClassicalGS_Mi355", C.c_int, [P(c_void_p), P(P(c_real)), C.c_int, C.c_int,
                                                    P(c_real), P(c_real), P(c_void_p)])
    _sig(lib, "SUNAMD_TridiagATimes_Mi355", C.c_int, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNAMD_TridiagPSetup_Mi355", C.c_int, [c_void_p])
    _sig(lib, "SUNAMD_TridiagPSolve_Mi355", C.c_int, [c_void_p, c_void_p, c_void_p, c_real, C.c_int])
    _sig(lib, "SUNAMD_AdrSolve_Mi355", C.c_int, [P(Adr), P(AdrOptions), P(c_real), P(c_real)])
    _sig(lib, "SUNAMD_KinSolve_Mi355", C.c_int, [P(KinOptions), P(c_real), P(c_real)])
    _sig(lib, "SUNAMD_KinNewtonSolve_Mi355", C.c_int,
         [P(KinNewtonOptions), P(c_real), P(c_real)])
    _batched_sigs(lib)
    PFl = P(C.c_float)
    _sig(lib, "SUNAMD_MLPCreate", c_void_p, [C.c_uint64, C.c_int])
    _sig(lib, "SUNAMD_MLPFree", None, [c_void_p])
    _sig(lib, "SUNAMD_MLPGetParams", C.c_int, [c_void_p, PFl])
    _sig(lib, "SUNAMD_MLPSetParams", C.c_int, [c_void_p, PFl])
    _sig(lib, "SUNAMD_MLPDeviceWeights", c_void_p, [c_void_p])
    _sig(lib, "SUNAMD_MLPForwardWave", C.c_int, [c_void_p, PFl, C.c_int, PFl])
    _sig(lib, "SUNAMD_MLPForward", C.c_int, [c_void_p, PFl, C.c_int, PFl])
    _sig(lib, "SUNAMD_MLPLossGrad", C.c_int, [c_void_p, PFl, PFl, PFl, C.c_int, PFl, PFl])
    _sig(lib, "SUNAMD_MLPTrain", C.c_int,
         [c_void_p, c_void_p, c_void_p, c_void_p, C.c_long, C.c_int, C.c_int, C.c_float, PFl])
    _hip = lib
    return lib


def load_oracle():
    """Load the CPU oracle (test infrastructure only)."""
    core()
    root = os.path.dirname(_HERE)
    path = os.path.join(root, "oracle", "_build", "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle not built: run __graft_entry__.build()")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    P = C.POINTER
    _sig(lib, "N_VNew_Serial", c_void_p, [c_index, c_void_p])
    _sig(lib, "N_VMake_Serial", c_void_p, [c_index, P(c_real), c_void_p])
    _sig(lib, "N_VEnableFusedOps_Serial", C.c_int, [c_void_p, C.c_int])
    _sig(lib, "SUNDenseMatrix", c_void_p, [c_index, c_index, c_void_p])
    _sig(lib, "SUNDenseMatrix_Data", P(c_real), [c_void_p])
    _sig(lib, "SUNLinSol_Dense", c_void_p, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "SUNBlockDenseMatrix_Oracle", c_void_p, [c_index, c_index, c_void_p])
    _sig(lib, "SUNBlockDenseMatrix_Oracle_Data", P(c_real), [c_void_p])
    _sig(lib, "SUNLinSol_BlockDense_Oracle", c_void_p, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "oracle_block_getrf", c_index, [P(c_real), c_index, P(c_index)])
    _sig(lib, "oracle_block_getrs", None, [P(c_real), c_index, P(c_index), P(c_real)])
    _sig(lib, "oracle_roberts_run", C.c_int, [C.c_int, P(c_real)])
    _sig(lib, "oracle_roberts_stats", C.c_int, [C.c_int, C.c_char_p, C.c_char_p])
    _sig(lib, "oracle_massaction_solve_coupled", C.c_int,
         [P(MassAction), c_real, c_real, c_real, P(c_real), P(c_real)])
    _sig(lib, "oracle_massaction_solve_per_cell", C.c_int,
         [P(MassAction), c_real, c_real, c_real, C.c_int, P(c_real), P(c_real), P(c_real)])
    _sig(lib, "oracle_tridiag_atimes", C.c_int, [c_void_p, c_void_p, c_void_p])
    _sig(lib, "oracle_tridiag_psetup", C.c_int, [c_void_p])
    _sig(lib, "oracle_tridiag_psolve", C.c_int, [c_void_p, c_void_p, c_void_p, c_real, C.c_int])
    _sig(lib, "oracle_adr_solve", C.c_int, [P(Adr), P(AdrOptions), P(c_real), P(c_real)])
    _sig(lib, "oracle_kin_solve", C.c_int, [P(KinOptions), P(c_real), P(c_real)])
    _sig(lib, "oracle_kin_newton_solve", C.c_int, [P(KinNewtonOptions), P(c_real), P(c_real)])
    _sig(lib, "oracle_kin_fertron_example", C.c_int, [P(c_real), P(c_real)])
    _sig(lib, "oracle_ark_dahlquist", C.c_int, [c_real, c_real, c_real, c_real, c_real, c_real,
                                                C.c_char_p, C.c_int, C.c_int, P(c_real),
                                                P(c_real)])
    _sig(lib, "oracle_ark_kpr", C.c_int, [C.c_int, P(c_real), P(c_real)])
    _sig(lib, "oracle_batch_solve", C.c_int,
         [P(Batch), c_real, c_real, c_real, C.c_int, P(c_real), P(C.c_long), P(c_real)])
    _sig(lib, "oracle_ark_dahlquist_predict", C.c_int,
         [c_real, c_real, c_real, c_real, c_real, C.c_int, C.c_int, P(c_real), P(c_real),
          P(C.c_long)])
    _sig(lib, "oracle_bench_vector_ops", C.c_int,
         [C.c_long, C.c_int, C.c_int, C.c_int, P(c_real)])
    return lib


def new_context():
    ctx = c_void_p()
    rc = core().SUNContext_Create(0, C.byref(ctx))
    if rc != 0:
        raise RuntimeError("SUNContext_Create failed: %d" % rc)
    return ctx


def free_context(ctx):
    core().SUNContext_Free(C.byref(ctx))
