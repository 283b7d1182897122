/* -----------------------------------------------------------------------------
 * sunlinsol_spgmr.h -- scaled preconditioned GMRES, the reference's
 * SUNLinSol_SPGMR (include/sunlinsol/sunlinsol_spgmr.h:44-46 defaults,
 * :88-95 constructor and setters), built on any N_Vector.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_SUNLINSOL_SPGMR_H
#define SUNAMD_SUNLINSOL_SPGMR_H

#include "sundials_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SUNSPGMR_MAXL_DEFAULT   5
#define SUNSPGMR_MAXRS_DEFAULT  0
#define SUNSPGMR_GSTYPE_DEFAULT SUN_MODIFIED_GS

/* sunlinsol_spgmr.h:88-95 */
SUNLinearSolver SUNLinSol_SPGMR(N_Vector y, int pretype, int maxl, SUNContext sunctx);
SUNErrCode SUNLinSol_SPGMRSetPrecType(SUNLinearSolver S, int pretype);
SUNErrCode SUNLinSol_SPGMRSetGSType(SUNLinearSolver S, int gstype);
SUNErrCode SUNLinSol_SPGMRSetMaxRestarts(SUNLinearSolver S, int maxrs);

/* Gram-Schmidt routine with SUNClassicalGS's signature
   (sundials_iterative.h:231-233) */
typedef SUNErrCode (*SUNAMD_ClassicalGSFn)(N_Vector* v, sunrealtype** h, int k, int p,
                                           sunrealtype* new_vk_norm, sunrealtype* stemp,
                                           N_Vector* vtemp);
/* no reference counterpart: installs the routine the solver uses when its
   Gram-Schmidt type is SUN_CLASSICAL_GS (NULL restores SUNClassicalGS) --
   how a device vector supplies a fused classical Gram-Schmidt step */
SUNErrCode SUNLinSol_SPGMRSetClassicalGSFn(SUNLinearSolver S, SUNAMD_ClassicalGSFn fn);

#ifdef __cplusplus
}
#endif

#endif
