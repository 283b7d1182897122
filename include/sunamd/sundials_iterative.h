/* -----------------------------------------------------------------------------
 * sundials_iterative.h -- Gram-Schmidt and Givens QR routines of the Krylov
 * solvers, with the reference's signatures:
 *   SUNModifiedGS   include/sundials/sundials_iterative.h:207-208
 *   SUNClassicalGS  include/sundials/sundials_iterative.h:231-233
 *   SUNQRfact       include/sundials/sundials_iterative.h:268
 *   SUNQRsol        include/sundials/sundials_iterative.h:303
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_SUNDIALS_ITERATIVE_H
#define SUNAMD_SUNDIALS_ITERATIVE_H

#include "sundials_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

SUNErrCode SUNModifiedGS(N_Vector* v, sunrealtype** h, int k, int p, sunrealtype* new_vk_norm);
SUNErrCode SUNClassicalGS(N_Vector* v, sunrealtype** h, int k, int p, sunrealtype* new_vk_norm,
                          sunrealtype* stemp, N_Vector* vtemp);
int SUNQRfact(int n, sunrealtype** h, sunrealtype* q, int job);
int SUNQRsol(int n, sunrealtype** h, sunrealtype* q, sunrealtype* b);

#ifdef __cplusplus
}
#endif

#endif
