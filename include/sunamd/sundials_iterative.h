/* -----------------------------------------------------------------------------
 * sundials_iterative.h -- Gram-Schmidt and Givens QR routines of the Krylov
 * solvers, with the reference's signatures:
 *   SUNModifiedGS   include/sundials/sundials_iterative.h:207-208
 *   SUNClassicalGS  include/sundials/sundials_iterative.h:231-233
 *   SUNQRfact       include/sundials/sundials_iterative.h:268
 *   SUNQRsol        include/sundials/sundials_iterative.h:303
 *   SUNQRAddFn, SUNQRAdd_MGS / _ICWY / _CGS2 / _DCGS2
 *                   include/sundials/sundials_iterative.h:167-168, 334-458
 *   SUNQRData       src/sundials/sundials_iterative_impl.h:34-41
 * (the QR updates of an Anderson-acceleration space: append column df to the
 * thin QR factorisation Q R of the m columns before it; R is column-major
 * with leading dimension mMax)
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

typedef struct _SUNQRData* SUNQRData;
struct _SUNQRData
{
  N_Vector vtemp;
  N_Vector vtemp2;
  sunrealtype* temp_array;
};
typedef int (*SUNQRAddFn)(N_Vector* Q, sunrealtype* R, N_Vector f, int m, int mMax,
                          void* QR_data);
SUNErrCode SUNQRAdd_MGS(N_Vector* Q, sunrealtype* R, N_Vector df, int m, int mMax, void* QRdata);
SUNErrCode SUNQRAdd_ICWY(N_Vector* Q, sunrealtype* R, N_Vector df, int m, int mMax,
                         void* QRdata);
SUNErrCode SUNQRAdd_CGS2(N_Vector* Q, sunrealtype* R, N_Vector df, int m, int mMax,
                         void* QRdata);
SUNErrCode SUNQRAdd_DCGS2(N_Vector* Q, sunrealtype* R, N_Vector df, int m, int mMax,
                          void* QRdata);

#ifdef __cplusplus
}
#endif

#endif
