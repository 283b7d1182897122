/* -----------------------------------------------------------------------------
 * tridiag_mi355.h -- the Krylov test problem of the reference's SPGMR unit
 * test (test/unit_tests/sunlinsol/spgmr/serial/test_sunlinsol_spgmr_serial.c
 * :433-503): A = S1^{-1} Ahat S2 with Ahat = tridiag(-1, 5, -1), and a
 * Jacobi preconditioner z = r / d, as device callbacks for SPGMR on the
 * MI355X vector.  Every element is formed with the test's own operation
 * order, so the products equal the serial ones bit for bit.
 * ---------------------------------------------------------------------------*/
#ifndef SUNAMD_TRIDIAG_MI355_H
#define SUNAMD_TRIDIAG_MI355_H

#include "sundials_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  sunindextype N;
  N_Vector d;  /* preconditioner diagonal */
  N_Vector s1; /* row scaling */
  N_Vector s2; /* column scaling */
} SUNAMD_TridiagData;

/* SUNATimesFn / SUNPSetupFn / SUNPSolveFn on MI355X vectors */
int SUNAMD_TridiagATimes_Mi355(void* data, N_Vector v, N_Vector z);
int SUNAMD_TridiagPSetup_Mi355(void* data);
int SUNAMD_TridiagPSolve_Mi355(void* data, N_Vector r, N_Vector z, sunrealtype tol, int lr);

#ifdef __cplusplus
}
#endif

#endif
