// -----------------------------------------------------------------------------
// Interface between the coupled mass-action driver and the device-stepped
// coupled integrator (coupled_device.hip).
// -----------------------------------------------------------------------------
#ifndef SUNAMD_COUPLED_DEVICE_H
#define SUNAMD_COUPLED_DEVICE_H

#include <hip/hip_runtime.h>

#include "sunamd/massaction_mi355.h"

// One integration of a coupled system of ncells cells of m species from t = 0
// to tout, all buffers on the device.
struct CoupledDeviceJob
{
  int m, R;
  long ncells;
  const int* idx;      // 4R reaction indices
  const double* kf;    // rate constants, cell-major (cell * R + r)
  const double* kb;
  const double* y0;    // ncells * m
  double* yout;        // ncells * m
  void* lu;            // ncells * m * m entries of the factor precision (scratch)
  int* piv;            // ncells * m (scratch)
  double rtol, atol, tout;
  long mxstep;
  int lu_precision;    // SUNAMD_LU_FP64 / SUNAMD_LU_FP32
  int sequential;      // WRMS norms in the serial vector's order
  hipStream_t stream;
};

// Runs the job; *state keeps the integrator's work buffers between runs of
// the same problem (NULL on the first call).  Returns the CVode flag of the
// integration, or -100 on a HIP error / unsupported size / lost grid barrier.
int sunamd_coupled_device_run(const CoupledDeviceJob& job, void** state, SUNAMD_SolveStats* st);
void sunamd_coupled_device_free(void* state);

#endif
