// Batched BDF kernels for systems of 10, 12 species (FP64 and FP32 LU).
#include "bdf_device.h"

SUNAMD_BDF_INSTANTIATE(10)
SUNAMD_BDF_INSTANTIATE(12)
