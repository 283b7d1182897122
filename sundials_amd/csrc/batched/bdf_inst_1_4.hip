// Batched BDF kernels for systems of 1, 2, 3, 4 species (FP64 and FP32 LU).
#include "bdf_device.h"

SUNAMD_BDF_INSTANTIATE(1)
SUNAMD_BDF_INSTANTIATE(2)
SUNAMD_BDF_INSTANTIATE(3)
SUNAMD_BDF_INSTANTIATE(4)
