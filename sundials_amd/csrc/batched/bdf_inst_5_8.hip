// Batched BDF kernels for systems of 5, 6, 7, 8 species (FP64 and FP32 LU).
#include "bdf_device.h"

SUNAMD_BDF_INSTANTIATE(5)
SUNAMD_BDF_INSTANTIATE(6)
SUNAMD_BDF_INSTANTIATE(7)
SUNAMD_BDF_INSTANTIATE(8)
