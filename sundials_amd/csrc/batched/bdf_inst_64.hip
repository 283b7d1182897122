// Batched BDF kernels for systems of 64 species (FP64 and FP32 LU).
#include "bdf_device.h"

SUNAMD_BDF_INSTANTIATE(64)
