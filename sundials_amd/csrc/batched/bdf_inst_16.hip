// Batched BDF kernels for systems of 16 species (FP64 and FP32 LU).
#include "bdf_device.h"

SUNAMD_BDF_INSTANTIATE(16)
