// Internal hooks used by kernel launch sites to time a kernel family with
// HIP events on the launching stream (no-ops unless profiling is enabled).
#ifndef SUNAMD_PROFILER_INTERNAL_H
#define SUNAMD_PROFILER_INTERNAL_H

#include <hip/hip_runtime.h>

void sunamd_prof_begin(int family, hipStream_t s);
void sunamd_prof_end(int family, hipStream_t s);

#endif
