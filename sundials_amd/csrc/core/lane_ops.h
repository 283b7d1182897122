// -----------------------------------------------------------------------------
// lane_ops.h -- cross-lane primitives of the register-row LU kernels (the
// block solver's and the batched integrator's): DPP row permutations, lane
// reads, and the pivot search of a group of G lanes as a (value, row)
// maximum with ties to the smaller row.
// -----------------------------------------------------------------------------
#ifndef SUNAMD_LANE_OPS_H
#define SUNAMD_LANE_OPS_H

#include <hip/hip_runtime.h>

namespace sunamd_lane
{

// DPP row permutations used by the group reductions
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x)
{
  return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_v(double x)
{
  const long long u = __double_as_longlong(x);
  const int lo      = dpp_i<CTRL>((int)(u & 0xffffffffLL));
  const int hi      = dpp_i<CTRL>((int)(u >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

template <int CTRL>
__device__ __forceinline__ float dpp_v(float x)
{
  return __int_as_float(dpp_i<CTRL>(__float_as_int(x)));
}

// lane l's value of x, for a lane index that is the same in the whole wavefront
__device__ __forceinline__ double readlane_d(double x, int l)
{
  const long long v = __double_as_longlong(x);
  const int lo      = __builtin_amdgcn_readlane((int)(v & 0xffffffffLL), l);
  const int hi      = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

__device__ __forceinline__ float readlane_d(float x, int l)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// ---- pivot search: largest |a_ik| over the group, ties to the smallest row
// (the dense GETRF's first maximum), as a reduction on (value, row) pairs:
// DPP exchanges inside rows of 16 lanes, lane reads across rows --------------

// (ov, op) replaces (v, p) when ov is larger, or equal with a smaller row;
// evaluated without short-circuit so it compiles to compares and selects
template <typename V>
__device__ __forceinline__ void amax_take(V& v, int& p, V ov, int op)
{
  const bool take = (ov > v) | ((ov == v) & (op < p));
  v               = take ? ov : v;
  p               = take ? op : p;
}

template <int CTRL, typename V>
__device__ __forceinline__ void amax_dpp(V& v, int& p)
{
  const V ov   = dpp_v<CTRL>(v);
  const int op = dpp_i<CTRL>(p);
  amax_take(v, p, ov, op);
}

// every lane of a G-lane group (G a power of two, groups aligned to G lanes)
// ends with the group's maximum pair
template <int G, typename V>
__device__ __forceinline__ void g_argmax(V& v, int& p)
{
  if constexpr (G >= 2) { amax_dpp<0xB1>(v, p); }  // quad_perm [1,0,3,2]
  if constexpr (G >= 4) { amax_dpp<0x4E>(v, p); }  // quad_perm [2,3,0,1]
  if constexpr (G >= 8) { amax_dpp<0x141>(v, p); } // row_half_mirror
  if constexpr (G >= 16) { amax_dpp<0x140>(v, p); } // row_mirror
  if constexpr (G >= 32)
  {
    const V r0 = readlane_d(v, 0), r1 = readlane_d(v, 16);
    const V r2 = readlane_d(v, 32), r3 = readlane_d(v, 48);
    const int p0 = __builtin_amdgcn_readlane(p, 0), p1 = __builtin_amdgcn_readlane(p, 16);
    const int p2 = __builtin_amdgcn_readlane(p, 32), p3 = __builtin_amdgcn_readlane(p, 48);
    V a = r0, b = r2;
    int pa = p0, pb = p2;
    amax_take(a, pa, r1, p1);
    amax_take(b, pb, r3, p3);
    if constexpr (G == 64) { amax_take(a, pa, b, pb); }
    const bool hi_half = (G == 32) && (__lane_id() >= 32);
    v = hi_half ? b : a;
    p = hi_half ? pb : pa;
  }
}

// wavefront-level ordering point for a group's LDS exchange
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

} // namespace sunamd_lane

#endif
