// Sums over small lane groups of a wavefront with DPP (data-parallel
// primitives: cross-lane moves inside the VALU), every lane of the group
// receiving the total.  Groups are aligned runs of G = 4, 8 or 16 lanes; the
// steps are lane permutations within a row of 16 lanes:
//   quad_perm [1,0,3,2]  (lane ^ 1)
//   quad_perm [2,3,0,1]  (lane ^ 2)
//   row_half_mirror      (lane i <-> 7 - i in each half row: the other quad)
//   row_mirror           (lane i <-> 15 - i in the row: the other half row)
// After each step every lane holds the sum of its aligned run of 2, 4, 8, 16
// lanes, so the result is the same on every lane of the group.  All lanes of
// the group must be active.
#pragma once

#include <hip/hip_runtime.h>

namespace toeddpp {

template <int CTRL> __device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int G, typename T> __device__ __forceinline__ T group_sum(T v) {
    static_assert(G == 4 || G == 8 || G == 16, "groups of 4, 8 or 16 lanes");
    v += dpp<0xB1>(v);                       // quad_perm [1,0,3,2]
    v += dpp<0x4E>(v);                       // quad_perm [2,3,0,1]
    if (G >= 8) v += dpp<0x141>(v);          // row_half_mirror
    if (G >= 16) v += dpp<0x140>(v);         // row_mirror
    return v;
}

}  // namespace toeddpp
