// msm_lanepair.h -- lane-pair gathers of 32-byte records (DPP swaps between adjacent lanes).
//
// Lanes 2m and 2m + 1 of a wave form a pair (sub = lane & 1).  To gather one 32-byte record for
// each lane, the two lanes load the two 16-byte halves of the SAME record in one instruction
// (the even lane half 0, the odd lane half 1), once for the even lane's record and once for the
// odd lane's, and a DPP quad_perm [1,0,3,2] move hands each lane the half it is missing.  The L1
// merges accesses of neighbouring lanes to one 128-byte line, so each gather instruction costs
// 32 line accesses instead of 64; lanes 32 apart are not merged, which is why the pair is two
// adjacent lanes.  Both lanes of a pair must execute every call together.
#pragma once

#include "msm_device.h"

namespace msm {
namespace {

// Value of the other lane of the pair (DPP quad_perm [1,0,3,2]; both lanes active).
__device__ __forceinline__ unsigned pair_other(unsigned v) {
    return static_cast<unsigned>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false));
}

__device__ __forceinline__ double pair_other(double v) {
    const unsigned long long a = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = pair_other(static_cast<unsigned>(a));
    const unsigned hi = pair_other(static_cast<unsigned>(a >> 32));
    return __longlong_as_double(
        static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// On entry lo (hi) holds, in every lane, the half `sub` of the record of the even (odd) lane of
// its pair; on exit lo and hi hold halves 0 and 1 of the lane's own record.  The even lane sends
// its hi (half 0 of the odd lane's record) and receives the odd lane's lo (half 1 of its own).
__device__ __forceinline__ void pair_swap(double& lo, double& hi, int sub) {
    const double r = pair_other(sub ? lo : hi);
    if (sub)
        lo = r;
    else
        hi = r;
}

// Value of the even lane and of the odd lane (both lanes receive both).
__device__ __forceinline__ void pair_both(int v, int sub, int& vlo, int& vhi) {
    const int o = static_cast<int>(pair_other(static_cast<unsigned>(v)));
    vlo = sub ? o : v;
    vhi = sub ? v : o;
}

// Record r (32 bytes) of a double4 array, gathered by the pair: returns halves 0 and 1 of the
// lane's own record.
__device__ __forceinline__ void pair_gather(const double4* base, int r, int sub,
                                            double2& h0, double2& h1) {
    int rlo, rhi;
    pair_both(r, sub, rlo, rhi);
    const double2* p = reinterpret_cast<const double2*>(base);
    double2 a = p[2 * static_cast<size_t>(rlo) + sub];
    double2 b = p[2 * static_cast<size_t>(rhi) + sub];
    pair_swap(a.x, b.x, sub);
    pair_swap(a.y, b.y, sub);
    h0 = a;
    h1 = b;
}

}  // namespace
}  // namespace msm
