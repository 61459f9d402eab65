// msm_lanepair.h -- lane-pair gathers of 32-byte records (gfx950 v_permlane32_swap).
//
// Lanes l and l + 32 of a wave form a pair.  To gather one 32-byte record for each lane, the two
// lanes load the two 16-byte halves of the SAME record in one instruction (lane l half 0, lane
// l + 32 half 1), once for the low lane's record and once for the high lane's, and
// v_permlane32_swap hands each lane the half it is missing.  Each gather instruction then
// touches 32 records instead of 64, halving the L1 tag lookups.  Both lanes of a pair must
// execute every call together.
#pragma once

#include "msm_device.h"

namespace msm {
namespace {

// On entry lo (hi) holds, in every lane, the half `lane >> 5` of the record of the low (high)
// lane of its pair; on exit lo and hi hold halves 0 and 1 of the lane's own record.
// permlane32_swap(vdst, vsrc) exchanges lanes 32..63 of vdst with lanes 0..31 of vsrc.
__device__ __forceinline__ void pair_swap(double& lo, double& hi) {
    const unsigned long long a = static_cast<unsigned long long>(__double_as_longlong(lo));
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(hi));
    const auto w0 = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(a),
                                                     static_cast<unsigned>(b), false, false);
    const auto w1 = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(a >> 32),
                                                     static_cast<unsigned>(b >> 32), false, false);
    lo = __longlong_as_double(static_cast<long long>(
        (static_cast<unsigned long long>(w1[0]) << 32) | w0[0]));
    hi = __longlong_as_double(static_cast<long long>(
        (static_cast<unsigned long long>(w1[1]) << 32) | w0[1]));
}

// Index of the low lane's and the high lane's value (both lanes receive both).
__device__ __forceinline__ void pair_both(int v, int& vlo, int& vhi) {
    const auto w = __builtin_amdgcn_permlane32_swap(static_cast<unsigned>(v),
                                                    static_cast<unsigned>(v), false, false);
    vlo = static_cast<int>(w[0]);
    vhi = static_cast<int>(w[1]);
}

// Record r (32 bytes) of a double4 array, gathered by the pair: returns halves 0 and 1 of the
// lane's own record.
__device__ __forceinline__ void pair_gather(const double4* base, int r, int sub,
                                            double2& h0, double2& h1) {
    int rlo, rhi;
    pair_both(r, rlo, rhi);
    const double2* p = reinterpret_cast<const double2*>(base);
    double2 a = p[2 * static_cast<size_t>(rlo) + sub];
    double2 b = p[2 * static_cast<size_t>(rhi) + sub];
    pair_swap(a.x, b.x);
    pair_swap(a.y, b.y);
    h0 = a;
    h1 = b;
}

}  // namespace
}  // namespace msm
