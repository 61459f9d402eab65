// msm_thermal.hip -- Maxwell re-thermalisation of the velocities on the device
// (Thermalizing_MC_DEV, MDLIB/sor/CommonGPU/MD_DiffScheme_GPU.F90:1451-1805).
//
// The reference draws its uniform numbers from cuRAND's XORWOW states, one state per thread.
// Here every number comes from a counter-based generator, Philox4x32-10 (Salmon, Moraes, Dror and
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11): one 128-bit block per
// (atom, component), keyed by the caller's 64-bit seed, counter = (caller's atom index,
// component, 64-bit draw number).  The velocities therefore depend only on the seed, the draw
// number and the caller's atom numbering -- not on the cell-sorted order, the grid or the
// device.  The two uniforms of a component are (0, 1] doubles with 53 random bits each, as
// curand_uniform's (0, 1] range keeps -ln z finite.
//
// The transformation is the reference kernel's: for an active atom, each component whose FIXVEL
// and FIXPOS bits are clear becomes V0 sqrt(-ln z1) cos(2 pi z2), V0 = sqrt(2 T kB / m), the
// other components keep their value; an inactive atom gets zero velocity.  Then, box by box, the
// mass-weighted mean velocity of all the box's atoms is subtracted from every atom of the box.
#include "msm_device.h"

namespace msm {

namespace {

struct U4 {
    unsigned x, y, z, w;
};

// Philox4x32 with 10 rounds: multipliers 0xD2511F53, 0xCD9E8D57 and Weyl key increments
// 0x9E3779B9, 0xBB67AE85 of the published generator.
__device__ __forceinline__ U4 philox4x32_10(U4 c, unsigned k0, unsigned k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const unsigned long long p0 = static_cast<unsigned long long>(0xD2511F53u) * c.x;
        const unsigned long long p1 = static_cast<unsigned long long>(0xCD9E8D57u) * c.z;
        const unsigned hi0 = static_cast<unsigned>(p0 >> 32), lo0 = static_cast<unsigned>(p0);
        const unsigned hi1 = static_cast<unsigned>(p1 >> 32), lo1 = static_cast<unsigned>(p1);
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

// (0, 1] double from 53 bits of (lo, hi): (m + 1) 2^-53, m in [0, 2^53).
__device__ __forceinline__ double unit_open0(unsigned lo, unsigned hi) {
    const unsigned long long m =
        (static_cast<unsigned long long>(hi & 0x1FFFFFu) << 32) | static_cast<unsigned long long>(lo);
    return static_cast<double>(m + 1ull) * 0x1.0p-53;
}

__global__ void k_thermalize(AtomsDev a, TabDev t, const int* __restrict__ gid, double tkb,
                             unsigned long long seed, unsigned long long draw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int st = a.statu[i];
    if ((st & STATU_ACTIVE) != STATU_ACTIVE) {
        a.vx[i] = 0.0;
        a.vy[i] = 0.0;
        a.vz[i] = 0.0;
        return;
    }
    const double v0 = sqrt(2.0 * tkb / t.cm[a.ityp[i]]);
    const unsigned k0 = static_cast<unsigned>(seed), k1 = static_cast<unsigned>(seed >> 32);
    const unsigned g = static_cast<unsigned>(gid[i]);
    const unsigned d0 = static_cast<unsigned>(draw), d1 = static_cast<unsigned>(draw >> 32);
    double* v[3] = {a.vx, a.vy, a.vz};
    const int fix[3] = {STATU_FIXVELX | STATU_FIXPOSX, STATU_FIXVELY | STATU_FIXPOSY,
                        STATU_FIXVELZ | STATU_FIXPOSZ};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (st & fix[k]) continue;
        const U4 r = philox4x32_10(U4{g, static_cast<unsigned>(k), d0, d1}, k0, k1);
        const double z1 = unit_open0(r.x, r.y), z2 = unit_open0(r.z, r.w);
        v[k][i] = v0 * sqrt(-log(z1)) * cos(6.283185307179586 * z2);
    }
}

// One workgroup per box: sum of m v (3 components) and of m over the box's atoms -- its cells in
// sorted order, then the atoms of the out-of-box bucket that belong to it -- as fixed-order
// per-thread partials and a fixed tree.  out[4 b .. 4 b + 3] = {sum m vx, sum m vy, sum m vz, M}.
__global__ __launch_bounds__(kBlock) void k_box_momentum(AtomsDev a, TabDev t,
                                                         const int* __restrict__ start,
                                                         const BoxDev* __restrict__ boxes,
                                                         int ncell_total,
                                                         double* __restrict__ out) {
    __shared__ double ls[4][kBlock];
    const int b = blockIdx.x;
    const BoxDev B = boxes[b];
    const int c0 = B.cell_off, c1 = B.cell_off + B.ncell[0] * B.ncell[1] * B.ncell[2];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    auto add = [&](int i) {
        const double m = t.cm[a.ityp[i]];
        s[0] += m * a.vx[i];
        s[1] += m * a.vy[i];
        s[2] += m * a.vz[i];
        s[3] += m;
    };
    for (int i = start[c0] + threadIdx.x; i < start[c1]; i += kBlock) add(i);
    for (int i = start[ncell_total] + threadIdx.x; i < a.n; i += kBlock)
        if (a.boxid[i] == b) add(i);
#pragma unroll
    for (int k = 0; k < 4; ++k) ls[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int d = kBlock / 2; d > 0; d >>= 1) {
        if (threadIdx.x < d)
#pragma unroll
            for (int k = 0; k < 4; ++k) ls[k][threadIdx.x] += ls[k][threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) out[4 * b + k] = ls[k][0];
}

// v -= (sum m v) / M of the atom's box, every atom and component.
__global__ void k_sub_momentum(AtomsDev a, const double* __restrict__ mom) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double* p = mom + 4 * a.boxid[i];
    const double m = p[3];
    if (!(m > 0.0)) return;
    a.vx[i] = a.vx[i] - p[0] / m;
    a.vy[i] = a.vy[i] - p[1] / m;
    a.vz[i] = a.vz[i] - p[2] / m;
}

}  // namespace

void launch_thermalize(const AtomsDev& a, const TabDev& t, const int* gid, double tkb,
                       unsigned long long seed, unsigned long long draw, const int* start,
                       const BoxDev* boxes, int nbox, int ncell_total, double* mom,
                       hipStream_t s) {
    if (a.n <= 0 || nbox <= 0) return;
    hipLaunchKernelGGL(k_thermalize, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, t, gid, tkb, seed,
                       draw);
    hipLaunchKernelGGL(k_box_momentum, dim3(nbox), dim3(kBlock), 0, s, a, t, start, boxes,
                       ncell_total, mom);
    hipLaunchKernelGGL(k_sub_momentum, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, mom);
}

}  // namespace msm
