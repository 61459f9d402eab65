// msm_brick.h -- bricks of cells staged in LDS: geometry of a brick, its halo table, the walk
// over its halo slots and the map from home-atom number to cell (shared by the list build and
// the density/force passes).
//
// Atoms are cell-sorted, so each cell is a contiguous index range.  A brick is at most 4x4x4
// cells of one box; its halo is the brick plus one cell layer (periodic images resolved), at most
// 6x6x6 = 216 cells.
#pragma once

#include "msm_device.h"
#include "msm_internal.h"

namespace msm {
namespace {

struct BrickInfo {
    int c0[3], e[3], H[3];
    int nh, nhc;
};

__device__ __forceinline__ void brick_info(const BoxDev& B, int bi, BrickInfo& g) {
    int loc = bi - B.brick_off;
    const int kx = loc % B.nbrick[0];
    loc /= B.nbrick[0];
    const int ky = loc % B.nbrick[1];
    const int kz = loc / B.nbrick[1];
    g.c0[0] = kx * B.bdim[0];
    g.c0[1] = ky * B.bdim[1];
    g.c0[2] = kz * B.bdim[2];
    for (int k = 0; k < 3; ++k) {
        g.e[k] = min(B.bdim[k], B.ncell[k] - g.c0[k]);
        g.H[k] = g.e[k] + 2;
    }
    g.nh = g.H[0] * g.H[1] * g.H[2];
    g.nhc = g.e[0] * g.e[1] * g.e[2];
}

// Halo table of a brick (all threads of the workgroup call it): global cell of each halo cell
// (-1 outside a free boundary), image shift code (-1/0/+1 per direction, when hsh != nullptr),
// exclusive prefix of atom counts hoff[0..nh] and of home-cell counts homeoff[0..nhc].  Returns
// the number of halo atoms; LDS tables are valid on return.
__device__ __forceinline__ int brick_halo(const BoxDev& B, const BrickInfo& g, const int* __restrict__ start,
                          int* hcell, signed char* hsh, int* hoff, int* homeoff, int* scan_lds) {
    const int tid = threadIdx.x;
    int cnt = 0;
    if (tid < g.nh) {
        const int h3[3] = {tid % g.H[0], (tid / g.H[0]) % g.H[1], tid / (g.H[0] * g.H[1])};
        int g3[3];
        bool ok = true;
        for (int k = 0; k < 3; ++k) {
            int c = g.c0[k] - 1 + h3[k];
            signed char sh = 0;
            if (B.ifpd[k]) {
                if (c < 0) {
                    c += B.ncell[k];
                    sh = -1;
                } else if (c >= B.ncell[k]) {
                    c -= B.ncell[k];
                    sh = 1;
                }
            }
            ok = ok && c >= 0 && c < B.ncell[k];
            g3[k] = c;
            if (hsh) hsh[3 * tid + k] = sh;
        }
        const int gc = ok ? B.cell_off + cell_local(B, g3[0], g3[1], g3[2]) : -1;
        hcell[tid] = gc;
        cnt = ok ? start[gc + 1] - start[gc] : 0;
    }
    const int incl = block_incl_scan(cnt, scan_lds);
    if (tid < g.nh) hoff[tid + 1] = incl;
    if (tid == 0) hoff[0] = 0;
    __syncthreads();
    // prefix over the home cells (x fastest), one thread per home cell
    int hc = 0;
    if (tid < g.nhc) {
        const int ix = tid % g.e[0], iy = (tid / g.e[0]) % g.e[1], iz = tid / (g.e[0] * g.e[1]);
        const int h = (ix + 1) + g.H[0] * ((iy + 1) + g.H[1] * (iz + 1));
        hc = hoff[h + 1] - hoff[h];
    }
    const int hincl = block_incl_scan(hc, scan_lds);
    if (tid < g.nhc) homeoff[tid + 1] = hincl;
    if (tid == 0) homeoff[0] = 0;
    __syncthreads();
    return hoff[g.nh];
}

// Visit every halo slot s of a staged brick with its sorted atom index q and halo cell h.  Halo
// rows (fixed y, z) are contiguous runs of slots; waves take rows round-robin and lanes take
// consecutive slots, so the global loads of a row are coalesced, and a lane finds its cell by a
// short forward scan within the row (at most 6 cells).
template <typename F>
__device__ __forceinline__ void for_halo_slots(const BrickInfo& g, const int* __restrict__ start,
                                               const int* hcell, const int* hoff, F&& f) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nrows = g.H[1] * g.H[2];
    for (int r = wave; r < nrows; r += static_cast<int>(blockDim.x >> 6)) {
        const int h0 = r * g.H[0];
        const int s1 = hoff[h0 + g.H[0]];
        for (int s = hoff[h0] + lane; s < s1; s += 64) {
            int h = h0;
            while (hoff[h + 1] <= s) ++h;
            f(s, start[hcell[h]] + (s - hoff[h]), h);
        }
    }
}

// Home atom number t of the brick -> halo cell index h and offset within that cell.
__device__ __forceinline__ void home_atom(int t, const BrickInfo& g, const int* homeoff, int& h,
                                          int& off) {
    int lo = 0, hi = g.nhc;  // homeoff[lo] <= t < homeoff[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (homeoff[mid] <= t) lo = mid; else hi = mid;
    }
    const int ix = lo % g.e[0], iy = (lo / g.e[0]) % g.e[1], iz = lo / (g.e[0] * g.e[1]);
    h = (ix + 1) + g.H[0] * ((iy + 1) + g.H[1] * (iz + 1));
    off = t - homeoff[lo];
}

// Neighbours processed per group in the density and force passes.
constexpr int kUnroll = 8;

__device__ __forceinline__ double shift_of(signed char c, double len) {
    return c ? (c > 0 ? len : -len) : 0.0;
}

struct BrickLds {
    int hoff[kNlMaxHalo + 1];
    int hcell[kNlMaxHalo];
    int homeoff[65];
    int scan[kNlBlock / 64];
};

}  // namespace
}  // namespace msm
