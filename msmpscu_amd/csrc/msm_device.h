// msm_device.h -- device-side helpers shared by the kernel translation units.
#pragma once

#include "msm_internal.h"

namespace msm {

static constexpr int STATU_ACTIVE = 1;
static constexpr int STATU_FIXPOSX = 2, STATU_FIXPOSY = 4, STATU_FIXPOSZ = 8;
static constexpr int STATU_FIXVELX = 16, STATU_FIXVELY = 32, STATU_FIXVELZ = 64;
static constexpr int STATU_OUTOFBOX = 65536, STATU_REFLECT = 131072, STATU_TRANSMIT = 262144,
              STATU_PASSBOUND = 524288;

// Workgroup id -> logical block so that XCD x (receiving ids x, x+8, ...) gets the contiguous
// logical range [x*q + min(x,r), ...) of a bijection over [0, nblk).
static __device__ __forceinline__ int xcd_block(int nblk) {
    const int b = blockIdx.x;
    const int q = nblk >> 3, r = nblk & 7;
    const int x = b & 7, k = b >> 3;
    const int base = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    return base + k;
}

static __device__ __forceinline__ int atom_index() {
    return xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
}

static inline unsigned grid_for(int n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// Local id of cell (ix, iy, iz) in box b.  Cells are numbered in bricks of cbs^3 cells (x fastest
// inside a brick; bricks x fastest, then y, then z; the last brick of a direction is cut to the
// grid), so runs of consecutive cell ids -- and of cell-sorted atoms -- are compact 3-D blocks
// rather than rows.  cbs = 1 is the plain row-major order.  The numbering is a bijection onto
// [0, ncell[0]*ncell[1]*ncell[2]).
__host__ __device__ __forceinline__ int cell_local(const BoxDev& b, int ix, int iy, int iz) {
    const int bs = b.cbs, nx = b.ncell[0], ny = b.ncell[1], nz = b.ncell[2];
    const int bx = ix / bs, by = iy / bs, bz = iz / bs;
    const int lx = ix - bx * bs, ly = iy - by * bs, lz = iz - bz * bs;
    const int ex = min(bs, nx - bx * bs), ey = min(bs, ny - by * bs), ez = min(bs, nz - bz * bs);
    return nx * ny * (bz * bs) + nx * (by * bs) * ez + (bx * bs) * ey * ez + lx + ex * (ly + ey * lz);
}

// Inverse of cell_local.
__host__ __device__ __forceinline__ void cell_coords(const BoxDev& b, int lc, int& ix, int& iy,
                                                     int& iz) {
    const int bs = b.cbs, nx = b.ncell[0], ny = b.ncell[1], nz = b.ncell[2];
    const int bz = lc / (nx * ny * bs);
    int r = lc - bz * nx * ny * bs;
    const int ez = min(bs, nz - bz * bs);
    const int by = r / (nx * bs * ez);
    r -= by * nx * bs * ez;
    const int ey = min(bs, ny - by * bs);
    const int bx = r / (bs * ey * ez);
    r -= bx * bs * ey * ez;
    const int ex = min(bs, nx - bx * bs);
    ix = bx * bs + r % ex;
    iy = by * bs + (r / ex) % ey;
    iz = bz * bs + r / (ex * ey);
}

// Offset of slot 0 of atom p in a blocked neighbour list (slot iw adds iw * kListRow).
static __device__ __forceinline__ size_t list_base(int p, int mx) {
    return static_cast<size_t>(p / kListRow) * static_cast<size_t>(mx) * kListRow + (p % kListRow);
}

// Minimum image as the reference writes it: "if periodic and |s| > L/2 then s - sign(L, s)" --
// one box length at most, whatever |s| (quench trial points can lie many box lengths out).
// Branch-free: the shifted value is formed and selected (one rounding, as in the reference).
// ilen = 0 marks a free direction (no shift).
static __device__ __forceinline__ double min_image(double s, double len, double ilen) {
    const double hm = ilen > 0.0 ? 0.5 * len : __builtin_huge_val();
    const double t = s - __builtin_copysign(len, s);
    return __builtin_fabs(s) > hm ? t : s;
}

// 1/sqrt(x), x > 0: the hardware estimate y (relative error below 6e-8 on gfx950) refined by one
// second-order step of the binomial series of (1 - e)^(-1/2), e = 1 - x y^2:
//   y (1 + e/2 + 3e^2/8)
// whose truncation error 5e^3/16 is below 1e-21; the result is within one ulp over the pair
// ranges of the passes (two Newton steps, which cost two more operations, reach two).  Used in
// place of the reference's two correctly rounded square roots and two divisions per pair; the
// difference is at the 1e-16 relative level, far inside the 1e-10 eV/A force bar.
static __device__ __forceinline__ double rsqrt_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// F'(rho) of the Finnis-Sinclair embedding F = -sqrt(rho): -1/2 rho^(-1/2) for rho > 0 (the
// reference's -0.5/DSQRT(DEN)), from the refined reciprocal square root (within two ulp of the
// correctly rounded quotient).
static __device__ __forceinline__ double fs_fprime(double rho) {
    return -0.5 * rsqrt_nr(rho);
}

// Table coordinate of a pair: SK = sqrt(sqrt(r2)) * CSI (MD_FS_ForceTable.F90:248-252), with
// 1/r = rsqrt(r2) and sqrt(r) = rsqrt(1/r).  Returns 1/r through inv_r.
static __device__ __forceinline__ double table_coord(double r2, double csi, double& inv_r) {
    inv_r = rsqrt_nr(r2);
    return rsqrt_nr(inv_r) * csi;
}

template <int MB>
__device__ __forceinline__ const BoxDev& box_of(const BoxDev& box0, const BoxDev* boxes,
                                                const int* boxid, int i) {
    if constexpr (MB) return boxes[boxid[i]];
    return box0;
}

static __device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

static __device__ __forceinline__ int block_incl_scan(int v, int* lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = wave_incl_scan(v);
    if (lane == 63) lds[wid] = v;
    __syncthreads();
    if (wid == 0) {
        int w = (lane < (int)(blockDim.x >> 6)) ? lds[lane] : 0;
        w = wave_incl_scan(w);
        if (lane < (int)(blockDim.x >> 6)) lds[lane] = w;
    }
    __syncthreads();
    if (wid > 0) v += lds[wid - 1];
    __syncthreads();
    return v;
}

// Full-stencil order of MD_NeighborsList.F90:427-429 (entry 0 = own cell).
static __constant__ int c_nix[27] = {0, -1, -1, -1, 0, 0, -1, 1, -1, 0, 1, -1, 0, 1,
                              1, 1, 1, 0, 0, -1, 1, -1, 0, 1, -1, 0, 1};
static __constant__ int c_niy[27] = {0, 0, -1, 1, 1, 0, 0, 0, -1, -1, -1, 1, 1, 1,
                              0, 1, -1, -1, 0, 0, 0, -1, -1, -1, 1, 1, 1};
static __constant__ int c_niz[27] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                              0, 0, 0, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1};

// Velocity and accumulated displacement of one atom, read by a force pass before the atom's pair
// loop so that the integrator at its end does not wait on them (only the atom's own lane reads or
// writes them during the pass).
struct AtomState {
    double vx, vy, vz, dx, dy, dz;
};

__device__ __forceinline__ void load_atom_state(const AtomsDev& a, const StepParams& sp, int i,
                                                AtomState& s) {
    s.vx = a.vx[i];
    s.vy = a.vy[i];
    s.vz = a.vz[i];
    s.dx = s.dy = s.dz = 0.0;
    if (sp.track_dis) {
        s.dx = a.dx[i];
        s.dy = a.dy[i];
        s.dz = a.dz[i];
    }
}

// Corrector of this step (Correction_KERNEL, MD_DiffScheme_GPU.F90:686-758) and, for MODE 2, the
// predictor of the next step (Predictor_KERNEL0, :242-384) with the same forces, for atom i.
// Positions of MODE 2 go to a.pos_out (the next step's buffer).  pre: the atom's velocity and
// displacement as read by load_atom_state, or null to read them here.
template <int MODE>
__device__ __forceinline__ void integrate_atom(const AtomsDev& a, const TabDev& t, const BoxDev& b,
                                               const StepParams& sp, int i, int st, bool active,
                                               const double4& pi, int ti, double fpx, double fpy,
                                               double fpz, const AtomState* pre = nullptr) {
        if (active) {
#pragma clang fp contract(off)
            const double cm0 = t.cm[ti];
            const double ax = fpx / cm0, ay = fpy / cm0, az = fpz / cm0;
            double vx = pre ? pre->vx : a.vx[i], vy = pre ? pre->vy : a.vy[i],
                   vz = pre ? pre->vz : a.vz[i];
            // Correction_KERNEL
            if (!(st & (STATU_FIXVELX | STATU_FIXPOSX))) vx = vx + sp.hs2 * ax;
            if (!(st & (STATU_FIXVELY | STATU_FIXPOSY))) vy = vy + sp.hs2 * ay;
            if (!(st & (STATU_FIXVELZ | STATU_FIXPOSZ))) vz = vz + sp.hs2 * az;
            if constexpr (MODE == 2) {
                // Predictor_KERNEL0 of the next step with the same forces
                int stat = st;
                double x = pi.x, y = pi.y, z = pi.z;
                double dx = sp.h * vx + sp.h2s2 * ax;
                if (stat & STATU_FIXPOSX) dx = 0.0;
                x = x + dx;
                if (b.ifpd[0]) {
                    if (x > b.lo[0] + b.len[0]) { x = x - b.len[0]; stat |= STATU_PASSBOUND; }
                    else if (x < b.lo[0])       { x = x + b.len[0]; stat |= STATU_PASSBOUND; }
                }
                double dy = sp.h * vy + sp.h2s2 * ay;
                if (stat & STATU_FIXPOSY) dy = 0.0;
                y = y + dy;
                if (b.ifpd[1]) {
                    if (y > b.lo[1] + b.len[1]) { y = y - b.len[1]; stat |= STATU_PASSBOUND; }
                    else if (y < b.lo[1])       { y = y + b.len[1]; stat |= STATU_PASSBOUND; }
                }
                double dz = sp.h * vz + sp.h2s2 * az;
                if (stat & STATU_FIXPOSZ) dz = 0.0;
                z = z + dz;
                if (b.ifpd[2]) {
                    if (z > b.lo[2] + b.len[2]) { z = z - b.len[2]; stat |= STATU_PASSBOUND; }
                    else if (z < b.lo[2])       { z = z + b.len[2]; stat |= STATU_PASSBOUND; }
                }
                if (!(stat & (STATU_FIXVELX | STATU_FIXPOSX))) vx = vx + sp.hs2 * ax;
                if (!(stat & (STATU_FIXVELY | STATU_FIXPOSY))) vy = vy + sp.hs2 * ay;
                if (!(stat & (STATU_FIXVELZ | STATU_FIXPOSZ))) vz = vz + sp.hs2 * az;
                if (sp.track_dis) {
                    a.dx[i] = (pre ? pre->dx : a.dx[i]) + dx;
                    a.dy[i] = (pre ? pre->dy : a.dy[i]) + dy;
                    a.dz[i] = (pre ? pre->dz : a.dz[i]) + dz;
                }
                if (x > b.lo[0] + b.len[0] || y > b.lo[1] + b.len[1] || z > b.lo[2] + b.len[2])
                    stat = STATU_OUTOFBOX | STATU_REFLECT;
                else if (x < b.lo[0] || y < b.lo[1] || z < b.lo[2])
                    stat = STATU_OUTOFBOX | STATU_TRANSMIT;
                if (stat != st) a.statu[i] = stat;
                a.pos_out[i] = make_double4(x, y, z, 0.0);
            }
            a.vx[i] = vx;
            a.vy[i] = vy;
            a.vz[i] = vz;
        } else if constexpr (MODE == 2) {
            a.pos_out[i] = make_double4(pi.x, pi.y, pi.z, 0.0);
        }
    }

// Full list of one atom p (local cell lc of box b) read straight from global memory.
static __device__ void nlist_atom_global(int p, int lc, const BoxDev& b, const double4* __restrict__ pos,
                                  const int* __restrict__ ityp, const int* __restrict__ start,
                                  int ngroup, const double* __restrict__ rnl2, int mx,
                                  int* __restrict__ kvois, int* __restrict__ indi, int stride,
                                  int* err) {
#pragma clang fp contract(off)
    int ix, iy, iz;
    cell_coords(b, lc, ix, iy, iz);
    const double4 xi = pos[p];
    const int ti = ityp[p];
    int cnt = 0;
    for (int s = 0; s < 27; ++s) {
        int j3[3] = {ix + c_nix[s], iy + c_niy[s], iz + c_niz[s]};
        double cx[3] = {0.0, 0.0, 0.0};
        bool skip = false;
        for (int k = 0; k < 3; ++k) {
            if (b.ifpd[k]) {
                if (j3[k] >= b.ncell[k]) {
                    j3[k] = 0;
                    cx[k] = b.len[k];
                } else if (j3[k] < 0) {
                    j3[k] = b.ncell[k] - 1;
                    cx[k] = -b.len[k];
                }
            }
            skip = skip || j3[k] < 0 || j3[k] >= b.ncell[k];
        }
        if (skip) continue;
        const int g = b.cell_off + cell_local(b, j3[0], j3[1], j3[2]);
        const int qe = start[g + 1];
        for (int q = start[g]; q < qe; ++q) {
            if (q == p) continue;
            const double4 xj = pos[q];
            // XPT(I) - (XP(J) + CXYZ), squared and summed left to right, no contraction
            const double dx = xi.x - (xj.x + cx[0]);
            const double dy = xi.y - (xj.y + cx[1]);
            const double dz = xi.z - (xj.z + cx[2]);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < rnl2[ti * ngroup + ityp[q]]) {
                if (cnt < mx) indi[list_base(p, mx) + static_cast<size_t>(cnt) * stride] = q;
                ++cnt;
            }
        }
    }
    if (cnt > mx) {
        atomicOr(&err[kErrOverflow], 1);
        cnt = mx;
    }
    kvois[p] = cnt;
}

}  // namespace msm
