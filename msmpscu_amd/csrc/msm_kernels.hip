// msm_kernels.hip -- gfx950 kernels of the tabulated FS/EAM force path (FP64).
//
// One thread per atom over cell-sorted atoms; workgroups of 256 (four wave64s).  Workgroup ids
// are remapped so that each of the eight XCDs (which receive workgroups round-robin) walks one
// contiguous eighth of the sorted atoms: the neighbour gathers of an XCD then stay within one
// spatial slab and hit that XCD's own L2.
//
// Reference kernels restated (semantics, not code):
//   cell ids         MDLIB/sor/CommonGPU/MD_NeighborsList_GPU.F90:759-819 with the host formula
//                    of MDLIB/sor/Common/MD_NeighborsList.F90:255-276 (FP64, eps = 1e-4)
//   list             MD_NeighborsList.F90:396-633 (27-cell stencil :427-429, strict <, shifted
//                    images x_j + L) -- the host predicate, exactly, rather than the FP32 device one
//   density          MD_FS_ForceTable_GPU.F90:341-516, MD_EAM_ForceTable_GPU.F90:369-548
//   force            MD_FS_ForceTable_GPU.F90:623-806, MD_EAM_ForceTable_GPU.F90:646-829
//   energy           MD_FS_ForceTable_GPU.F90:1451-1609, MD_EAM_ForceTable_GPU.F90:1470-1636
//   integrator       MD_DiffScheme_GPU.F90:242-384 (Predictor_KERNEL0), :686-758 (Correction_KERNEL)
#include <algorithm>
#include <climits>

#include "msm_internal.h"
#include "msm_device.h"

namespace msm {


// --------------------------------------------------------------------------------------------
// utilities
// --------------------------------------------------------------------------------------------

__global__ void k_fill_int(int* p, int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_int(int* p, int v, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n)), dim3(kBlock), 0, s, p, v, n);
}

// --------------------------------------------------------------------------------------------
// neighbour-list build: cell ids, counting sort, deterministic in-cell order, gather, list
// --------------------------------------------------------------------------------------------

__global__ void k_cell_assign(const double4* __restrict__ pos, const int* __restrict__ statu,
                              const int* __restrict__ boxid, const BoxDev* __restrict__ boxes,
                              int n, int ncell_total, int* __restrict__ key,
                              int* __restrict__ slot, int* __restrict__ count, int* err) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = ncell_total;  // sentinel bucket: atoms outside every cell grid
    if ((statu[i] & STATU_OUTOFBOX) != STATU_OUTOFBOX) {
        const BoxDev& b = boxes[boxid[i]];
        const double4 p = pos[i];
        const double xyz[3] = {p.x, p.y, p.z};
        int ic[3];
        bool in = true;
        for (int k = 0; k < 3; ++k) {
            // (x - LB)/L * NC - eps, truncated toward zero, as MD_NeighborsList.F90:258
            const double s = (xyz[k] - b.lo[k]) / b.len[k] * static_cast<double>(b.ncell[k]) - 1.0e-4;
            ic[k] = static_cast<int>(s);
            in = in && ic[k] >= 0 && ic[k] < b.ncell[k];
        }
        if (in)
            c = b.cell_off + cell_local(b, ic[0], ic[1], ic[2]);
        else
            atomicOr(&err[kErrOutOfBox], 1);
    }
    key[i] = c;
    // one atomic per distinct cell of the wave (after the first build the atoms are cell-sorted,
    // so a wave spans a few cells): the lanes of each cell are grouped by ballots first, then
    // the leader of every cell adds the cell's lane count, all leaders in one atomic
    // instruction (one round trip per wave), and every lane takes base + its rank among its
    // cell's lanes.  Slot order inside a cell is made deterministic afterwards by k_cell_order.
    const int lane = threadIdx.x & 63;
    unsigned long long todo = __ballot(1);
    int lead = lane, rank = 0, cnt = 0;
    while (todo) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const int lc = __shfl(c, leader, 64);
        const unsigned long long same = __ballot(c == lc) & todo;
        if (c == lc) {
            lead = leader;
            rank = __popcll(same & ((1ull << lane) - 1ull));
            if (lane == leader) cnt = __popcll(same);
        }
        todo &= ~same;
    }
    int base = 0;
    if (cnt > 0) base = atomicAdd(&count[c], cnt);
    slot[i] = __shfl(base, lead, 64) + rank;
}

void launch_cell_assign(const double4* pos, const int* statu, const int* boxid,
                        const BoxDev* boxes, int n, int ncell_total, SortBufs sb, int* err,
                        hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_assign, dim3(grid_for(n)), dim3(kBlock), 0, s, pos, statu, boxid,
                       boxes, n, ncell_total, sb.key, sb.slot, sb.count, err);
}

// Exclusive scan of int: block-level inclusive scans (wave64 shuffles + LDS), one-block scan of
// the block totals, then a fix-up pass.
constexpr int kScanBlock = 1024;

__global__ void k_scan_blocks(const int* __restrict__ in, int* __restrict__ out,
                              int* __restrict__ sums, int n) {
    __shared__ int lds[kScanBlock / 64];
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    const int v = (i < n) ? in[i] : 0;
    const int inc = block_incl_scan(v, lds);
    if (i < n) out[i] = inc - v;  // exclusive within the block
    if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = inc;
}

__global__ void k_scan_sums(int* sums, int nb) {
    __shared__ int lds[kScanBlock / 64];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += kScanBlock) {
        const int i = base + threadIdx.x;
        const int v = (i < nb) ? sums[i] : 0;
        const int inc = block_incl_scan(v, lds);
        const int c = carry;
        if (i < nb) sums[i] = c + inc - v;
        __syncthreads();
        if (threadIdx.x == kScanBlock - 1) carry = c + inc;
        __syncthreads();
    }
}

__global__ void k_scan_add(int* out, const int* __restrict__ sums, int n) {
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    if (i < n) out[i] += sums[blockIdx.x];
}

void launch_scan_exclusive(const int* in, int* out, int n, int* tmp, hipStream_t s) {
    if (n <= 0) return;
    const int nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanBlock), 0, s, in, out, tmp, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, s, tmp, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, s, out, tmp, n);
}

__global__ void k_scatter(int n, const int* __restrict__ key, const int* __restrict__ slot,
                          const int* __restrict__ start, int* __restrict__ perm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[start[key[i]] + slot[i]] = i;
}

// One thread per cell: order the cell's atoms by caller index, which makes the sorted order,
// the list order and every summation order independent of atomic arrival order.  A cell of at
// most kCellSortReg atoms is loaded in two independent rounds (entries, then their caller
// indices) and sorted in registers by an odd-even transposition network; a larger cell takes an
// insertion sort in memory.
constexpr int kCellSortReg = 16;

__global__ void k_cell_order(int ncells, const int* __restrict__ start, int* __restrict__ perm,
                             const int* __restrict__ gid) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const int b = start[c], e = start[c + 1];
    const int n = e - b;
    if (n <= 1) return;
    if (n <= kCellSortReg) {
        int v[kCellSortReg], g[kCellSortReg];
#pragma unroll
        for (int k = 0; k < kCellSortReg; ++k) v[k] = k < n ? perm[b + k] : 0;
#pragma unroll
        for (int k = 0; k < kCellSortReg; ++k) g[k] = k < n ? gid[v[k]] : INT_MAX;
        auto cswap = [&](int k) {
            const bool sw = g[k] > g[k + 1];
            const int g0 = sw ? g[k + 1] : g[k], g1 = sw ? g[k] : g[k + 1];
            const int v0 = sw ? v[k + 1] : v[k], v1 = sw ? v[k] : v[k + 1];
            g[k] = g0;
            g[k + 1] = g1;
            v[k] = v0;
            v[k + 1] = v1;
        };
#pragma unroll
        for (int r = 0; r < kCellSortReg / 2; ++r) {  // even phase, then odd phase
#pragma unroll
            for (int k = 0; k + 1 < kCellSortReg; k += 2) cswap(k);
#pragma unroll
            for (int k = 1; k + 1 < kCellSortReg; k += 2) cswap(k);
        }
#pragma unroll
        for (int k = 0; k < kCellSortReg; ++k)
            if (k < n) perm[b + k] = v[k];
        return;
    }
    for (int a = b + 1; a < e; ++a) {
        const int v = perm[a];
        const int g = gid[v];
        int q = a - 1;
        while (q >= b && gid[perm[q]] > g) {
            perm[q + 1] = perm[q];
            --q;
        }
        perm[q + 1] = v;
    }
}

void launch_scatter_sort(int n, int ncell_total, SortBufs sb, const int* gid, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(kBlock), 0, s, n, sb.key, sb.slot,
                       sb.start, sb.perm);
    const int nc = ncell_total + 1;
    hipLaunchKernelGGL(k_cell_order, dim3(grid_for(nc)), dim3(kBlock), 0, s, nc, sb.start,
                       sb.perm, gid);
}

// F = 0: forces are not carried (a force pass overwrites them before they are read).
template <int F>
__global__ void k_gather(int n, const int* __restrict__ perm, AtomsDev src,
                         const int* __restrict__ gid_src, AtomsDev dst, int* __restrict__ gid_dst,
                         const int* __restrict__ key, int* __restrict__ cell_dst) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = perm[p];
    double4 x = src.pos[i];
    x.w = 0.0;
    dst.pos[p] = x;
    dst.vx[p] = src.vx[i];
    dst.vy[p] = src.vy[i];
    dst.vz[p] = src.vz[i];
    if constexpr (F) {
        dst.fx[p] = src.fx[i];
        dst.fy[p] = src.fy[i];
        dst.fz[p] = src.fz[i];
    }
    dst.dx[p] = src.dx[i];
    dst.dy[p] = src.dy[i];
    dst.dz[p] = src.dz[i];
    dst.ityp[p] = src.ityp[i];
    dst.statu[p] = src.statu[i];
    dst.boxid[p] = src.boxid[i];
    gid_dst[p] = gid_src[i];
    cell_dst[p] = key[i];
}

void launch_gather(int n, int n_alloc, const int* perm, const AtomsDev& src, const int* gid_src,
                   AtomsDev& dst, int* gid_dst, const int* key, int* cell_dst, int keep_f,
                   hipStream_t s) {
    (void)n_alloc;
    if (n <= 0) return;
    if (keep_f)
        hipLaunchKernelGGL(k_gather<1>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, perm, src,
                           gid_src, dst, gid_dst, key, cell_dst);
    else
        hipLaunchKernelGGL(k_gather<0>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, perm, src,
                           gid_src, dst, gid_dst, key, cell_dst);
}

__global__ __launch_bounds__(kBlock) void k_nlist(
    int n, const double4* __restrict__ pos, const int* __restrict__ ityp,
    const int* __restrict__ statu, const int* __restrict__ boxid, const int* __restrict__ cell,
    const int* __restrict__ start, const BoxDev* __restrict__ boxes, int ngroup,
    const double* __restrict__ rnl2, int mx, int* __restrict__ kvois, int* __restrict__ indi,
    int stride, int* err) {
    const int p = atom_index();
    if (p >= n) return;
    if ((statu[p] & STATU_OUTOFBOX) == STATU_OUTOFBOX) {
        kvois[p] = 0;
        return;
    }
    const BoxDev& b = boxes[boxid[p]];
    nlist_atom_global(p, cell[p] - b.cell_off, b, pos, ityp, start, ngroup, rnl2, mx, kvois, indi,
                      stride, err);
}

void launch_nlist(int n, const double4* pos, const int* ityp, const int* statu, const int* boxid,
                  const int* cell, const int* start, const BoxDev* boxes, int ngroup,
                  const double* rnl2, int mx, int* kvois, int* indi, int stride, int* err,
                  hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_nlist, dim3(grid_for(n)), dim3(kBlock), 0, s, n, pos, ityp, statu, boxid,
                       cell, start, boxes, ngroup, rnl2, mx, kvois, indi, stride, err);
}

// --------------------------------------------------------------------------------------------
// density pass: rho_i = sum_j POTB(r_ij); F'(rho_i) stored in pos[i].w
// --------------------------------------------------------------------------------------------

template <int MB, int MT, int EAM>
__global__ __launch_bounds__(kBlock) void k_density(AtomsDev a, ListDev l, TabDev t, BoxDev box0,
                                                    const BoxDev* __restrict__ boxes, int* err) {
    const int i = atom_index();
    if (i >= a.n) return;
    double den0 = 0.0;
    if ((a.statu[i] & STATU_ACTIVE) == STATU_ACTIVE) {
        const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
        const double2 pxy = *reinterpret_cast<const double2*>(&a.pos[i]);
        const double pz = a.pos[i].z;
        const int ti = MT ? a.ityp[i] : 0;
        const int kv = l.kvois[i];
        const int* __restrict__ li = l.indi + list_base(i, l.mx);
        int jn = kv > 0 ? li[0] : 0;
        for (int iw = 0; iw < kv; ++iw) {
            const int j = jn;
            if (iw + 1 < kv) jn = li[static_cast<size_t>(iw + 1) * l.stride];
            const double2 qxy = *reinterpret_cast<const double2*>(&a.pos[j]);
            const double qz = a.pos[j].z;
            const double sx = min_image(pxy.x - qxy.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pxy.y - qxy.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pz - qz, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            if (r2 <= t.ru2max) {
                const int kind = MT ? t.kpair[ti * t.ngroup + a.ityp[j]] : t.kpair[0];
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), t.ntab2 - 1);
                const double2 e = t.b2[kind * t.ntab2 + kk];
                den0 = den0 + (e.x + (sk - kk) * e.y);
            }
        }
        if constexpr (!EAM) {
            if (den0 > 0.0) den0 = fs_fprime(den0);
        } else {
            if (den0 > 0.0) {
                const double sk = den0 / t.rhod + 1.0;
                int kk = static_cast<int>(sk + 0.000001);
                if (kk > t.nembd) {
                    atomicOr(&err[kErrEmbd], 1);
                    kk = t.nembd;
                }
                const double4 m = t.m4[t.kembd[ti] * t.nembd2 + kk];
                den0 = m.z + (sk - kk) * m.w;
            }
        }
    }
    a.pos[i].w = den0;
}

void launch_density(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                    const BoxDev* boxes, int multibox, int* err, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(grid_for(a.n)), b(kBlock);
    const bool mt = t.ngroup > 1, eam = t.pot_type != 0;
#define MSM_DEN(MB, MT, E) \
    hipLaunchKernelGGL((k_density<MB, MT, E>), g, b, 0, s, a, l, t, box0, boxes, err)
    if (multibox) {
        if (mt) { if (eam) MSM_DEN(1, 1, 1); else MSM_DEN(1, 1, 0); }
        else    { if (eam) MSM_DEN(1, 0, 1); else MSM_DEN(1, 0, 0); }
    } else {
        if (mt) { if (eam) MSM_DEN(0, 1, 1); else MSM_DEN(0, 1, 0); }
        else    { if (eam) MSM_DEN(0, 0, 1); else MSM_DEN(0, 0, 0); }
    }
#undef MSM_DEN
}

// --------------------------------------------------------------------------------------------
// force pass (+ optional fused corrector of this step and predictor of the next step)
// --------------------------------------------------------------------------------------------

template <int MB, int MT, int MODE>
__global__ __launch_bounds__(kBlock) void k_force(AtomsDev a, ListDev l, TabDev t, BoxDev box0,
                                                  const BoxDev* __restrict__ boxes, StepParams sp) {
    const int i = atom_index();
    if (i >= a.n) return;
    const int st = a.statu[i];
    const bool active = (st & STATU_ACTIVE) == STATU_ACTIVE;
    const double4 pi = a.pos[i];
    const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
    double fpx = 0.0, fpy = 0.0, fpz = 0.0;
    const int ti = MT ? a.ityp[i] : 0;
    if (active) {
        const double denki = pi.w;
        const int kv = l.kvois[i];
        const int* __restrict__ li = l.indi + list_base(i, l.mx);
        int jn = kv > 0 ? li[0] : 0;
        for (int iw = 0; iw < kv; ++iw) {
            const int j = jn;
            if (iw + 1 < kv) jn = li[static_cast<size_t>(iw + 1) * l.stride];
            const double4 pj = a.pos[j];
            const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            if (r2 <= t.ru2max) {
                int kind, kind1;
                if constexpr (MT) {
                    const int tj = a.ityp[j];
                    kind = t.kpair[ti * t.ngroup + tj];
                    kind1 = t.kpair[tj * t.ngroup + ti];
                } else {
                    kind = kind1 = t.kpair[0];
                }
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), t.ntab2 - 1);
                const double dk = sk - kk;
                const double4 f = t.f4[kind * t.ntab2 + kk];
                double fb1;
                if constexpr (MT) {
                    const double4 f1 = t.f4[kind1 * t.ntab2 + kk];
                    fb1 = f1.z + dk * f1.w;
                } else {
                    fb1 = f.z + dk * f.w;
                }
                // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r  (CALFORCE_KERNEL :811-813)
                const double fortot = (f.x + dk * f.y) * (inv_r * inv_r) +
                                      ((f.z + dk * f.w) * denki + fb1 * pj.w) * inv_r;
                fpx = fpx + fortot * sx;
                fpy = fpy + fortot * sy;
                fpz = fpz + fortot * sz;
            }
        }
    }
    if (sp.write_f) {
        a.fx[i] = fpx;
        a.fy[i] = fpy;
        a.fz[i] = fpz;
    }
    if constexpr (MODE >= 1)
        integrate_atom<MODE>(a, t, b, sp, i, st, active, pi, ti, fpx, fpy, fpz);
}

void launch_force(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                  const BoxDev* boxes, int multibox, const StepParams& sp, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(grid_for(a.n)), b(kBlock);
    const bool mt = t.ngroup > 1;
#define MSM_F(MB, MT, M) \
    hipLaunchKernelGGL((k_force<MB, MT, M>), g, b, 0, s, a, l, t, box0, boxes, sp)
#define MSM_FM(MB, MT)                      \
    do {                                    \
        if (sp.mode == 0) MSM_F(MB, MT, 0); \
        else if (sp.mode == 1) MSM_F(MB, MT, 1); \
        else MSM_F(MB, MT, 2);              \
    } while (0)
    if (multibox) {
        if (mt) MSM_FM(1, 1); else MSM_FM(1, 0);
    } else {
        if (mt) MSM_FM(0, 1); else MSM_FM(0, 0);
    }
#undef MSM_FM
#undef MSM_F
}

// --------------------------------------------------------------------------------------------
// energy pass
// --------------------------------------------------------------------------------------------

// Two lanes per atom (lanes 2i and 2i+1 take the even and odd list slots; the lane pair's sums
// are added once at the end), the first lane finishes the atom.
template <int MB, int EAM>
__global__ __launch_bounds__(kBlock) void k_epot(AtomsDev a, ListDev l, TabDev t, BoxDev box0,
                                                 const BoxDev* __restrict__ boxes, int* err,
                                                 double* __restrict__ rho_out) {
    const int gt = atom_index();
    const int i = gt >> 1, sub = gt & 1;
    const bool have = i < a.n;  // both lanes of a pair agree, so the shuffle below is paired
    double er0 = 0.0, den0 = 0.0;
    const bool active = have && (a.statu[i] & STATU_ACTIVE) == STATU_ACTIVE;
    if (active) {
        const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
        const double4 pi = a.pos[i];
        const int ti = a.ityp[i];
        const int kv = l.kvois[i];
        for (int iw = sub; iw < kv; iw += 2) {
            const int j = l.indi[list_base(i, l.mx) + static_cast<size_t>(iw) * l.stride];
            const double4 pj = a.pos[j];
            const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
            const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
            const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            if (r2 <= t.ru2max) {
                const int kind = t.kpair[ti * t.ngroup + a.ityp[j]];
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), t.ntab2 - 1);
                const double dk = sk - kk;
                const double4 e = t.e4[kind * t.ntab2 + kk];
                er0 = er0 + (e.x + dk * e.y) * inv_r;
                den0 = den0 + (e.z + dk * e.w);
            }
        }
    }
    er0 = er0 + __shfl_xor(er0, 1, 64);
    den0 = den0 + __shfl_xor(den0, 1, 64);
    if (!have || sub != 0) return;
    if (rho_out) rho_out[i] = den0;
    if (active) {
        if constexpr (!EAM) {
            den0 = -sqrt(den0);
        } else {
            const double sk = den0 / t.rhod + 1.0;
            int kk = static_cast<int>(sk + 0.000001);
            if (kk < 0) kk = 0;
            if (kk > t.nembd) {
                atomicOr(&err[kErrEmbd], 1);
                kk = t.nembd;
            }
            const double4 m = t.m4[t.kembd[a.ityp[i]] * t.nembd2 + kk];
            den0 = m.x + (sk - kk) * m.y;
        }
    }
    a.epot[i] = er0 + den0;
}

void launch_epot(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                 const BoxDev* boxes, int multibox, int* err, hipStream_t s, double* rho_out) {
    if (a.n <= 0) return;
    const dim3 g(grid_for(2 * a.n)), b(kBlock);
    const bool eam = t.pot_type != 0;
    if (multibox) {
        if (eam) hipLaunchKernelGGL((k_epot<1, 1>), g, b, 0, s, a, l, t, box0, boxes, err, rho_out);
        else     hipLaunchKernelGGL((k_epot<1, 0>), g, b, 0, s, a, l, t, box0, boxes, err, rho_out);
    } else {
        if (eam) hipLaunchKernelGGL((k_epot<0, 1>), g, b, 0, s, a, l, t, box0, boxes, err, rho_out);
        else     hipLaunchKernelGGL((k_epot<0, 0>), g, b, 0, s, a, l, t, box0, boxes, err, rho_out);
    }
}

// --------------------------------------------------------------------------------------------
// stand-alone integrator kernels (Predictor_DEV / Correction_DEV)
// --------------------------------------------------------------------------------------------

template <int MB>
__global__ __launch_bounds__(kBlock) void k_predictor(AtomsDev a, TabDev t, BoxDev box0,
                                                      const BoxDev* __restrict__ boxes, double h) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    int stat = a.statu[i];
    if ((stat & STATU_ACTIVE) != STATU_ACTIVE) return;
    const int st0 = stat;
    const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
    const double hs2 = h * 0.5, h2s2 = h * hs2;
    const double cm0 = t.cm[a.ityp[i]];
    const double ax = a.fx[i] / cm0, ay = a.fy[i] / cm0, az = a.fz[i] / cm0;
    double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
    const double4 p = a.pos[i];
    double x = p.x, y = p.y, z = p.z;
    double dx = h * vx + h2s2 * ax;
    if (stat & STATU_FIXPOSX) dx = 0.0;
    x = x + dx;
    if (b.ifpd[0]) {
        if (x > b.lo[0] + b.len[0]) { x = x - b.len[0]; stat |= STATU_PASSBOUND; }
        else if (x < b.lo[0])       { x = x + b.len[0]; stat |= STATU_PASSBOUND; }
    }
    double dy = h * vy + h2s2 * ay;
    if (stat & STATU_FIXPOSY) dy = 0.0;
    y = y + dy;
    if (b.ifpd[1]) {
        if (y > b.lo[1] + b.len[1]) { y = y - b.len[1]; stat |= STATU_PASSBOUND; }
        else if (y < b.lo[1])       { y = y + b.len[1]; stat |= STATU_PASSBOUND; }
    }
    double dz = h * vz + h2s2 * az;
    if (stat & STATU_FIXPOSZ) dz = 0.0;
    z = z + dz;
    if (b.ifpd[2]) {
        if (z > b.lo[2] + b.len[2]) { z = z - b.len[2]; stat |= STATU_PASSBOUND; }
        else if (z < b.lo[2])       { z = z + b.len[2]; stat |= STATU_PASSBOUND; }
    }
    if (!(stat & (STATU_FIXVELX | STATU_FIXPOSX))) vx = vx + hs2 * ax;
    if (!(stat & (STATU_FIXVELY | STATU_FIXPOSY))) vy = vy + hs2 * ay;
    if (!(stat & (STATU_FIXVELZ | STATU_FIXPOSZ))) vz = vz + hs2 * az;
    a.pos[i] = make_double4(x, y, z, p.w);
    a.vx[i] = vx;
    a.vy[i] = vy;
    a.vz[i] = vz;
    a.dx[i] = a.dx[i] + dx;
    a.dy[i] = a.dy[i] + dy;
    a.dz[i] = a.dz[i] + dz;
    if (x > b.lo[0] + b.len[0] || y > b.lo[1] + b.len[1] || z > b.lo[2] + b.len[2])
        stat = STATU_OUTOFBOX | STATU_REFLECT;
    else if (x < b.lo[0] || y < b.lo[1] || z < b.lo[2])
        stat = STATU_OUTOFBOX | STATU_TRANSMIT;
    if (stat != st0) a.statu[i] = stat;
}

void launch_predictor(const AtomsDev& a, const TabDev& t, const BoxDev& box0,
                      const BoxDev* boxes, int multibox, double h, hipStream_t s) {
    if (a.n <= 0) return;
    if (multibox)
        hipLaunchKernelGGL(k_predictor<1>, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, t, box0,
                           boxes, h);
    else
        hipLaunchKernelGGL(k_predictor<0>, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, t, box0,
                           boxes, h);
}

__global__ __launch_bounds__(kBlock) void k_correction(AtomsDev a, TabDev t, double h) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int st = a.statu[i];
    if ((st & STATU_ACTIVE) != STATU_ACTIVE) return;
    const double hs2 = h * 0.5;
    const double cm0 = t.cm[a.ityp[i]];
    if (!(st & (STATU_FIXVELX | STATU_FIXPOSX))) a.vx[i] = a.vx[i] + hs2 * (a.fx[i] / cm0);
    if (!(st & (STATU_FIXVELY | STATU_FIXPOSY))) a.vy[i] = a.vy[i] + hs2 * (a.fy[i] / cm0);
    if (!(st & (STATU_FIXVELZ | STATU_FIXPOSZ))) a.vz[i] = a.vz[i] + hs2 * (a.fz[i] / cm0);
}

void launch_correction(const AtomsDev& a, const TabDev& t, double h, hipStream_t s) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_correction, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, t, h);
}

// --------------------------------------------------------------------------------------------
// thermalisation and quench around the force path (MD_DiffScheme_GPU.F90)
// --------------------------------------------------------------------------------------------

// DAMPING_KERNEL (:125-187): a velocity component opposite to its force component is zeroed, and
// components with FIXPOS set are zeroed (every atom, as the reference).
__global__ void k_damping(AtomsDev a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int st = a.statu[i];
    double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
    if (vx * a.fx[i] < 0.0) vx = 0.0;
    if (vy * a.fy[i] < 0.0) vy = 0.0;
    if (vz * a.fz[i] < 0.0) vz = 0.0;
    if (st & STATU_FIXPOSX) vx = 0.0;
    if (st & STATU_FIXPOSY) vy = 0.0;
    if (st & STATU_FIXPOSZ) vz = 0.0;
    a.vx[i] = vx;
    a.vy[i] = vy;
    a.vz[i] = vz;
}

void launch_damping(const AtomsDev& a, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_damping, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a);
}

// CheckTimestep_KERNEL (MD_DiffScheme_GPU.F90:1068-1139): flag when an active atom's step
// displacement th v + h2s2 F/m (components without a FIXPOS bit) has a square above mxd2; one
// atomic per wave that sees such an atom.
__global__ void k_check_dt(AtomsDev a, TabDev t, double th, double h2s2, double mxd2,
                           int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool over = false;
    if (i < a.n) {
        const int st = a.statu[i];
        if ((st & STATU_ACTIVE) == STATU_ACTIVE) {
            const double cm0 = t.cm[a.ityp[i]];
            double dx = 0.0, dy = 0.0, dz = 0.0;
            if ((st & STATU_FIXPOSX) == 0) dx = th * a.vx[i] + h2s2 * (a.fx[i] / cm0);
            if ((st & STATU_FIXPOSY) == 0) dy = th * a.vy[i] + h2s2 * (a.fy[i] / cm0);
            if ((st & STATU_FIXPOSZ) == 0) dz = th * a.vz[i] + h2s2 * (a.fz[i] / cm0);
            over = dx * dx + dy * dy + dz * dz > mxd2;
        }
    }
    const unsigned long long b = __ballot(over);
    if (b != 0ull && static_cast<int>(threadIdx.x & 63) == __ffsll(b) - 1) atomicOr(flag, 1);
}

void launch_check_dt(const AtomsDev& a, const TabDev& t, double th, double mxd2, int* flag,
                     hipStream_t s) {
    if (a.n > 0)
        hipLaunchKernelGGL(k_check_dt, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, t, th,
                           0.5 * th * th, mxd2, flag);
}

// CALEKIN_KERNEL (:851-904): EK = CM/2 |v|^2 for active atoms with no FIXPOS bit, else -1e32.
__global__ void k_ekin(AtomsDev a, TabDev t, double* __restrict__ ek) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int st = a.statu[i];
    double e = -1.0e32;
    if ((st & STATU_ACTIVE) == STATU_ACTIVE &&
        (st & (STATU_FIXPOSX | STATU_FIXPOSY | STATU_FIXPOSZ)) == 0) {
        const double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
        e = 0.5 * t.cm[a.ityp[i]] * (vx * vx + vy * vy + vz * vz);
    }
    ek[i] = e;
}

void launch_ekin(const AtomsDev& a, const TabDev& t, double* ek, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_ekin, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, t, ek);
}

// One workgroup per box: sum and count of the non-negative EK of the box's atoms (fixed-order
// strided partial sums, then a fixed-shape tree).
__global__ __launch_bounds__(kBlock) void k_box_ekin(const double* __restrict__ ek,
                                                     const int* __restrict__ start,
                                                     const BoxDev* __restrict__ boxes,
                                                     double* __restrict__ sum,
                                                     double* __restrict__ cnt) {
    __shared__ double ls[kBlock], lc[kBlock];
    const BoxDev b = boxes[blockIdx.x];
    const int c0 = b.cell_off, c1 = b.cell_off + b.ncell[0] * b.ncell[1] * b.ncell[2];
    const int a0 = start[c0], a1 = start[c1];
    double sv = 0.0, cv = 0.0;
    for (int i = a0 + threadIdx.x; i < a1; i += kBlock) {
        const double e = ek[i];
        if (e >= 0.0) {
            sv += e;
            cv += 1.0;
        }
    }
    ls[threadIdx.x] = sv;
    lc[threadIdx.x] = cv;
    __syncthreads();
    for (int d = kBlock / 2; d > 0; d >>= 1) {
        if (threadIdx.x < d) {
            ls[threadIdx.x] += ls[threadIdx.x + d];
            lc[threadIdx.x] += lc[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sum[blockIdx.x] = ls[0];
        cnt[blockIdx.x] = lc[0];
    }
}

void launch_box_ekin(const double* ek, const int* start, const BoxDev* boxes, int nbox,
                     double* sum, double* cnt, hipStream_t s) {
    if (nbox > 0)
        hipLaunchKernelGGL(k_box_ekin, dim3(nbox), dim3(kBlock), 0, s, ek, start, boxes, sum, cnt);
}

// VelScaling_KERNEL (:1262-1326): active atoms' velocities times their box's factor; components
// with FIXPOS or FIXVEL set become zero.
__global__ void k_velscale(AtomsDev a, const double* __restrict__ scal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int st = a.statu[i];
    if ((st & STATU_ACTIVE) != STATU_ACTIVE) return;
    const double f = scal[a.boxid[i]];
    a.vx[i] = (st & (STATU_FIXPOSX | STATU_FIXVELX)) ? 0.0 : a.vx[i] * f;
    a.vy[i] = (st & (STATU_FIXPOSY | STATU_FIXVELY)) ? 0.0 : a.vy[i] * f;
    a.vz[i] = (st & (STATU_FIXPOSZ | STATU_FIXVELZ)) ? 0.0 : a.vz[i] * f;
}

void launch_velscale(const AtomsDev& a, const double* scal, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_velscale, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, scal);
}

// max_i |epot[i] - epot0[i]| into *out (bit pattern of a non-negative double, which orders as an
// unsigned integer); *out must be zero on entry.  Then epot0 = epot.
__global__ void k_epot_maxdelta(const double* __restrict__ epot, double* __restrict__ epot0,
                                int n, unsigned long long* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n) {
        d = fabs(epot[i] - epot0[i]);
        epot0[i] = epot[i];
    }
    for (int o = 32; o > 0; o >>= 1) d = fmax(d, __shfl_xor(d, o, 64));
    if ((threadIdx.x & 63) == 0 && d > 0.0)
        atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(d)));
}

void launch_epot_maxdelta(const double* epot, double* epot0, int n, unsigned long long* out,
                          hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_epot_maxdelta, dim3(grid_for(n)), dim3(kBlock), 0, s, epot, epot0, n,
                           out);
}

// --------------------------------------------------------------------------------------------
// steepest-descent quench (MD_SteepestScheme_GPU.F90) on 3-component vectors held as three
// arrays of n doubles, like the forces; the reductions of DevDot_product / DevMaxAbsVal
// (MSM_MultiGPU_Basic.F90) are deterministic: fixed grid, fixed-order partial sums, fixed tree
// --------------------------------------------------------------------------------------------

int quench_blocks(int n) { return std::max(1, std::min(kQuenchBlocks, (n + kBlock - 1) / kBlock)); }

template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* __restrict__ part) {
    __shared__ double ls[NV][kBlock];
#pragma unroll
    for (int c = 0; c < NV; ++c) ls[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int d = kBlock / 2; d > 0; d >>= 1) {
        if (threadIdx.x < d) {
#pragma unroll
            for (int c = 0; c < NV; ++c) ls[c][threadIdx.x] += ls[c][threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) part[NV * blockIdx.x + c] = ls[c][0];
    }
}

// partial sums of u . v
__global__ __launch_bounds__(kBlock) void k_dot3(P3 u, P3 v, int n, double* __restrict__ part) {
    double acc[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        acc[0] += u.x[i] * v.x[i] + u.y[i] * v.y[i] + u.z[i] * v.z[i];
    block_sums<1>(acc, part);
}

// partial sums of dx . (p - f) and (p - f) . (p - f)  (DOTDXDF and DOTDF of Do_Steepest0)
__global__ __launch_bounds__(kBlock) void k_dxdf(P3 dx, P3 p, P3 f, int n,
                                                 double* __restrict__ part) {
    double acc[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double gx = p.x[i] - f.x[i], gy = p.y[i] - f.y[i], gz = p.z[i] - f.z[i];
        acc[0] += dx.x[i] * gx + dx.y[i] * gy + dx.z[i] * gz;
        acc[1] += gx * gx + gy * gy + gz * gz;
    }
    block_sums<2>(acc, part);
}

// out[c] = sum of the nblk partials of component c, in block order then a fixed tree
template <int NV>
__global__ __launch_bounds__(kBlock) void k_sum_parts(const double* __restrict__ part, int nblk,
                                                      double* __restrict__ out) {
    double acc[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) {
#pragma unroll
        for (int c = 0; c < NV; ++c) acc[c] += part[NV * b + c];
    }
    block_sums<NV>(acc, out);
}

void launch_dot3(P3 u, P3 v, int n, double* part, double* out, hipStream_t s) {
    const int nb = quench_blocks(n);
    hipLaunchKernelGGL(k_dot3, dim3(nb), dim3(kBlock), 0, s, u, v, n, part);
    hipLaunchKernelGGL(k_sum_parts<1>, dim3(1), dim3(kBlock), 0, s, part, nb, out);
}

void launch_dxdf(P3 dx, P3 p, P3 f, int n, double* part, double* out, hipStream_t s) {
    const int nb = quench_blocks(n);
    hipLaunchKernelGGL(k_dxdf, dim3(nb), dim3(kBlock), 0, s, dx, p, f, n, part);
    hipLaunchKernelGGL(k_sum_parts<2>, dim3(1), dim3(kBlock), 0, s, part, nb, out);
}

// dst = c * src (DevMultiply_noshift; dst may be src)
__global__ void k_scale3(P3 src, double c, int n, P3 dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst.x[i] = c * src.x[i];
    dst.y[i] = c * src.y[i];
    dst.z[i] = c * src.z[i];
}

void launch_scale3(P3 src, double c, int n, P3 dst, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_scale3, dim3(grid_for(n)), dim3(kBlock), 0, s, src, c, n, dst);
}

// d = f + g * d  (the conjugate direction update of Do_CG0_Forsteps_DEV)
__global__ void k_dir_update(P3 f, double g, int n, P3 d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d.x[i] = f.x[i] + g * d.x[i];
    d.y[i] = f.y[i] + g * d.y[i];
    d.z[i] = f.z[i] + g * d.z[i];
}

void launch_dir_update(P3 f, double g, int n, P3 d, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_dir_update, dim3(grid_for(n)), dim3(kBlock), 0, s, f, g, n, d);
}

// max |component| into *out as the bit pattern of a non-negative double (*out zero on entry)
__global__ void k_maxabs3(P3 v, int n, unsigned long long* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n) d = fmax(fabs(v.x[i]), fmax(fabs(v.y[i]), fabs(v.z[i])));
    for (int o = 32; o > 0; o >>= 1) d = fmax(d, __shfl_xor(d, o, 64));
    if ((threadIdx.x & 63) == 0 && d > 0.0)
        atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(d)));
}

void launch_maxabs3(P3 v, int n, unsigned long long* out, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_maxabs3, dim3(grid_for(n)), dim3(kBlock), 0, s, v, n, out);
}

// x += dx with the bounds of DevAdd (AddBD_DevVec_DF_KERNEL0, MSM_MultiGPU_Basic.F90:5444-5477):
// in a periodic direction a coordinate above BOXUP moves down by the box length, one below
// BOXLOW up by it; other directions are not bounded.  F' in pos.w is left alone.
template <int MB>
__global__ void k_move3(AtomsDev a, P3 dx, BoxDev box0, const BoxDev* __restrict__ boxes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
    double4 p = a.pos[i];
    double x[3] = {p.x + dx.x[i], p.y + dx.y[i], p.z + dx.z[i]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (!b.ifpd[k]) continue;
        const double lb = b.lo[k], hb = b.lo[k] + b.len[k];
        if (x[k] > hb)
            x[k] = x[k] - (hb - lb);
        else if (x[k] < lb)
            x[k] = x[k] + (hb - lb);
    }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    a.pos[i] = p;
}

void launch_move3(const AtomsDev& a, P3 dx, const BoxDev& box0, const BoxDev* boxes, bool mb,
                  hipStream_t s) {
    if (a.n <= 0) return;
    if (mb)
        hipLaunchKernelGGL(k_move3<1>, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, dx, box0, boxes);
    else
        hipLaunchKernelGGL(k_move3<0>, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, dx, box0, boxes);
}

// --------------------------------------------------------------------------------------------
// copy-out helpers (sorted -> caller order), box totals, list statistics
// --------------------------------------------------------------------------------------------

__global__ void k_unsort_d(const double* __restrict__ src, double* __restrict__ dst,
                           const int* __restrict__ gid, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[gid[p]] = src[p];
}
__global__ void k_unsort_i(const int* __restrict__ src, int* __restrict__ dst,
                           const int* __restrict__ gid, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[gid[p]] = src[p];
}
__global__ void k_unsort_pos(const double4* __restrict__ src, double* __restrict__ dst,
                             const int* __restrict__ gid, int n, int comp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double4 v = src[p];
    const double c = comp == 0 ? v.x : comp == 1 ? v.y : comp == 2 ? v.z : v.w;
    dst[gid[p]] = c;
}
// Caller-order coordinates (three arrays of n) into the sorted double4 records (w kept) --
// CopyXPFrom_Host_to_Devices without a list rebuild.
__global__ void k_sort_pos(const double* __restrict__ src, double4* __restrict__ dst,
                           const int* __restrict__ gid, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int q = gid[p];
    double4 v = dst[p];
    v.x = src[q];
    v.y = src[q + static_cast<size_t>(n)];
    v.z = src[q + 2 * static_cast<size_t>(n)];
    dst[p] = v;
}
// One caller-order array into sorted order.
__global__ void k_sort_d(const double* __restrict__ src, double* __restrict__ dst,
                         const int* __restrict__ gid, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[gid[p]];
}
__global__ void k_unsort_list(const int* __restrict__ kvois, const int* __restrict__ indi,
                              int stride, const int* __restrict__ gid, int n, int mx,
                              int* __restrict__ kv_out, int* __restrict__ indi_out, int ld) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int g = gid[p];
    const int kv = kvois[p];
    kv_out[g] = kv;
    for (int iw = 0; iw < kv && iw < mx; ++iw)
        indi_out[g + static_cast<size_t>(iw) * ld] = gid[indi[list_base(p, mx) + static_cast<size_t>(iw) * stride]] + 1;
}

// Per-atom stress AVP(N, 9) in the caller's order from the per-atom virial halves (6 components,
// component-major, sorted order): AVP(i, 3a+b) = 2 * vir(ab), the full-list sum
// sum_j s_a s_b FORTOT of CAL_FS_AtomicStress_KERNEL (MD_FS_ForceTable_GPU.F90:1746-1965),
// components P11, P12, P13, P21, ..., P33.  Doubling is exact, so this is that sum itself.
__global__ void k_unsort_stress(const double* __restrict__ vir, double* __restrict__ avp,
                                const int* __restrict__ gid, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    // vir components: xx, yy, zz, xy, xz, yz
    constexpr int kMap[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    const size_t i = static_cast<size_t>(gid[p]);
#pragma unroll
    for (int c = 0; c < 9; ++c)
        avp[static_cast<size_t>(c) * n + i] = 2.0 * vir[static_cast<size_t>(kMap[c]) * n + p];
}

void launch_unsort_stress(const double* vir, double* avp, const int* gid, int n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_unsort_stress, dim3(grid_for(n)), dim3(kBlock), 0, s, vir, avp, gid,
                           n);
}

void launch_unsort_d(const double* src, double* dst, const int* gid, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_unsort_d, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, gid, n);
}
void launch_unsort_i(const int* src, int* dst, const int* gid, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_unsort_i, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, gid, n);
}
void launch_unsort_pos(const double4* src, double* dst, const int* gid, int n, int comp,
                       hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_unsort_pos, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, gid, n, comp);
}
void launch_sort_pos(const double* src, double4* dst, const int* gid, int n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_sort_pos, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, gid, n);
}
__global__ void k_add_sorted(const double* __restrict__ src, double* __restrict__ dst,
                             const int* __restrict__ gid, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = dst[p] + src[gid[p]];
}
void launch_add_sorted(const double* src, double* dst, const int* gid, int n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_add_sorted, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, gid, n);
}
void launch_sort_d(const double* src, double* dst, const int* gid, int n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_sort_d, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, gid, n);
}
void launch_unsort_list(const int* kvois, const int* indi, int stride, const int* gid, int n,
                        int mx, int* kv_out, int* indi_out, int ld, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_unsort_list, dim3(grid_for(n)), dim3(kBlock), 0, s, kvois, indi,
                           stride, gid, n, mx, kv_out, indi_out, ld);
}

// Virial tensor (CALPTENSOR_KERNEL, MD_FS_ForceTable_GPU.F90:998-1272): per active atom i,
// sum over in-cutoff neighbours of FORTOT/2 * s (x) s (the full list sees each pair twice), in
// the six distinct components {xx, yy, zz, xy, xz, yz}; the atom's force is returned beside them,
// as the force pass computes it.  F' of the density pass is read from pos.w.
template <int MB, int MT>
__device__ inline void virial_atom(const AtomsDev& a, const ListDev& l, const TabDev& t,
                                   const BoxDev& box0, const BoxDev* __restrict__ boxes, int i,
                                   double (&p)[6], double& fpx, double& fpy, double& fpz) {
    if ((a.statu[i] & STATU_ACTIVE) != STATU_ACTIVE) return;
    const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
    const double4 pi = a.pos[i];
    const int ti = MT ? a.ityp[i] : 0;
    const int kv = l.kvois[i];
    for (int iw = 0; iw < kv; ++iw) {
        const int j = l.indi[list_base(i, l.mx) + static_cast<size_t>(iw) * l.stride];
        const double4 pj = a.pos[j];
        const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
        const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
        const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
        const double r2 = sx * sx + sy * sy + sz * sz;
        if (!(r2 <= t.ru2max)) continue;
        int kind = t.kpair[0], kind1 = t.kpair[0];
        if constexpr (MT) {
            const int tj = a.ityp[j];
            kind = t.kpair[ti * t.ngroup + tj];
            kind1 = t.kpair[tj * t.ngroup + ti];
        }
        double inv_r;
        const double sk = table_coord(r2, t.csi, inv_r);
        const int kk = min(static_cast<int>(sk), t.ntab2 - 1);
        const double dk = sk - kk;
        const double4 f = t.f4[kind * t.ntab2 + kk];
        const double4 f1 = t.f4[kind1 * t.ntab2 + kk];
        const double fortot = (f.x + dk * f.y) * (inv_r * inv_r) +
                              ((f.z + dk * f.w) * pi.w + (f1.z + dk * f1.w) * pj.w) * inv_r;
        fpx = fpx + fortot * sx;
        fpy = fpy + fortot * sy;
        fpz = fpz + fortot * sz;
        const double h = fortot * 0.5;
        p[0] = p[0] + sx * sx * h;
        p[1] = p[1] + sy * sy * h;
        p[2] = p[2] + sz * sz * h;
        p[3] = p[3] + sx * sy * h;
        p[4] = p[4] + sx * sz * h;
        p[5] = p[5] + sy * sz * h;
    }
}

// Per-atom virial, stored component-major (stride n), and the forces.
template <int MB, int MT>
__global__ __launch_bounds__(kBlock) void k_virial(AtomsDev a, ListDev l, TabDev t, BoxDev box0,
                                                   const BoxDev* __restrict__ boxes,
                                                   double* __restrict__ vir) {
    const int i = atom_index();
    if (i >= a.n) return;
    double p[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double fpx = 0.0, fpy = 0.0, fpz = 0.0;
    virial_atom<MB, MT>(a, l, t, box0, boxes, i, p, fpx, fpy, fpz);
    a.fx[i] = fpx;
    a.fy[i] = fpy;
    a.fz[i] = fpz;
    for (int c = 0; c < 6; ++c) vir[static_cast<size_t>(c) * a.n + i] = p[c];
}

// The same sums reduced without a per-atom buffer.  Block (x, y < nbox) takes atoms
// [a0 + 256 x, a0 + 256 x + 256) of box y's cell range [a0, a1) and writes its six partial sums
// (fixed-shape tree) to part[(y nchunk + x) 6 + c]; blocks with y = nbox write the forces of the
// atoms after the last cell (the out-of-box and inactive buckets), grid-strided, and no partials.
// Every atom's force is written exactly as k_virial writes it.
template <int MB, int MT>
__global__ __launch_bounds__(kBlock) void k_virial_part(AtomsDev a, ListDev l, TabDev t,
                                                        BoxDev box0,
                                                        const BoxDev* __restrict__ boxes,
                                                        const int* __restrict__ start, int nbox,
                                                        int ncell_total, int nchunk,
                                                        double* __restrict__ part) {
    __shared__ double lds[6][kBlock];
    const int by = blockIdx.y, tid = threadIdx.x;
    if (by == nbox) {
        for (int i = start[ncell_total] + blockIdx.x * kBlock + tid; i < a.n;
             i += nchunk * kBlock) {
            double p[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            double fpx = 0.0, fpy = 0.0, fpz = 0.0;
            virial_atom<MB, MT>(a, l, t, box0, boxes, i, p, fpx, fpy, fpz);
            a.fx[i] = fpx;
            a.fy[i] = fpy;
            a.fz[i] = fpz;
        }
        return;
    }
    const BoxDev& bx = boxes[by];
    const int c0 = bx.cell_off, c1 = bx.cell_off + bx.ncell[0] * bx.ncell[1] * bx.ncell[2];
    const int i = start[c0] + blockIdx.x * kBlock + tid;
    double p[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < start[c1]) {
        double fpx = 0.0, fpy = 0.0, fpz = 0.0;
        virial_atom<MB, MT>(a, l, t, box0, boxes, i, p, fpx, fpy, fpz);
        a.fx[i] = fpx;
        a.fy[i] = fpy;
        a.fz[i] = fpz;
    }
    for (int c = 0; c < 6; ++c) lds[c][tid] = p[c];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w)
            for (int c = 0; c < 6; ++c) lds[c][tid] += lds[c][tid + w];
        __syncthreads();
    }
    if (tid < 6) part[(static_cast<size_t>(by) * nchunk + blockIdx.x) * 6 + tid] = lds[tid][0];
}

// One workgroup per box: the box's nchunk partials in fixed order (strided sums, then a tree),
// written as the symmetric 3x3 tensor, column-major, at out[9 b].
__global__ __launch_bounds__(kBlock) void k_virial_reduce(const double* __restrict__ part,
                                                          int nchunk, double* __restrict__ out) {
    __shared__ double lds[6][kBlock];
    const int b = blockIdx.x, tid = threadIdx.x;
    double p[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = tid; k < nchunk; k += kBlock)
        for (int c = 0; c < 6; ++c) p[c] += part[(static_cast<size_t>(b) * nchunk + k) * 6 + c];
    for (int c = 0; c < 6; ++c) lds[c][tid] = p[c];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w)
            for (int c = 0; c < 6; ++c) lds[c][tid] += lds[c][tid + w];
        __syncthreads();
    }
    if (tid < 6) {
        const int row[6] = {0, 1, 2, 0, 0, 1}, col[6] = {0, 1, 2, 1, 2, 2};
        const double v = lds[tid][0];
        out[9 * b + row[tid] + 3 * col[tid]] = v;
        out[9 * b + col[tid] + 3 * row[tid]] = v;
    }
}

void launch_virial_boxes(const AtomsDev& a, const ListDev& l, const TabDev& t,
                         const BoxDev& box0, const BoxDev* boxes, int multibox, const int* start,
                         int nbox, int ncell_total, int nchunk, double* part, double* out,
                         hipStream_t s) {
    if (a.n <= 0 || nbox <= 0 || nchunk <= 0) return;
    const dim3 g(nchunk, nbox + 1), bl(kBlock);
    const bool mt = t.ngroup > 1;
    if (multibox) {
        if (mt) hipLaunchKernelGGL((k_virial_part<1, 1>), g, bl, 0, s, a, l, t, box0, boxes, start, nbox, ncell_total, nchunk, part);
        else    hipLaunchKernelGGL((k_virial_part<1, 0>), g, bl, 0, s, a, l, t, box0, boxes, start, nbox, ncell_total, nchunk, part);
    } else {
        if (mt) hipLaunchKernelGGL((k_virial_part<0, 1>), g, bl, 0, s, a, l, t, box0, boxes, start, nbox, ncell_total, nchunk, part);
        else    hipLaunchKernelGGL((k_virial_part<0, 0>), g, bl, 0, s, a, l, t, box0, boxes, start, nbox, ncell_total, nchunk, part);
    }
    hipLaunchKernelGGL(k_virial_reduce, dim3(nbox), bl, 0, s, part, nchunk, out);
}

void launch_virial(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                   const BoxDev* boxes, int multibox, double* vir, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(grid_for(a.n)), bl(kBlock);
    const bool mt = t.ngroup > 1;
    if (multibox) {
        if (mt) hipLaunchKernelGGL((k_virial<1, 1>), g, bl, 0, s, a, l, t, box0, boxes, vir);
        else    hipLaunchKernelGGL((k_virial<1, 0>), g, bl, 0, s, a, l, t, box0, boxes, vir);
    } else {
        if (mt) hipLaunchKernelGGL((k_virial<0, 1>), g, bl, 0, s, a, l, t, box0, boxes, vir);
        else    hipLaunchKernelGGL((k_virial<0, 0>), g, bl, 0, s, a, l, t, box0, boxes, vir);
    }
}

// One workgroup per box: fixed-order strided partial sums then a fixed-shape tree, so the total
// is the same on every run.
__global__ __launch_bounds__(kBlock) void k_box_sum(const double* __restrict__ epot,
                                                    const int* __restrict__ start,
                                                    const BoxDev* __restrict__ boxes,
                                                    double* __restrict__ out) {
    __shared__ double lds[kBlock];
    const BoxDev b = boxes[blockIdx.x];
    const int c0 = b.cell_off, c1 = b.cell_off + b.ncell[0] * b.ncell[1] * b.ncell[2];
    const int a0 = start[c0], a1 = start[c1];
    double s = 0.0;
    for (int i = a0 + threadIdx.x; i < a1; i += kBlock) s += epot[i];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = lds[0];
}

void launch_box_sum(const double* epot, const int* start, const BoxDev* boxes, int nbox,
                    double* out, hipStream_t s) {
    if (nbox > 0)
        hipLaunchKernelGGL(k_box_sum, dim3(nbox), dim3(kBlock), 0, s, epot, start, boxes, out);
}

// Sum and maximum of KVOIS: grid-stride partial results, reduced per wave and then per
// workgroup, so that only two atomics per workgroup reach the two global counters.
__global__ void k_list_stats(const int* __restrict__ kvois, int n, unsigned long long* total,
                             int* maxk) {
    __shared__ unsigned long long ws[kBlock / 64];
    __shared__ int wm[kBlock / 64];
    unsigned long long s = 0;
    int m = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int v = kvois[p];
        s += static_cast<unsigned long long>(v);
        m = max(m, v);
    }
    for (int d = 32; d > 0; d >>= 1) {
        s += __shfl_down(s, d, 64);
        m = max(m, __shfl_down(m, d, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ws[w] = s;
        wm[w] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < static_cast<int>(blockDim.x >> 6); ++k) {
            s += ws[k];
            m = max(m, wm[k]);
        }
        atomicAdd(total, s);
        atomicMax(maxk, m);
    }
}

void launch_list_stats(const int* kvois, int n, unsigned long long* total, int* maxk,
                       hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = grid_for(n) < 1024u ? grid_for(n) : 1024u;
    hipLaunchKernelGGL(k_list_stats, dim3(g), dim3(kBlock), 0, s, kvois, n, total, maxk);
}

}  // namespace msm
