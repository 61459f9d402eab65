// msm_bricks.hip -- LDS-staged kernels over bricks of cells: neighbour-list build, density pass
// and force pass (with the fused integrator).
//
// Atoms are cell-sorted, so each cell is a contiguous index range.  A brick is at most 4x4x4
// cells of one box; its halo is the brick plus one cell layer (periodic images resolved), at most
// 6x6x6 = 216 cells.  One workgroup per brick copies the halo atoms into LDS with
// coalesced loads (consecutive threads read consecutive atoms of a cell run) and then every home
// atom works out of LDS.  The list build records each neighbour both as a caller-visible sorted
// index (INDI) and as its 16-bit LDS slot in the home brick's halo; the density and force passes
// rebuild the same halo (cell ranges do not change between rebuilds), read the slot list and
// never gather from global memory.  A brick whose halo exceeds the LDS capacity of a kernel
// (kNlCap for the list build, kNlCapPass for the passes) is processed from global memory
// instead (same results, slower).
//
// Reference semantics: list predicate MD_NeighborsList.F90:396-633; density, force, integrator
// as in msm_kernels.hip (MD_FS/EAM_ForceTable_GPU.F90 PRECALFOR/CALFORCE, MD_DiffScheme_GPU.F90).
#include "msm_brick.h"

namespace msm {

// --------------------------------------------------------------------------------------------
// list build
// --------------------------------------------------------------------------------------------

// Small-brick list build (sparse cells): one kNlSmallBlock-thread workgroup per brick of at most
// kNlCapPass halo atoms, one lane per home atom.  The LDS tables stay near 40 KB (a neighbour's
// sorted index is its cell's base plus its halo slot, not a stored array), so four workgroups
// share a CU.  The candidates of each stencil cell are tested kNlGroup at a time (their LDS
// loads issued together) and the hits stored in candidate order.  SIDX: also store each
// neighbour's halo slot (read by the LDS-brick passes).  Measured alternatives: groups of 8,
// and hit masks with one store round per hit of the busiest lane -- both slower.  The list's slot
// stride is kListRow (the `stride` argument, which only the global fallback reads, equals it).
constexpr int kNlGroup = 4;
// 512 threads: a 4^3-cell brick of a bcc metal at the usual skin holds ~420 home atoms, which a
// 256-thread workgroup served in two rounds (the second 60% full); with 53 VGPRs and 40 KB of LDS
// four such workgroups still share a CU.
constexpr int kNlSmallBlock = 512;

template <int MT, int SIDX>
__global__ __launch_bounds__(kNlSmallBlock) void k_nlist_small(
    const int* __restrict__ brick_box, const double4* __restrict__ pos,
    const int* __restrict__ ityp, const int* __restrict__ start,
    const BoxDev* __restrict__ boxes, int ngroup, const double* __restrict__ rnl2, int mx,
    int* __restrict__ kvois, int* __restrict__ indi, unsigned short* __restrict__ sidx,
    int stride, int* err) {
#pragma clang fp contract(off)
    // about 40 KB for one atom type, so that four workgroups share a CU.  The x, y, z rows are
    // one array with kNlGroup - 1 slots of padding at its end: a candidate group reads up to
    // kNlGroup - 1 slots past its cell (the next cell's atoms, the next row's first ones or the
    // padding) and masks them, so the loads need no clamped index.
    __shared__ double lxyz[3 * kNlCapPass + kNlGroup - 1];
    __shared__ int lt[MT ? kNlCapPass + kNlGroup - 1 : 1];
    double* const lx = lxyz;
    double* const ly = lxyz + kNlCapPass;
    double* const lz = lxyz + 2 * kNlCapPass;
    __shared__ int hbase[kNlMaxHalo];  // sorted index of halo slot q in cell h: hbase[h] + q
    __shared__ signed char hsh[kNlMaxHalo * 3];
    __shared__ double srnl2[MT ? kMaxGroup * kMaxGroup : 1];
    __shared__ BrickLds L;

    const int tid = threadIdx.x;
    const int bi = xcd_block(gridDim.x);
    const BoxDev& B = boxes[brick_box[bi]];
    BrickInfo g;
    brick_info(B, bi, g);
    if constexpr (MT)
        for (int k = tid; k < ngroup * ngroup; k += blockDim.x) srnl2[k] = rnl2[k];
    const int total = brick_halo(B, g, start, L.hcell, hsh, L.hoff, L.homeoff, L.scan);
    const bool staged = total <= kNlCapPass;
    for (int k = tid; k < g.nh; k += blockDim.x)
        hbase[k] = L.hcell[k] >= 0 ? start[L.hcell[k]] - L.hoff[k] : 0;
    if (staged) {
        for_halo_slots(g, start, L.hcell, L.hoff, [&](int s, int q, int h) {
            const double4 pq = pos[q];
            // the image shift is added once, as XP(J) + CXYZ in MD_NeighborsList.F90:582
            lx[s] = pq.x + shift_of(hsh[3 * h], B.len[0]);
            ly[s] = pq.y + shift_of(hsh[3 * h + 1], B.len[1]);
            lz[s] = pq.z + shift_of(hsh[3 * h + 2], B.len[2]);
            if constexpr (MT) lt[s] = ityp[q];
        });
        // the types read past the last staged atom index srnl2: keep them valid
        if constexpr (MT)
            if (tid < kNlGroup - 1) lt[total + tid] = 0;
    }
    __syncthreads();
    const int nhome = L.homeoff[g.nhc];
    const double rsingle = rnl2[0];
    for (int t = tid; t < nhome; t += blockDim.x) {
        int h, off;
        home_atom(t, g, L.homeoff, h, off);
        if (!staged) {
            const int p = start[L.hcell[h]] + off;
            nlist_atom_global(p, L.hcell[h] - B.cell_off, B, pos, ityp, start, ngroup, rnl2, mx,
                              kvois, indi, stride, err);
            continue;
        }
        const int si = L.hoff[h] + off;
        const int p = hbase[h] + si;
        const double xi = lx[si], yi = ly[si], zi = lz[si];
        const int ti = MT ? lt[si] : 0;
        // slot c of this atom is li[c * kListRow]: the store pointer advances by kListRow per hit
        // and stops storing at the row's end (the count goes on, for the overflow flag)
        int* __restrict__ const li = indi + list_base(p, mx);
        int* const lend = li + static_cast<size_t>(mx) * kListRow;
        int* __restrict__ w = li;
        // one stencil cell, candidates in slot order; OWN: the home cell (stencil entry 0), the
        // only one that holds atom si itself
        auto scan = [&](auto own, int hn) {
            const int qe = L.hoff[hn + 1];
            const int gb = hbase[hn];
            for (int q0 = L.hoff[hn]; q0 < qe; q0 += kNlGroup) {
                bool hit[kNlGroup];
#pragma unroll
                for (int u = 0; u < kNlGroup; ++u) {
                    const int q = q0 + u;
                    // XPT(I) - (XP(J) + CXYZ), squared and summed left to right, no contraction
                    const double dx = xi - lx[q];
                    const double dy = yi - ly[q];
                    const double dz = zi - lz[q];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double lim = MT ? srnl2[ti * ngroup + lt[q]] : rsingle;
                    bool ok = u == 0 || q < qe;
                    if constexpr (decltype(own)::value) ok = ok && q != si;
                    hit[u] = ok && r2 < lim;
                }
#pragma unroll
                for (int u = 0; u < kNlGroup; ++u) {
                    if (hit[u]) {
                        if (w < lend) {
                            *w = gb + q0 + u;
                            if constexpr (SIDX)
                                sidx[w - indi] = static_cast<unsigned short>(q0 + u);
                        }
                        w += kListRow;
                    }
                }
            }
        };
        scan(std::true_type{}, h);
        for (int s = 1; s < 27; ++s)
            scan(std::false_type{}, h + c_nix[s] + g.H[0] * (c_niy[s] + g.H[1] * c_niz[s]));
        int c = static_cast<int>((w - li) / kListRow);
        if (c > mx) {
            atomicOr(&err[kErrOverflow], 1);
            c = mx;
        }
        kvois[p] = c;
    }
}


// Wide-brick list build (dense cells): one kNlBlock-thread workgroup per brick.  The halo (up to `cap` atoms, image shifts added as
// the single rounding XP(J) + CXYZ of MD_NeighborsList.F90:582) is staged in dynamic LDS; then T
// adjacent lanes serve one home atom: for each of its 27 stencil cells, in the reference stencil
// order, the T lanes test kNlUnroll*T consecutive candidates of the cell per round (lane `sub`
// takes candidates sub, sub+T, ...; their LDS loads are issued together), and wave ballots place
// the hits in candidate order, so every atom's list comes out in exactly the serial order
// (stencil cell, then slot) whatever T is.  A brick whose halo exceeds `cap` falls back to the
// global-memory scan.
constexpr int kNlUnroll = 4;
// Slots of padding after the staged rows: a round of kNlUnroll * T candidates (T <= 8) reads up
// to kNlUnroll * T - 1 slots past its cell and masks them.
constexpr int kNlPad = kNlUnroll * 8 - 1;

template <int MT, int T, int SIDX>
__global__ __launch_bounds__(kNlBlock, MT ? 4 : 8) void k_nlist_bricks(
    const int* __restrict__ brick_box, const double4* __restrict__ pos,
    const int* __restrict__ ityp, const int* __restrict__ start,
    const BoxDev* __restrict__ boxes, int ngroup, const double* __restrict__ rnl2, int mx,
    int* __restrict__ kvois, int* __restrict__ indi, unsigned short* __restrict__ sidx,
    int stride, int* err, int cap) {
#pragma clang fp contract(off)
    static_assert(kNlUnroll * T - 1 <= kNlPad, "padding too small for T");
    extern __shared__ double s_nl[];
    double* lx = s_nl;
    double* ly = lx + cap;
    double* lz = ly + cap;
    int* lt = reinterpret_cast<int*>(lz + cap + kNlPad);  // MT only
    __shared__ int hbase[kNlMaxHalo];  // sorted index of halo slot q in cell h: hbase[h] + q
    __shared__ signed char hsh[kNlMaxHalo * 3];
    __shared__ double srnl2[MT ? kMaxGroup * kMaxGroup : 1];
    __shared__ BrickLds L;

    const int tid = threadIdx.x;
    const int bi = xcd_block(gridDim.x);
    const BoxDev& B = boxes[brick_box[bi]];
    BrickInfo g;
    brick_info(B, bi, g);
    if constexpr (MT)
        for (int k = tid; k < ngroup * ngroup; k += blockDim.x) srnl2[k] = rnl2[k];
    const int total = brick_halo(B, g, start, L.hcell, hsh, L.hoff, L.homeoff, L.scan);
    const bool staged = total <= cap;
    for (int k = tid; k < g.nh; k += blockDim.x)
        hbase[k] = L.hcell[k] >= 0 ? start[L.hcell[k]] - L.hoff[k] : 0;
    if (staged) {
        for_halo_slots(g, start, L.hcell, L.hoff, [&](int s, int q, int h) {
            const double4 pq = pos[q];
            lx[s] = pq.x + shift_of(hsh[3 * h], B.len[0]);
            ly[s] = pq.y + shift_of(hsh[3 * h + 1], B.len[1]);
            lz[s] = pq.z + shift_of(hsh[3 * h + 2], B.len[2]);
            if constexpr (MT) lt[s] = ityp[q];
        });
        // the types read past the last staged atom index srnl2: keep them valid
        if constexpr (MT)
            if (tid < kNlPad) lt[total + tid] = 0;
    }
    __syncthreads();
    const int nhome = L.homeoff[g.nhc];
    const double rsingle = rnl2[0];
    const int sub = tid % T;
    const int gbase = (tid & 63) - sub;  // first lane of this atom's group in the wave
    const unsigned below = (1u << sub) - 1u, gmask = (T == 32) ? ~0u : ((1u << T) - 1u);
    for (int t0 = 0; t0 < nhome * T; t0 += blockDim.x) {
        const int a = (t0 + tid) / T;
        if (a >= nhome) continue;
        int h, off;
        home_atom(a, g, L.homeoff, h, off);
        if (!staged) {
            if (sub == 0) {
                const int p = start[L.hcell[h]] + off;
                nlist_atom_global(p, L.hcell[h] - B.cell_off, B, pos, ityp, start, ngroup, rnl2,
                                  mx, kvois, indi, stride, err);
            }
            continue;
        }
        const int si = L.hoff[h] + off;
        const int p = hbase[h] + si;
        const double xi = lx[si], yi = ly[si], zi = lz[si];
        const int ti = MT ? lt[si] : 0;
        const size_t lb = list_base(p, mx);
        int c = 0;
        // one stencil cell; OWN: the home cell (stencil entry 0), the only one holding atom si.
        // Candidates past the cell are read (the next cell, the next row or the padding) and
        // masked.
        auto scan = [&](auto own, int hn) {
            const int qe = L.hoff[hn + 1];
            const int gb = hbase[hn];
            for (int q0 = L.hoff[hn]; q0 < qe; q0 += kNlUnroll * T) {
                bool hit[kNlUnroll];
#pragma unroll
                for (int u = 0; u < kNlUnroll; ++u) {
                    const int q = q0 + u * T + sub;
                    // XPT(I) - (XP(J) + CXYZ), squared and summed left to right, no contraction
                    const double dx = xi - lx[q];
                    const double dy = yi - ly[q];
                    const double dz = zi - lz[q];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double lim = MT ? srnl2[ti * ngroup + lt[q]] : rsingle;
                    bool ok = q < qe;
                    if constexpr (decltype(own)::value) ok = ok && q != si;
                    hit[u] = ok && r2 < lim;
                }
#pragma unroll
                for (int u = 0; u < kNlUnroll; ++u) {
                    const unsigned grp =
                        static_cast<unsigned>(__ballot(hit[u]) >> gbase) & gmask;
                    if (hit[u]) {
                        const int w = c + __popc(grp & below);
                        if (w < mx) {
                            const int q = q0 + u * T + sub;
                            const size_t at = lb + static_cast<unsigned>(w) * kListRow;
                            indi[at] = gb + q;
                            if constexpr (SIDX) sidx[at] = static_cast<unsigned short>(q);
                        }
                    }
                    c += __popc(grp);
                }
            }
        };
        scan(std::true_type{}, h);
        for (int s = 1; s < 27; ++s)
            scan(std::false_type{}, h + c_nix[s] + g.H[0] * (c_niy[s] + g.H[1] * c_niz[s]));
        if (sub == 0) {
            if (c > mx) {
                atomicOr(&err[kErrOverflow], 1);
                c = mx;
            }
            kvois[p] = c;
        }
    }
}

size_t nlist_lds_bytes(int cap, int ngroup) {
    const size_t slots = static_cast<size_t>(cap) + kNlPad;
    return (static_cast<size_t>(cap) * 2 + slots) * sizeof(double) +
           (ngroup > 1 ? slots * sizeof(int) : 0);
}

__global__ void k_zero_range(int* __restrict__ p, const int* __restrict__ first, int n) {
    const int i = *first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_nlist_bricks(int wide, int nbricks, int lanes, int cap, int write_sidx,
                         const int* brick_box, int n,
                         int ncell_total, const double4* pos, const int* ityp, const int* start,
                         const BoxDev* boxes, int ngroup, const double* rnl2, int mx, int* kvois,
                         int* indi, unsigned short* sidx, int stride, int* err, hipStream_t s) {
    if (n <= 0) return;
    // atoms in the out-of-box bucket have no neighbours
    hipLaunchKernelGGL(k_zero_range, dim3(grid_for(n)), dim3(kBlock), 0, s, kvois,
                       start + ncell_total, n);
    if (nbricks <= 0) return;
    if (!wide) {
#define MSM_NS(MT, S)                                                                         \
    hipLaunchKernelGGL((k_nlist_small<MT, S>), dim3(nbricks), dim3(kNlSmallBlock), 0, s, brick_box, pos, \
                       ityp, start, boxes, ngroup, rnl2, mx, kvois, indi, sidx, stride, err)
        if (ngroup > 1) { if (write_sidx) MSM_NS(1, 1); else MSM_NS(1, 0); }
        else            { if (write_sidx) MSM_NS(0, 1); else MSM_NS(0, 0); }
#undef MSM_NS
        return;
    }
    const size_t lds = nlist_lds_bytes(cap, ngroup);
#define MSM_NL(MT, T, S)                                                                       \
    hipLaunchKernelGGL((k_nlist_bricks<MT, T, S>), dim3(nbricks), dim3(kNlBlock), lds, s,       \
                       brick_box, pos, ityp, start, boxes, ngroup, rnl2, mx, kvois, indi, sidx,  \
                       stride, err, cap)
#define MSM_NLT(MT, S)                                \
    do {                                              \
        if (lanes >= 8) MSM_NL(MT, 8, S);             \
        else if (lanes >= 4) MSM_NL(MT, 4, S);        \
        else if (lanes >= 2) MSM_NL(MT, 2, S);        \
        else MSM_NL(MT, 1, S);                        \
    } while (0)
    if (ngroup > 1) { if (write_sidx) MSM_NLT(1, 1); else MSM_NLT(1, 0); }
    else            { if (write_sidx) MSM_NLT(0, 1); else MSM_NLT(0, 0); }
#undef MSM_NLT
#undef MSM_NL
}

// --------------------------------------------------------------------------------------------
// density pass
// --------------------------------------------------------------------------------------------

template <int MT, int EAM>
__global__ __launch_bounds__(kBlock) void k_density_bricks(const int* __restrict__ brick_box,
                                                           AtomsDev a, ListDev l, TabDev t,
                                                           const BoxDev* __restrict__ boxes,
                                                           const int* __restrict__ start,
                                                           int* err) {
    __shared__ double lx[kNlCapPass], ly[kNlCapPass], lz[kNlCapPass];
    __shared__ int lt[MT ? kNlCapPass : 1];
    __shared__ BrickLds L;

    const int tid = threadIdx.x;
    const int bi = xcd_block(gridDim.x);
    const BoxDev& B = boxes[brick_box[bi]];
    BrickInfo g;
    brick_info(B, bi, g);
    const int total = brick_halo(B, g, start, L.hcell, nullptr, L.hoff, L.homeoff, L.scan);
    const bool staged = total <= min(kNlCapPass, l.scap);
    if (staged) {
        for_halo_slots(g, start, L.hcell, L.hoff, [&](int s, int q, int) {
            // x, y, z only: w (F') of other atoms is being written by this pass
            const double2 xy = *reinterpret_cast<const double2*>(&a.pos[q]);
            lx[s] = xy.x;
            ly[s] = xy.y;
            lz[s] = a.pos[q].z;
            if constexpr (MT) lt[s] = a.ityp[q];
        });
    }
    __syncthreads();
    const int nhome = L.homeoff[g.nhc];
    for (int tt = tid; tt < nhome; tt += kBlock) {
        int h, off;
        home_atom(tt, g, L.homeoff, h, off);
        const int p = start[L.hcell[h]] + off;
        const int si = L.hoff[h] + off;
        double den0 = 0.0;
        if ((a.statu[p] & STATU_ACTIVE) == STATU_ACTIVE) {
            const int ti = MT ? a.ityp[p] : 0;
            const int kv = l.kvois[p];
            if (staged) {
                // Neighbours in groups of kUnroll: the group's slot loads are issued together
                // (one memory latency per group), and the body is branch-free -- the table
                // index is clamped for pairs beyond the cutoff and their term is selected away
                // -- so the group's table loads are in flight together as well.  Terms are
                // still added in list order.
                const double xi = lx[si], yi = ly[si], zi = lz[si];
                const unsigned short* __restrict__ li = l.sidx + list_base(p, l.mx);
                const int kmax = t.ntab2 - 1;
                for (int base = 0; base < kv; base += kUnroll) {
                    int qs[kUnroll];
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u)
                        qs[u] = (base + u < kv) ? li[static_cast<size_t>(base + u) * l.stride] : si;
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u) {
                        const int q = qs[u];
                        const double sx = min_image(xi - lx[q], B.len[0], B.ilen[0]);
                        const double sy = min_image(yi - ly[q], B.len[1], B.ilen[1]);
                        const double sz = min_image(zi - lz[q], B.len[2], B.ilen[2]);
                        const double r2 = sx * sx + sy * sy + sz * sz;
                        const int kind = MT ? t.kpair[ti * t.ngroup + lt[q]] : t.kpair[0];
                        double inv_r;
                        const double sk = table_coord(r2, t.csi, inv_r);
                        const int kk = min(static_cast<int>(sk), kmax);
                        const double2 e = t.b2[kind * t.ntab2 + kk];
                        const double v = e.x + (sk - kk) * e.y;
                        den0 = den0 + ((base + u < kv && r2 <= t.ru2max) ? v : 0.0);
                    }
                }
            } else {
                const double4 pi = a.pos[p];
                const int* __restrict__ li = l.indi + list_base(p, l.mx);
                for (int iw = 0; iw < kv; ++iw) {
                    const int j = li[static_cast<size_t>(iw) * l.stride];
                    const double2 qxy = *reinterpret_cast<const double2*>(&a.pos[j]);
                    const double qz = a.pos[j].z;
                    const double sx = min_image(pi.x - qxy.x, B.len[0], B.ilen[0]);
                    const double sy = min_image(pi.y - qxy.y, B.len[1], B.ilen[1]);
                    const double sz = min_image(pi.z - qz, B.len[2], B.ilen[2]);
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
        a.pos[p].w = den0;
    }
}

// --------------------------------------------------------------------------------------------
// force pass (+ corrector of this step and predictor of the next)
// --------------------------------------------------------------------------------------------

template <int MT, int MODE>
__global__ __launch_bounds__(kBlock) void k_force_bricks(const int* __restrict__ brick_box,
                                                         AtomsDev a, ListDev l, TabDev t,
                                                         const BoxDev* __restrict__ boxes,
                                                         const int* __restrict__ start,
                                                         StepParams sp) {
    __shared__ double4 lp[kNlCapPass];
    __shared__ int lt[MT ? kNlCapPass : 1];
    __shared__ BrickLds L;

    const int tid = threadIdx.x;
    const int bi = xcd_block(gridDim.x);
    const BoxDev& B = boxes[brick_box[bi]];
    BrickInfo g;
    brick_info(B, bi, g);
    const int total = brick_halo(B, g, start, L.hcell, nullptr, L.hoff, L.homeoff, L.scan);
    const bool staged = total <= min(kNlCapPass, l.scap);
    if (staged) {
        for_halo_slots(g, start, L.hcell, L.hoff, [&](int s, int q, int) {
            lp[s] = a.pos[q];  // x, y, z and F'
            if constexpr (MT) lt[s] = a.ityp[q];
        });
    }
    __syncthreads();
    const int nhome = L.homeoff[g.nhc];
    for (int tt = tid; tt < nhome; tt += kBlock) {
        int h, off;
        home_atom(tt, g, L.homeoff, h, off);
        const int p = start[L.hcell[h]] + off;
        const int st = a.statu[p];
        const bool active = (st & STATU_ACTIVE) == STATU_ACTIVE;
        const double4 pi = staged ? lp[L.hoff[h] + off] : a.pos[p];
        const int ti = MT ? a.ityp[p] : 0;
        double fpx = 0.0, fpy = 0.0, fpz = 0.0;
        if (active) {
            const double denki = pi.w;
            const int kv = l.kvois[p];
            if (staged) {
                // grouped, branch-free neighbour loop (see the density pass)
                const unsigned short* __restrict__ li = l.sidx + list_base(p, l.mx);
                const int si = L.hoff[h] + off;
                const int kmax = t.ntab2 - 1;
                for (int base = 0; base < kv; base += kUnroll) {
                    int qs[kUnroll];
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u)
                        qs[u] = (base + u < kv) ? li[static_cast<size_t>(base + u) * l.stride] : si;
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u) {
                        const int q = qs[u];
                        const double4 pj = lp[q];
                        const double sx = min_image(pi.x - pj.x, B.len[0], B.ilen[0]);
                        const double sy = min_image(pi.y - pj.y, B.len[1], B.ilen[1]);
                        const double sz = min_image(pi.z - pj.z, B.len[2], B.ilen[2]);
                        const double r2 = sx * sx + sy * sy + sz * sz;
                        int kind, kind1;
                        if constexpr (MT) {
                            const int tj = lt[q];
                            kind = t.kpair[ti * t.ngroup + tj];
                            kind1 = t.kpair[tj * t.ngroup + ti];
                        } else {
                            kind = kind1 = t.kpair[0];
                        }
                        double inv_r;
                        const double sk = table_coord(r2, t.csi, inv_r);
                        const int kk = min(static_cast<int>(sk), kmax);
                        const double dk = sk - kk;
                        const double4 f = t.f4[kind * t.ntab2 + kk];
                        double fb1;
                        if constexpr (MT) {
                            const double4 f1 = t.f4[kind1 * t.ntab2 + kk];
                            fb1 = f1.z + dk * f1.w;
                        } else {
                            fb1 = f.z + dk * f.w;
                        }
                        const double ft = (f.x + dk * f.y) * (inv_r * inv_r) +
                                          ((f.z + dk * f.w) * denki + fb1 * pj.w) * inv_r;
                        const double fortot = (base + u < kv && r2 <= t.ru2max) ? ft : 0.0;
                        fpx = fpx + fortot * sx;
                        fpy = fpy + fortot * sy;
                        fpz = fpz + fortot * sz;
                    }
                }
            } else
            for (int iw = 0; iw < kv; ++iw) {
                const size_t a_iw = list_base(p, l.mx) + static_cast<size_t>(iw) * l.stride;
                const int j = l.indi[a_iw];
                const double4 pj = a.pos[j];
                int tj = 0;
                if constexpr (MT) tj = a.ityp[j];
                const double sx = min_image(pi.x - pj.x, B.len[0], B.ilen[0]);
                const double sy = min_image(pi.y - pj.y, B.len[1], B.ilen[1]);
                const double sz = min_image(pi.z - pj.z, B.len[2], B.ilen[2]);
                const double r2 = sx * sx + sy * sy + sz * sz;
                if (r2 <= t.ru2max) {
                    int kind, kind1;
                    if constexpr (MT) {
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
                    // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r  (CALFORCE_KERNEL)
                    const double fortot = (f.x + dk * f.y) * (inv_r * inv_r) +
                                          ((f.z + dk * f.w) * denki + fb1 * pj.w) * inv_r;
                    fpx = fpx + fortot * sx;
                    fpy = fpy + fortot * sy;
                    fpz = fpz + fortot * sz;
                }
            }
        }
        if (sp.write_f) {
            a.fx[p] = fpx;
            a.fy[p] = fpy;
            a.fz[p] = fpz;
        }
        if constexpr (MODE >= 1)
            integrate_atom<MODE>(a, t, B, sp, p, st, active, pi, ti, fpx, fpy, fpz);
    }
}

// Atoms of the out-of-box bucket belong to no brick: F' = 0, no force, position carried over.
__global__ void k_outside_atoms(AtomsDev a, const int* __restrict__ first, int write_f,
                                int copy_pos, int clear_den) {
    const int p = *first + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    if (clear_den) a.pos[p].w = 0.0;
    if (write_f) {
        a.fx[p] = 0.0;
        a.fy[p] = 0.0;
        a.fz[p] = 0.0;
    }
    if (copy_pos) {
        const double4 x = a.pos[p];
        a.pos_out[p] = make_double4(x.x, x.y, x.z, 0.0);
    }
}

void launch_outside_atoms(const AtomsDev& a, const int* first, int write_f, int copy_pos,
                          int clear_den, hipStream_t s) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_outside_atoms, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a, first,
                       write_f, copy_pos, clear_den);
}

void launch_density_bricks(int nbricks, const int* brick_box, const AtomsDev& a,
                           const ListDev& l, const TabDev& t, const BoxDev* boxes,
                           const int* start, int ncell_total, int* err, hipStream_t s) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_outside_atoms, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a,
                       start + ncell_total, 0, 0, 1);
    if (nbricks <= 0) return;
    const dim3 gr(nbricks), bl(kBlock);
    const bool mt = t.ngroup > 1, eam = t.pot_type != 0;
#define MSM_DB(MT, E) \
    hipLaunchKernelGGL((k_density_bricks<MT, E>), gr, bl, 0, s, brick_box, a, l, t, boxes, start, err)
    if (mt) { if (eam) MSM_DB(1, 1); else MSM_DB(1, 0); }
    else    { if (eam) MSM_DB(0, 1); else MSM_DB(0, 0); }
#undef MSM_DB
}

void launch_force_bricks(int nbricks, const int* brick_box, const AtomsDev& a, const ListDev& l,
                         const TabDev& t, const BoxDev* boxes, const int* start, int ncell_total,
                         const StepParams& sp, hipStream_t s) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_outside_atoms, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a,
                       start + ncell_total, sp.write_f, sp.mode == 2 ? 1 : 0, 0);
    if (nbricks <= 0) return;
    const dim3 gr(nbricks), bl(kBlock);
    const bool mt = t.ngroup > 1;
#define MSM_FB(MT, M) \
    hipLaunchKernelGGL((k_force_bricks<MT, M>), gr, bl, 0, s, brick_box, a, l, t, boxes, start, sp)
    if (mt) {
        if (sp.mode == 0) MSM_FB(1, 0); else if (sp.mode == 1) MSM_FB(1, 1); else MSM_FB(1, 2);
    } else {
        if (sp.mode == 0) MSM_FB(0, 0); else if (sp.mode == 1) MSM_FB(0, 1); else MSM_FB(0, 2);
    }
#undef MSM_FB
}

}  // namespace msm
