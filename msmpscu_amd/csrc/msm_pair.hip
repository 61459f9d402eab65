// msm_pair.hip -- density and force passes over the global neighbour list, two lanes per atom,
// with each 32-byte record gathered by a lane pair.
//
// A wave serves 32 atoms: lane l (l < 32) and lane l + 32 share atom wave*32 + l and split its
// list by slot parity, exactly as the two-lane kernels of msm_tpa.hip do, so the sums (and the
// results) are bit-identical to that path.  What changes is the gather.  A neighbour record
// (position + F') and a force-table knot are 32 bytes each, which a lane needs two 16-byte loads
// for.  Here the two lanes of a pair load the two halves of the SAME record in one instruction
// (the low lane half 0, the high lane half 1), once for the low lane's neighbour and once for the
// high lane's, and gfx950's v_permlane32_swap hands each lane the half it is missing.  Every
// gather instruction then touches 32 records instead of 64, which halves the L1 tag lookups that
// bound these passes (one lookup per cache line per instruction).
//
// Same semantics as msm_tpa.hip (MD_FS/EAM_ForceTable_GPU.F90 PRECALFOR_KERNEL /
// CALFORCE_KERNEL; MD_DiffScheme_GPU.F90 Correction/Predictor kernels).
#include "msm_internal.h"
#include "msm_lanepair.h"

namespace msm {

namespace {

constexpr int kGroup = 8;

__device__ __forceinline__ double pair_sum(double v) {
    // each lane adds its partner's value, in the order of the two-lane kernels
    return v + __shfl_xor(v, 1, 64);
}

// lanes 2m and 2m + 1 of the grid share atom m
__device__ __forceinline__ void pair_atom(int& i, int& sub) {
    const int gt = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    i = gt >> 1;
    sub = gt & 1;
}

}  // namespace

template <int MB, int MT, int EAM>
__global__ __launch_bounds__(kBlock) void k_density_pair(AtomsDev a, ListDev l, TabDev t,
                                                         BoxDev box0,
                                                         const BoxDev* __restrict__ boxes,
                                                         int* err) {
    int i, sub;
    pair_atom(i, sub);
    const bool have = i < a.n;
    const bool active = have && (a.statu[i] & STATU_ACTIVE) == STATU_ACTIVE;
    double den0 = 0.0;
    int ti = 0;
    if (active) {
        const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
        const double2 pxy = *reinterpret_cast<const double2*>(&a.pos[i]);
        const double pz = a.pos[i].z;
        ti = MT ? a.ityp[i] : 0;
        const int kv = l.kvois[i];
        const int* __restrict__ li = l.indi + list_base(i, l.mx);
        const int kmax = t.ntab2 - 1;
        // both lanes of a pair run the same number of groups (same atom, same kv)
        for (int base = 0; base < kv; base += 2 * kGroup) {
            int js[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int iw = base + 2 * u + sub;
                js[u] = iw < kv ? li[static_cast<size_t>(iw) * l.stride] : i;
            }
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int j = js[u];
                double2 qxy, qzw;
                pair_gather(a.pos, j, sub, qxy, qzw);
                const double sx = min_image(pxy.x - qxy.x, b.len[0], b.ilen[0]);
                const double sy = min_image(pxy.y - qxy.y, b.len[1], b.ilen[1]);
                const double sz = min_image(pz - qzw.x, b.len[2], b.ilen[2]);
                const double r2 = sx * sx + sy * sy + sz * sz;
                const int kind = MT ? t.kpair[ti * t.ngroup + a.ityp[j]] : t.kpair[0];
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), kmax);
                const bool in = base + 2 * u + sub < kv && r2 <= t.ru2max;
                double v = 0.0;
                if (in) {  // knots are fetched for pairs inside the cutoff only
                    const double2 e = t.b2[kind * t.ntab2 + kk];
                    v = e.x + (sk - kk) * e.y;
                }
                den0 = den0 + v;
            }
        }
    }
    den0 = pair_sum(den0);
    if (sub != 0 || !have) return;
    if (active) {
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

template <int MB, int MT, int MODE>
__global__ __launch_bounds__(kBlock) void k_force_pair(AtomsDev a, ListDev l, TabDev t,
                                                       BoxDev box0,
                                                       const BoxDev* __restrict__ boxes,
                                                       StepParams sp) {
    int i, sub;
    pair_atom(i, sub);
    const bool have = i < a.n;
    const int st = have ? a.statu[i] : 0;
    const bool active = have && (st & STATU_ACTIVE) == STATU_ACTIVE;
    const double4 pi = have ? a.pos[i] : make_double4(0, 0, 0, 0);
    const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, have ? i : 0);
    const int ti = (MT && have) ? a.ityp[i] : 0;
    double fpx = 0.0, fpy = 0.0, fpz = 0.0;
    if (active) {
        const double denki = pi.w;
        const int kv = l.kvois[i];
        const int* __restrict__ li = l.indi + list_base(i, l.mx);
        const int kmax = t.ntab2 - 1;
        for (int base = 0; base < kv; base += 2 * kGroup) {
            int js[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int iw = base + 2 * u + sub;
                js[u] = iw < kv ? li[static_cast<size_t>(iw) * l.stride] : i;
            }
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int j = js[u];
                double2 qxy, qzw;
                pair_gather(a.pos, j, sub, qxy, qzw);
                const double sx = min_image(pi.x - qxy.x, b.len[0], b.ilen[0]);
                const double sy = min_image(pi.y - qxy.y, b.len[1], b.ilen[1]);
                const double sz = min_image(pi.z - qzw.x, b.len[2], b.ilen[2]);
                const double r2 = sx * sx + sy * sy + sz * sz;
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
                const int kk = min(static_cast<int>(sk), kmax);
                const double dk = sk - kk;
                double2 fr, fb;  // {FPOTR, dFPOTR}, {FPOTB, dFPOTB} of knot kk of KTAB
                pair_gather(t.f4, kind * t.ntab2 + kk, sub, fr, fb);
                double fb1;
                if constexpr (MT) {
                    const double2 f1 = reinterpret_cast<const double2*>(t.f4)
                        [2 * (static_cast<size_t>(kind1) * t.ntab2 + kk) + 1];
                    fb1 = f1.x + dk * f1.y;
                } else {
                    fb1 = fb.x + dk * fb.y;
                }
                // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r  (CALFORCE_KERNEL)
                const double ft = (fr.x + dk * fr.y) * (inv_r * inv_r) +
                                  ((fb.x + dk * fb.y) * denki + fb1 * qzw.y) * inv_r;
                const double fortot = (base + 2 * u + sub < kv && r2 <= t.ru2max) ? ft : 0.0;
                fpx = fpx + fortot * sx;
                fpy = fpy + fortot * sy;
                fpz = fpz + fortot * sz;
            }
        }
    }
    fpx = pair_sum(fpx);
    fpy = pair_sum(fpy);
    fpz = pair_sum(fpz);
    if (sub != 0 || !have) return;
    if (sp.write_f) {
        a.fx[i] = fpx;
        a.fy[i] = fpy;
        a.fz[i] = fpz;
    }
    if constexpr (MODE >= 1)
        integrate_atom<MODE>(a, t, b, sp, i, st, active, pi, ti, fpx, fpy, fpz);
}

namespace {
inline unsigned grid_pair(int n) {
    const long long th = (static_cast<long long>(n) + 31) / 32 * 64;
    return static_cast<unsigned>((th + kBlock - 1) / kBlock);
}
}  // namespace

void launch_density_pair(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                         const BoxDev* boxes, int multibox, int* err, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(grid_pair(a.n)), bl(kBlock);
    const bool mt = t.ngroup > 1, eam = t.pot_type != 0;
#define MSM_DP(MB, MT, E) \
    hipLaunchKernelGGL((k_density_pair<MB, MT, E>), g, bl, 0, s, a, l, t, box0, boxes, err)
    if (multibox) {
        if (mt) { if (eam) MSM_DP(1, 1, 1); else MSM_DP(1, 1, 0); }
        else    { if (eam) MSM_DP(1, 0, 1); else MSM_DP(1, 0, 0); }
    } else {
        if (mt) { if (eam) MSM_DP(0, 1, 1); else MSM_DP(0, 1, 0); }
        else    { if (eam) MSM_DP(0, 0, 1); else MSM_DP(0, 0, 0); }
    }
#undef MSM_DP
}

void launch_force_pair(const AtomsDev& a, const ListDev& l, const TabDev& t, const BoxDev& box0,
                       const BoxDev* boxes, int multibox, const StepParams& sp, hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(grid_pair(a.n)), bl(kBlock);
    const bool mt = t.ngroup > 1;
#define MSM_FP(MB, MT, M) \
    hipLaunchKernelGGL((k_force_pair<MB, MT, M>), g, bl, 0, s, a, l, t, box0, boxes, sp)
#define MSM_FPM(MB, MT)                          \
    do {                                         \
        if (sp.mode == 0) MSM_FP(MB, MT, 0);     \
        else if (sp.mode == 1) MSM_FP(MB, MT, 1); \
        else MSM_FP(MB, MT, 2);                  \
    } while (0)
    if (multibox) {
        if (mt) MSM_FPM(1, 1); else MSM_FPM(1, 0);
    } else {
        if (mt) MSM_FPM(0, 1); else MSM_FPM(0, 0);
    }
#undef MSM_FPM
#undef MSM_FP
}

}  // namespace msm
