// msm_tpa.hip -- density and force passes over the global neighbour list with TPA lanes per atom.
//
// Each atom's list is split round-robin over TPA adjacent lanes of a wave; a lane loads the
// indices of up to kGroup of its neighbours at once, then gathers their positions and table
// knots without branches (pairs beyond the cutoff keep a clamped table index and contribute an
// exact zero), so a wave keeps many independent loads in flight.  The partial sums of the TPA
// lanes are combined with a fixed butterfly of wave shuffles, which keeps the result
// deterministic.  The lane with sub-index 0 finishes the atom (F'(rho) for the density pass; the
// fused corrector/predictor for the force pass).
//
// Same semantics as the per-atom kernels of msm_kernels.hip (MD_FS/EAM_ForceTable_GPU.F90
// PRECALFOR_KERNEL / CALFORCE_KERNEL; MD_DiffScheme_GPU.F90 Correction/Predictor kernels).
#include "msm_device.h"
#include "msm_internal.h"

namespace msm {

namespace {

constexpr int kGroup = 8;

template <int TPA>
__device__ __forceinline__ double lane_sum(double v) {
#pragma unroll
    for (int d = TPA / 2; d >= 1; d >>= 1) v = v + __shfl_xor(v, d, 64);
    return v;
}

}  // namespace

template <int MB, int MT, int EAM, int TPA>
__global__ __launch_bounds__(kBlock) void k_density_tpa(AtomsDev a, ListDev l, TabDev t,
                                                        BoxDev box0,
                                                        const BoxDev* __restrict__ boxes,
                                                        int* err) {
    const int gt = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    const int i = gt / TPA, sub = gt % TPA;
    const bool have = i < a.n;
    const bool active = have && (a.statu[i] & STATU_ACTIVE) == STATU_ACTIVE;
    double den0 = 0.0;
    int ti = 0;
    int nin = 0;
    if (active) {
        const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, i);
        const double2 pxy = *reinterpret_cast<const double2*>(&a.pos[i]);
        const double pz = a.pos[i].z;
        ti = MT ? a.ityp[i] : 0;
        const int kv = l.kvois[i];
        const int* __restrict__ li = l.indi + list_base(i, l.mx);
        const int kmax = t.ntab2 - 1;
        for (int base = sub; base < kv; base += TPA * kGroup) {
            int js[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int iw = base + u * TPA;
                js[u] = iw < kv ? li[static_cast<size_t>(iw) * l.stride] : i;
            }
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int j = js[u];
                const double2 qxy = *reinterpret_cast<const double2*>(&a.pos[j]);
                const double qz = a.pos[j].z;
                const double sx = min_image(pxy.x - qxy.x, b.len[0], b.ilen[0]);
                const double sy = min_image(pxy.y - qxy.y, b.len[1], b.ilen[1]);
                const double sz = min_image(pz - qz, b.len[2], b.ilen[2]);
                const double r2 = sx * sx + sy * sy + sz * sz;
                const int kind = MT ? t.kpair[ti * t.ngroup + a.ityp[j]] : t.kpair[0];
                double inv_r;
                const double sk = table_coord(r2, t.csi, inv_r);
                const int kk = min(static_cast<int>(sk), kmax);
                const bool in = base + u * TPA < kv && r2 <= t.ru2max;
                double v = 0.0;
                if (in) {  // knots are fetched for pairs inside the cutoff only
                    const double2 e = t.b2[kind * t.ntab2 + kk];
                    v = e.x + (sk - kk) * e.y;
                }
                den0 = den0 + v;
                if (in && l.cidx) {
                    // this step's in-cutoff pairs, in list order, for the force pass
                    l.cidx[list_base(i, l.mx) + static_cast<size_t>(sub + TPA * nin) * l.stride] = j;
                    ++nin;
                }
            }
        }
    }
    if (l.cidx && have) l.ckv[gt] = nin;
    den0 = lane_sum<TPA>(den0);
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

template <int MB, int MT, int MODE, int TPA>
__global__ __launch_bounds__(kBlock) void k_force_tpa(AtomsDev a, ListDev l, TabDev t, BoxDev box0,
                                                      const BoxDev* __restrict__ boxes,
                                                      StepParams sp) {
    const int gt = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    const int i = gt / TPA, sub = gt % TPA;
    const bool have = i < a.n;
    const int st = have ? a.statu[i] : 0;
    const bool active = have && (st & STATU_ACTIVE) == STATU_ACTIVE;
    const double4 pi = have ? a.pos[i] : make_double4(0, 0, 0, 0);
    const BoxDev& b = box_of<MB>(box0, boxes, a.boxid, have ? i : 0);
    const int ti = (MT && have) ? a.ityp[i] : 0;
    double fpx = 0.0, fpy = 0.0, fpz = 0.0;
    if (active) {
        const double denki = pi.w;
        // With the compacted list of the density pass this lane's entries sit at slots
        // sub + TPA*c, c < its count, which the loop below visits as iw < sub + TPA*count.
        const bool cmp = l.cidx != nullptr;
        const int kv = cmp ? sub + TPA * l.ckv[gt] : l.kvois[i];
        const int* __restrict__ li = (cmp ? l.cidx : l.indi) + list_base(i, l.mx);
        const int kmax = t.ntab2 - 1;
        for (int base = sub; base < kv; base += TPA * kGroup) {
            int js[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int iw = base + u * TPA;
                js[u] = iw < kv ? li[static_cast<size_t>(iw) * l.stride] : i;
            }
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                const int j = js[u];
                const double4 pj = a.pos[j];
                const double sx = min_image(pi.x - pj.x, b.len[0], b.ilen[0]);
                const double sy = min_image(pi.y - pj.y, b.len[1], b.ilen[1]);
                const double sz = min_image(pi.z - pj.z, b.len[2], b.ilen[2]);
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
                double fortot = 0.0;
                if (base + u * TPA < kv && r2 <= t.ru2max) {
                    // knots are fetched for pairs inside the cutoff only
                    const double4 f = t.f4[kind * t.ntab2 + kk];
                    double fb1;
                    if constexpr (MT) {
                        const double4 f1 = t.f4[kind1 * t.ntab2 + kk];
                        fb1 = f1.z + dk * f1.w;
                    } else {
                        fb1 = f.z + dk * f.w;
                    }
                    // FPOTR/r^2 + (FPOTB(KTAB) F'_i + FPOTB(KTAB1) F'_j)/r  (CALFORCE_KERNEL)
                    fortot = (f.x + dk * f.y) * (inv_r * inv_r) +
                             ((f.z + dk * f.w) * denki + fb1 * pj.w) * inv_r;
                }
                fpx = fpx + fortot * sx;
                fpy = fpy + fortot * sy;
                fpz = fpz + fortot * sz;
            }
        }
    }
    fpx = lane_sum<TPA>(fpx);
    fpy = lane_sum<TPA>(fpy);
    fpz = lane_sum<TPA>(fpz);
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
inline unsigned grid_tpa(int n, int tpa) {
    const long long th = static_cast<long long>(n) * tpa;
    return static_cast<unsigned>((th + kBlock - 1) / kBlock);
}
}  // namespace

void launch_density_tpa(int tpa, const AtomsDev& a, const ListDev& l, const TabDev& t,
                        const BoxDev& box0, const BoxDev* boxes, int multibox, int* err,
                        hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(grid_tpa(a.n, tpa)), bl(kBlock);
    const bool mt = t.ngroup > 1, eam = t.pot_type != 0;
#define MSM_DT(MB, MT, E, P) \
    hipLaunchKernelGGL((k_density_tpa<MB, MT, E, P>), g, bl, 0, s, a, l, t, box0, boxes, err)
#define MSM_DTP(MB, MT, E)                 \
    do {                                   \
        if (tpa == 1) MSM_DT(MB, MT, E, 1);      \
        else if (tpa == 2) MSM_DT(MB, MT, E, 2); \
        else if (tpa == 4) MSM_DT(MB, MT, E, 4); \
        else MSM_DT(MB, MT, E, 8);               \
    } while (0)
    if (multibox) {
        if (mt) { if (eam) MSM_DTP(1, 1, 1); else MSM_DTP(1, 1, 0); }
        else    { if (eam) MSM_DTP(1, 0, 1); else MSM_DTP(1, 0, 0); }
    } else {
        if (mt) { if (eam) MSM_DTP(0, 1, 1); else MSM_DTP(0, 1, 0); }
        else    { if (eam) MSM_DTP(0, 0, 1); else MSM_DTP(0, 0, 0); }
    }
#undef MSM_DTP
#undef MSM_DT
}

void launch_force_tpa(int tpa, const AtomsDev& a, const ListDev& l, const TabDev& t,
                      const BoxDev& box0, const BoxDev* boxes, int multibox, const StepParams& sp,
                      hipStream_t s) {
    if (a.n <= 0) return;
    const dim3 g(grid_tpa(a.n, tpa)), bl(kBlock);
    const bool mt = t.ngroup > 1;
#define MSM_FT(MB, MT, M, P) \
    hipLaunchKernelGGL((k_force_tpa<MB, MT, M, P>), g, bl, 0, s, a, l, t, box0, boxes, sp)
#define MSM_FTP(MB, MT, M)                 \
    do {                                   \
        if (tpa == 1) MSM_FT(MB, MT, M, 1);      \
        else if (tpa == 2) MSM_FT(MB, MT, M, 2); \
        else if (tpa == 4) MSM_FT(MB, MT, M, 4); \
        else MSM_FT(MB, MT, M, 8);               \
    } while (0)
#define MSM_FTM(MB, MT)                                  \
    do {                                                 \
        if (sp.mode == 0) MSM_FTP(MB, MT, 0);            \
        else if (sp.mode == 1) MSM_FTP(MB, MT, 1);       \
        else MSM_FTP(MB, MT, 2);                         \
    } while (0)
    if (multibox) {
        if (mt) MSM_FTM(1, 1); else MSM_FTM(1, 0);
    } else {
        if (mt) MSM_FTM(0, 1); else MSM_FTM(0, 0);
    }
#undef MSM_FTM
#undef MSM_FTP
#undef MSM_FT
}

}  // namespace msm
